"""CPU restatement of the reference's runtime synthesis for the banked path
(test infrastructure only): the walking schedule of internal/source/source.go
and sensing.Reference (internal/sensing/sensing.go:344-386), the monolithic
sum that the reference's own tests hold the streaming engine to.

    walk_contact_at, walk_contacts, step_period   source.go:164-252
    range_span                                    source.go:254-263
    impulse_length                                sensing.go:135-148
    quantized_range                               sensing.go:175-190
    reference                                     sensing.go:344-386, with
        FromBank's impulse responses (oracle.bank_impulse) and
        stream.Monolithic (stream_oracle.Monolithic)

Forces come in as callables (the contact's vertical force at absolute time
t), as the reference's Profile interface hands them over; the bank's radial
response is zero, so only the vertical force is convolved.

Nothing in the product imports this module.  Pinned by
tests/test_sensing_oracle.py against the reference's source tests
(source_test.go known answers) and closed forms.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

RANGE_QUANTUM = 0.02
STANCE_FRACTION = 0.60
DEFAULT_WIDTH = 0.12


@dataclass
class WalkSpec:
    """The fields of source.Walk the schedule reads."""

    duration: float  # Stance.Duration
    speed: float
    start_x: float = 0.0
    start_y: float = 0.0
    heading: float = 0.0
    stride: float = 0.0
    width: float = 0.0
    until: float = 0.0
    first: str = "left_foot"
    second: str = "right_foot"


@dataclass
class OContact:
    id: str
    x: float
    y: float
    start: float
    end: float


def stride_or_default(w: WalkSpec) -> float:
    return w.stride if w.stride > 0 else w.speed * w.duration / STANCE_FRACTION


def step_period(w: WalkSpec) -> float:
    return stride_or_default(w) / 2 / w.speed


def walk_contact_at(w: WalkSpec, n: int) -> OContact:
    """Walk.ContactAt (source.go:195-217)."""
    along = float(n) * (stride_or_default(w) / 2)
    side = (w.width if w.width > 0 else DEFAULT_WIDTH) / 2
    if n % 2 == 1:
        side = -side
    c, s = math.cos(w.heading), math.sin(w.heading)
    start = float(n) * step_period(w)
    return OContact(w.second if n % 2 == 1 else w.first, w.start_x + along * c - side * s,
                    w.start_y + along * s + side * c, start, start + w.duration)


def walk_contacts(w: WalkSpec, frm: float, to: float) -> List[OContact]:
    """Walk.Contacts (source.go:219-252): starts in [frm, to), in order."""
    if to <= frm:
        return []
    p = step_period(w)
    if p <= 0 or (w.until > 0 and frm >= w.until):
        return []
    if w.until > 0:
        to = min(to, w.until)
    out = []
    for n in range(max(math.ceil(frm / p), 0), math.ceil(to / p)):
        c = walk_contact_at(w, n)
        if frm <= c.start < to:
            out.append(c)
    return out


def range_span(contacts: List[OContact], rx: float, ry: float):
    if not contacts:
        return math.inf, 0.0, False
    r = [math.hypot(rx - c.x, ry - c.y) for c in contacts]
    return min(r), max(r), True


def impulse_length(far: float, rayleigh_velocity: float, fs: float, transform: int) -> int:
    """NewEngineWith's impulse-response length (sensing.go:135-148)."""
    n = int(math.ceil((far / rayleigh_velocity + 2) * fs))
    return min(n, 3 * transform // 4) if transform > 0 else n


def quantized_range(r: float) -> float:
    """The range responseFor builds at (sensing.go:186-190): r rounded half
    away from zero to a RangeQuantum, at least one quantum."""
    x = r / RANGE_QUANTUM
    k = math.floor(x)
    if x - k >= 0.5:  # exact for a double: math.Round's half away from zero
        k += 1
    return max(float(k) * RANGE_QUANTUM, RANGE_QUANTUM)


def reference(contacts: List[OContact], forces: List[Callable[[np.ndarray], np.ndarray]], n: int, fs: float,
              impulse: Callable[[float], np.ndarray], rx: float = 0.0, ry: float = 0.0) -> np.ndarray:
    """sensing.Reference (sensing.go:353-386) for a receiver at (rx, ry): per
    contact beginning in [0, n / fs), its vertical force sampled at i / fs
    convolved with the impulse response at its quantized range, truncated to
    n samples, summed.  impulse(rq) returns the response at range rq."""
    import stream_oracle

    span = float(n) / fs
    out = np.zeros(n)
    t = np.arange(n) / fs
    cache = {}
    for c, force in zip(contacts, forces):
        if not (0 <= c.start < span):
            continue
        rq = quantized_range(math.hypot(rx - c.x, ry - c.y))
        if rq not in cache:
            cache[rq] = impulse(rq)
        out += stream_oracle.Monolithic(force(t), cache[rq])
    return out
