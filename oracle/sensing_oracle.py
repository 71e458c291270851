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
    rayleigh_velocity                             soil.go:136-163
    eigenfunctions, energy_integral               green.go:58-117
    velocity_response, radial_force_response      green.go:169-266, 375-388
    causal_impulse_response                       dsp.go:232-260
    reference_xyz                                 sensing.go:344-386 with the
        radial-force response (Analytic propagation)

Forces come in as callables (the contact's vertical force at absolute time
t), as the reference's Profile interface hands them over; the bank's radial
response is zero, so only the vertical force is convolved.

Nothing in the product imports this module.  Pinned by
tests/test_sensing_oracle.py against the reference's source tests
(source_test.go known answers) and closed forms.
"""

from __future__ import annotations

import functools
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


# ---- the analytic propagation: soil.HalfSpace and green.HalfSpaceGF ------


@dataclass
class OHalfSpace:
    vp: float
    vs: float
    rho: float
    qs: float
    qp: float = 0.0


def rayleigh_velocity(h: OHalfSpace) -> float:
    """soil.HalfSpace.RayleighVelocity (soil.go:136-163): bisection on
    (2 - x)^2 - 4 sqrt(1 - x) sqrt(1 - x k^2), x = (c / Vs)^2."""
    return _rayleigh(h.vp, h.vs)


@functools.lru_cache(maxsize=None)
def _rayleigh(vp: float, vs: float) -> float:
    k = vs / vp
    k2 = k * k

    def f(x):
        return (2 - x) * (2 - x) - 4 * math.sqrt(1 - x) * math.sqrt(1 - x * k2)

    lo, hi = 1e-9, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    return math.sqrt(0.5 * (lo + hi)) * vs


def eigenfunctions(h: OHalfSpace, omega: float):
    """NewEigenfunctions (green.go:58-80): (k, q, s, xq, xs, zq, zs)."""
    c = rayleigh_velocity(h)
    k = omega / c
    q = k * math.sqrt(1 - (c * c) / (h.vp * h.vp))
    s = k * math.sqrt(1 - (c * c) / (h.vs * h.vs))
    k2, s2 = k * k, s * s
    return k, q, s, k, -2 * k * q * s / (k2 + s2), -q, 2 * k2 * q / (k2 + s2)


def energy_integral(e, rho: float) -> float:
    k, q, s, xq, xs, zq, zs = e

    def sq(a, b):
        return a * a / (2 * q) + 2 * a * b / (q + s) + b * b / (2 * s)

    return 0.5 * rho * (sq(xq, xs) + sq(zq, zs))


def velocity_response(h: OHalfSpace, r: float, f: float, ref: float = 30.0) -> complex:
    """HalfSpaceGF.VelocityResponse (green.go:218-266) with ComplexWavenumber
    and PhaseVelocity (green.go:169-198)."""
    if f <= 0:
        return 0j
    cr = rayleigh_velocity(h)
    gam = math.atan(1 / h.qs) / math.pi
    c0 = cr * math.cos(math.pi * gam / 2)
    omega = 2 * math.pi * f
    kc = (omega / c0 * (f / ref) ** (-gam)) * complex(math.cos(-math.pi * gam / 2), math.sin(-math.pi * gam / 2))
    c = cr * (f / ref) ** gam
    e = eigenfunctions(h, omega)
    i1 = energy_integral(e, h.rho)
    uz = e[5] + e[6]
    amp = uz * uz / (8 * c * c * i1) * math.sqrt(2 / (math.pi * kc.real * r))
    prop = np.exp(complex(0, -math.pi / 4) - 1j * kc * r)
    return complex(0, omega) * amp * prop


def radial_force_response(h: OHalfSpace, r: float, f: float, ref: float = 30.0) -> complex:
    """HalfSpaceGF.RadialForceResponse (green.go:375-388): i (H(0)/V(0)) times
    the vertical response."""
    v = velocity_response(h, r, f, ref)
    if v == 0:
        return 0j
    e = eigenfunctions(h, 2 * math.pi * f)
    return 1j * ((e[3] + e[4]) / (e[5] + e[6])) * v


def causal_impulse_response(n: int, length: int, fs: float, response) -> np.ndarray:
    """dsp.CausalImpulseResponseAt (dsp.go:232-260): n = 0 picks
    NextPow2(4 length); taper above 0.8 of Nyquist; numpy's inverse real
    transform in place of dsp.IRFFT; the first `length` samples."""
    if n == 0:
        n = 1
        while n < 4 * length:
            n <<= 1
    nyq = fs / 2
    frm = 0.8 * nyq
    coeff = np.zeros(n // 2 + 1, dtype=complex)
    for k in range(n // 2 + 1):
        f = k * fs / n
        v = response(f)
        if f > frm:
            v = v * (0.5 * (1 + math.cos(math.pi * (f - frm) / (nyq - frm))))
        coeff[k] = v
    return np.fft.irfft(coeff, n)[:length]


def reference_xyz(contacts: List[OContact], forces, n: int, fs: float, impulse, rx: float = 0.0,
                  ry: float = 0.0) -> np.ndarray:
    """sensing.Reference (sensing.go:353-386) with both components: forces[i]
    maps times to (fx, fy, fz); impulse(rq) returns (vertical, radial)
    responses; the radial one is driven by the force along the contact-to-
    receiver line."""
    import stream_oracle

    span = float(n) / fs
    out = np.zeros(n)
    t = np.arange(n) / fs
    cache = {}
    for c, force in zip(contacts, forces):
        if not (0 <= c.start < span):
            continue
        dx, dy = rx - c.x, ry - c.y
        r = math.hypot(dx, dy)
        rq = quantized_range(r)
        if rq not in cache:
            cache[rq] = impulse(rq)
        hz, hr = cache[rq]
        fx, fy, fz = force(t)
        out += stream_oracle.Monolithic(fz, hz) + stream_oracle.Monolithic(fx * (dx / r) + fy * (dy / r), hr)
    return out
