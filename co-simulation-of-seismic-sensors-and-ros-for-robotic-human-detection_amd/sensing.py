"""Python face of runtime synthesis in libgeofdtd: the reference's sensing
engine driven by a bank (internal/sensing/sensing.go with
sensing.FromBank, propagation.go:79-116) and the walking schedule it
consumes (internal/source/source.go), over the C ABI.

    Walk, Footfall, Contact, RangeSpan   source.go:39-263 (host logic: which
                                         contacts begin when, and where)
    FromBank(bank)                       propagation.go:79-84
    HalfSpace, HalfSpaceGF, Analytic     soil.go:25-163, green.go:120-130,
                                         propagation.go:38-62
    NewEngine(...)                       sensing.go:110-113 (analytic)
    Geophone, SM24                       geophone.go:36-203 (the sensor)
    NewEngineWith(...)                   sensing.go:117-170
    Engine.Next, ChunkSize, SampleRate, ChunkDuration, StartTime,
    ActiveVoices, Propagation            sensing.go:224-330
    Reference(...)                       sensing.go:344-386
    RangeQuantum                         sensing.go:49

The engine's state (impulse responses per range quantum, each voice's force
record) lives in HBM and every chunk is mixed on the GPU for all receivers
in one launch.  The reference's NewEngineWith takes a resolved config; this
mirror takes the numbers it reads from one: the sampling rate, the walk's
duration, the geometry's range (the fallback furthest range), the soil's
Rayleigh velocity, and the sensor's frequency response.  Errors the Go code
returns raise fdtd.FdtdError with the same text.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import fdtd, stream
from .workloads import Stance, _DUTY

RangeQuantum = 0.02
StanceFraction = _DUTY


# ------------------------------------------------------------------ schedule


@dataclass
class Footfall:
    """source.Footfall (source.go:77-107): a stance resolved to a heading."""

    Stance: Stance
    Heading: float = 0.0

    def Duration(self) -> float:
        return self.Stance.Duration

    def ForceAt(self, since):
        """(fx, fy, fz) in newtons at times since heel strike (scalar or
        array); zero outside (0, Duration)."""
        since = np.asarray(since, dtype=np.float64)
        tau = since / self.Stance.Duration
        w = self.Stance.body_weight()
        inside = (since > 0) & (since < self.Stance.Duration)
        ap = np.where(inside, self.Stance.anterior_posterior_at(tau) * w, 0.0)
        fz = np.where(inside, self.Stance.profile_at(tau) * w, 0.0)
        return ap * math.cos(self.Heading), ap * math.sin(self.Heading), fz


@dataclass
class Contact:
    """source.Contact (source.go:39-58)."""

    ID: str
    X: float
    Y: float
    Start: float
    Profile: Footfall

    def End(self) -> float:
        return self.Start + self.Profile.Duration()

    def ForceAt(self, t):
        return self.Profile.ForceAt(np.asarray(t, dtype=np.float64) - self.Start)


@dataclass
class Walk:
    """source.Walk (source.go:109-252): two feet alternating along a line."""

    Stance: Stance
    Speed: float
    StartX: float = 0.0
    StartY: float = 0.0
    Heading: float = 0.0
    StrideLength: float = 0.0
    Width: float = 0.0
    FirstFoot: str = ""
    SecondFoot: str = ""
    Until: float = 0.0

    def Validate(self) -> None:
        if self.Speed <= 0:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID,
                                 f"source: walking speed must be positive, got {fdtd._go_g(self.Speed)} m/s")
        if self.Width < 0:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID,
                                 f"source: stance width must not be negative, got {fdtd._go_g(self.Width)} m")
        if self.StrideLength < 0:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, "source: stride length must not be negative, got "
                                 f"{fdtd._go_g(self.StrideLength)} m")
        self.Stance.Validate()

    def StrideOrDefault(self) -> float:
        if self.StrideLength > 0:
            return self.StrideLength
        return self.Speed * self.Stance.Duration / StanceFraction

    def width(self) -> float:
        return self.Width if self.Width > 0 else 0.12

    def feet(self) -> Tuple[str, str]:
        return self.FirstFoot or "left_foot", self.SecondFoot or "right_foot"

    def StepPeriod(self) -> float:
        return self.StrideOrDefault() / 2 / self.Speed

    def ContactAt(self, n: int) -> Contact:
        step = self.StrideOrDefault() / 2
        along = float(n) * step
        side = self.width() / 2
        if n % 2 == 1:
            side = -side
        cos, sin = math.cos(self.Heading), math.sin(self.Heading)
        first, second = self.feet()
        return Contact(ID=second if n % 2 == 1 else first, X=self.StartX + along * cos - side * sin,
                       Y=self.StartY + along * sin + side * cos, Start=float(n) * self.StepPeriod(),
                       Profile=Footfall(self.Stance, self.Heading))

    def Contacts(self, frm: float, to: float) -> List[Contact]:
        if to <= frm:
            return []
        period = self.StepPeriod()
        if period <= 0:
            return []
        if self.Until > 0 and frm >= self.Until:
            return []
        if self.Until > 0 and to > self.Until:
            to = self.Until
        lo = max(int(math.ceil(frm / period)), 0)
        hi = int(math.ceil(to / period))
        out = []
        for n in range(lo, hi):
            c = self.ContactAt(n)
            if frm <= c.Start < to:
                out.append(c)
        return out


@dataclass
class Fixed:
    """A fixed list of contacts as a schedule (what the reference's tests use)."""

    contacts: List[Contact] = field(default_factory=list)

    def Contacts(self, frm: float, to: float) -> List[Contact]:
        return [c for c in self.contacts if frm <= c.Start < to]


def RangeSpan(s, rx: float, ry: float, frm: float, to: float):
    """source.RangeSpan (source.go:254-263): (min, max, ok)."""
    lo, hi, ok = math.inf, 0.0, False
    for c in s.Contacts(frm, to):
        r = math.hypot(rx - c.X, ry - c.Y)
        lo, hi, ok = min(lo, r), max(hi, r), True
    return lo, hi, ok


# ------------------------------------------------------------------ sensor


@dataclass
class Coupling:
    """geophone.Coupling (geophone.go:88-97): the spike-to-ground mount as a
    second-order low-pass; ResonanceFreq 0 is perfect coupling."""

    ResonanceFreq: float = 0.0
    Damping: float = 0.0


def WellPlantedSpike() -> Coupling:
    return Coupling(200.0, 0.5)


@dataclass
class Geophone:
    """geophone.Geophone (geophone.go:36-62): a moving-coil velocity
    transducer behind its ground coupling; Response (geophone.go:201-203) is
    what the engine multiplies every propagation response by
    (sensing.go:189-215)."""

    NaturalFreq: float
    Sensitivity: float
    CoilResistance: float
    ShuntResistance: float = 0.0
    OpenCircuitDamping: float = 0.0
    MovingMass: float = 0.0
    Coupling: Coupling = field(default_factory=Coupling)

    def Validate(self) -> None:
        """geophone.go:104-121."""
        g = fdtd._go_g
        checks = [(self.NaturalFreq <= 0, f"geophone: natural frequency must be positive, got {g(self.NaturalFreq)} Hz"),
                  (self.Sensitivity <= 0, f"geophone: sensitivity must be positive, got {g(self.Sensitivity)} V/(m/s)"),
                  (self.CoilResistance <= 0,
                   f"geophone: coil resistance must be positive, got {g(self.CoilResistance)} ohm"),
                  (self.MovingMass <= 0, f"geophone: moving mass must be positive, got {g(self.MovingMass)} kg"),
                  (self.OpenCircuitDamping <= 0,
                   f"geophone: open-circuit damping must be positive, got {g(self.OpenCircuitDamping)}"),
                  (self.Coupling.ResonanceFreq < 0,
                   f"geophone: coupling resonance must not be negative, got {g(self.Coupling.ResonanceFreq)} Hz")]
        for bad, message in checks:
            if bad:
                raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, message)

    def _w0(self) -> float:
        return 2 * math.pi * self.NaturalFreq

    def Damping(self) -> float:
        """geophone.go:134-144: open-circuit damping plus the electrical
        damping of the coil and shunt load."""
        if self.ShuntResistance <= 0:
            return self.OpenCircuitDamping
        total = self.CoilResistance + self.ShuntResistance
        return self.OpenCircuitDamping + self.Sensitivity ** 2 / (2 * self._w0() * self.MovingMass * total)

    def ShuntForDamping(self, want: float) -> float:
        """geophone.go:146-165."""
        if want <= self.OpenCircuitDamping:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, f"geophone: damping {fdtd._go_g(want)} is not above the "
                                 f"open-circuit value {fdtd._go_g(self.OpenCircuitDamping)}; a shunt can only add "
                                 "damping")
        total = self.Sensitivity ** 2 / (2 * self._w0() * self.MovingMass * (want - self.OpenCircuitDamping))
        rs = total - self.CoilResistance
        if rs <= 0:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, f"geophone: damping {fdtd._go_g(want)} needs a load of "
                                 f"{fdtd._go_g(total)} ohm, below the coil's own {fdtd._go_g(self.CoilResistance)} ohm")
        return rs

    def _division(self) -> float:
        if self.ShuntResistance <= 0:
            return 1.0
        return self.ShuntResistance / (self.CoilResistance + self.ShuntResistance)

    def TransducerResponse(self, f):
        """geophone.go:175-186: G s^2 / (s^2 + 2 zeta w0 s + w0^2), s = i w,
        volts per (m/s)."""
        s = 1j * (2 * math.pi * np.asarray(f, dtype=np.float64))
        w0 = self._w0()
        return (self.Sensitivity * self._division()) * s * s / (s * s + (2 * self.Damping() * w0) * s + w0 * w0)

    def CouplingResponse(self, f):
        """geophone.go:190-199."""
        f = np.asarray(f, dtype=np.float64)
        if self.Coupling.ResonanceFreq <= 0:
            return np.ones_like(f, dtype=np.complex128)
        s = 1j * (2 * math.pi * f)
        wc = 2 * math.pi * self.Coupling.ResonanceFreq
        return (wc * wc) / (s * s + (2 * self.Coupling.Damping * wc) * s + wc * wc)

    def Response(self, f):
        return self.CouplingResponse(f) * self.TransducerResponse(f)

    __call__ = Response


def SM24() -> Geophone:
    """geophone.SM24 (geophone.go:64-75): the common 4.5 Hz moving-coil
    geophone, well planted."""
    return Geophone(NaturalFreq=4.5, Sensitivity=28.8, CoilResistance=375, ShuntResistance=0,
                    OpenCircuitDamping=0.34, MovingMass=0.011, Coupling=WellPlantedSpike())


# -------------------------------------------------------------- propagation


class FromBank:
    """sensing.FromBank (propagation.go:79-116) over a device bank."""

    def __init__(self, b):
        if b is None:
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, "sensing: bank is nil")
        self.bank = b

    def TransformSize(self) -> int:
        return self.bank.header.Samples

    def Describe(self) -> str:
        h = self.bank.header
        return (f"bank: {h.Solver}, {h.Count} ranges {fdtd._go_g(h.MinM)}-{fdtd._go_g(h.MaxM)} m, "
                f"{h.Component}")


@dataclass
class HalfSpace:
    """soil.HalfSpace (soil.go:25-40)."""

    Vp: float
    Vs: float
    Density: float
    Qs: float
    Qp: float = 0.0

    def to_c(self) -> fdtd.GfHalfspace:
        return fdtd.GfHalfspace(self.Vp, self.Vs, self.Density, self.Qs, self.Qp)

    def RayleighVelocity(self) -> float:
        """soil.go:136-163, on the library's bisection; Validate's errors."""
        out = C.c_double()
        h = self.to_c()
        fdtd._check(fdtd.load_library().gf_rayleigh_velocity(C.byref(h), C.byref(out)))
        return out.value


def Loam() -> HalfSpace:
    """soil.Loam (soil.go:50-52)."""
    return HalfSpace(500, 200, 1700, 25)


@dataclass
class HalfSpaceGF:
    """green.HalfSpaceGF (green.go:120-130): RefFreq 0 means 30 Hz."""

    Soil: HalfSpace
    RefFreq: float = 0.0


class Analytic:
    """sensing.Analytic (propagation.go:38-62): the homogeneous half-space in
    far-field form, with its radial-force response; the engine builds its
    responses on the device."""

    def __init__(self, gf: HalfSpaceGF):
        self.gf = gf

    def TransformSize(self) -> int:
        return 0

    def Describe(self) -> str:
        h = self.gf.Soil
        vp, vs = h.Vp, h.Vs
        nu = (vp * vp - 2 * vs * vs) / (2 * (vp * vp - vs * vs))
        return (f"homogeneous half-space, far-field Rayleigh (Vp={vp:.0f} Vs={vs:.0f} rho={h.Density:.0f} "
                f"Qs={h.Qs:.0f} nu={nu:.3f})")


# ------------------------------------------------------------------- engine


class Engine:
    """sensing.Engine (sensing.go:51-80) for one or many receivers."""

    def __init__(self, handle, prop, schedule, chunk, fs, receivers):
        self._h = handle
        self._prop = prop
        self._schedule = schedule
        self.chunk = chunk
        self.fs = fs
        self.receivers = receivers
        self._single = receivers.shape[0] == 1
        self.scheduled = 0.0
        self._records = {}  # (ID, Start, X, Y) -> (admission sample, force record(s))
        self._xyz = isinstance(prop, Analytic)

    def close(self):
        if getattr(self, "_h", None):
            fdtd.load_library().gf_engine_destroy(self._h)
            self._h = None

    __del__ = close

    def _info(self):
        em, act, resp, irl = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32()
        fdtd._check(fdtd.load_library().gf_engine_info(self._h, C.byref(em), C.byref(act), C.byref(resp),
                                                       C.byref(irl)))
        return em.value, act.value, resp.value, irl.value

    def Propagation(self) -> str:
        return self._prop.Describe()

    def ChunkSize(self) -> int:
        return self.chunk

    def SampleRate(self) -> float:
        return self.fs

    def ChunkDuration(self) -> float:
        return float(self.chunk) / self.fs

    def StartTime(self) -> float:
        return float(self._info()[0]) / self.fs

    def ActiveVoices(self) -> int:
        return self._info()[1]

    def Responses(self) -> int:
        return self._info()[2]

    def ImpulseLength(self) -> int:
        return self._info()[3]

    def _warm(self, frm: float, to: float) -> None:
        cs = self._schedule.Contacts(frm, to)
        if not cs:
            return
        xy = np.ascontiguousarray([[c.X, c.Y] for c in cs], dtype=np.float64)
        fdtd._check(fdtd.load_library().gf_engine_warm(self._h, fdtd._dptr(xy), len(cs)))
        # Each contact's force record too: its admission chunk is fixed by
        # its start, so no chunk has to sample one.
        for c in cs:
            key = (c.ID, c.Start, c.X, c.Y)
            a = self._admission(c)
            if a is not None and a >= self._info()[0] and key not in self._records:
                self._records[key] = (a, self._force(c, a))

    def _admission(self, c: Contact) -> Optional[int]:
        """The first sample of the chunk whose window [t0, t1) holds the
        contact's start, with the window bounds Next computes."""
        k0 = int(c.Start * self.fs) // self.chunk
        for k in (k0 - 1, k0, k0 + 1):
            if k < 0:
                continue
            t0 = float(k * self.chunk) / self.fs
            if t0 <= c.Start < t0 + self.ChunkDuration():
                return k * self.chunk
        return None

    def _force(self, c: Contact, emitted: int):
        """The contact's force at the instants the reference's Next evaluates
        it (t0 + i / fs per chunk, t0 = emitted / fs), from this chunk until
        the force has surely ended: the vertical component, or (fx, fy, fz)
        for an engine that also carries the radial-force response."""
        last = int(math.ceil(c.End() * self.fs)) + 2
        n = max(last - emitted, 1)
        k = np.arange(n)
        starts = emitted + (k // self.chunk) * self.chunk
        t = starts / self.fs + (k % self.chunk) / self.fs
        f = c.ForceAt(t)
        if self._xyz:
            return tuple(np.ascontiguousarray(np.broadcast_to(v, t.shape), dtype=np.float64) for v in f)
        return np.ascontiguousarray(f[2], dtype=np.float64)

    def Next(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """One chunk of output (volts, float32): [chunk] for one receiver,
        [receivers][chunk] for several.  out, when given, is filled."""
        lib = fdtd.load_library()
        if out is not None and np.shape(out)[-1:] != (self.chunk,):
            n_out = np.shape(out)[-1] if np.ndim(out) else 0
            raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, f"sensing: output length {n_out}, want {self.chunk}")
        emitted = self._info()[0]
        t0 = float(emitted) / self.fs
        t1 = t0 + self.ChunkDuration()
        # admit (sensing.go:295-330): keep the responses warmed ahead, then a
        # voice for every contact beginning in this chunk.
        lookahead = 2.0
        if t1 + lookahead > self.scheduled:
            self._warm(self.scheduled, t1 + lookahead)
            self.scheduled = t1 + lookahead
        for c in self._schedule.Contacts(t0, t1):
            a, f = self._records.pop((c.ID, c.Start, c.X, c.Y), (None, None))
            if a != emitted:
                f = self._force(c, emitted)
            if self._xyz:
                fx, fy, fz = f
                fdtd._check(lib.gf_engine_admit_xyz(self._h, c.ID.encode(), c.X, c.Y, fdtd._dptr(fx), fdtd._dptr(fy),
                                                    fdtd._dptr(fz), fz.size, c.End()))
            else:
                fdtd._check(lib.gf_engine_admit(self._h, c.ID.encode(), c.X, c.Y, fdtd._dptr(f), f.size,
                                                c.End()))
        buf = np.zeros((self.receivers.shape[0], self.chunk), dtype=np.float32)
        fdtd._check(lib.gf_engine_next(self._h, buf.ctypes.data_as(C.POINTER(C.c_float)), self.chunk))
        res = buf[0] if self._single else buf
        if out is not None:
            out[...] = res
        return res

    def Stats(self) -> dict:
        mix, tot, macs, build, built = C.c_double(), C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        fdtd._check(fdtd.load_library().gf_engine_stats(self._h, C.byref(mix), C.byref(tot), C.byref(macs),
                                                        C.byref(build), C.byref(built)))
        return {"mix_ms": mix.value, "mix_ms_total": tot.value, "macs": macs.value, "build_ms": build.value,
                "built": built.value}


def _go_hypot(p: float, q: float) -> float:
    """math.Hypot as Go computes it (the range every quantum is taken from)."""
    p, q = abs(p), abs(q)
    if math.isinf(p) or math.isinf(q):
        return math.inf
    if p < q:
        p, q = q, p
    if p == 0:
        return 0.0
    q = q / p
    return p * math.sqrt(1 + q * q)


def _quantized_range(r: float) -> float:
    """responseFor's range (sensing.go:186-190): r to the nearest
    RangeQuantum, halves away from zero (math.Round), at least one quantum."""
    x = r / RangeQuantum
    k = math.floor(x)
    if x - k >= 0.5:
        k += 1
    return max(float(k) * RangeQuantum, RangeQuantum)


def Reference(schedule, n: int, prop: FromBank, *, sample_rate: float, walk_duration: float, geometry_range: float,
              rayleigh_velocity: float, sensor: Optional[Callable[[np.ndarray], np.ndarray]] = None,
              receiver: Sequence[float] = (0.0, 0.0)) -> np.ndarray:
    """sensing.Reference (sensing.go:344-386) on the device: the same signal an
    engine produces, in one piece.  For every contact beginning in [0, n/fs),
    its vertical force at i/fs convolved with the impulse response at its
    quantized range over the whole record (the responses built in one batch,
    every contact's convolution in one launch), summed in contact order;
    float64 [n]."""
    fs = sample_rate
    if prop is None:
        raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, "sensing: a propagation model is required")
    if not isinstance(prop, FromBank):
        raise fdtd.FdtdError(fdtd.GF_ERR_ARGUMENT, "geofdtd: Reference on the device takes a bank's propagation")
    rx, ry = float(receiver[0]), float(receiver[1])
    transform = prop.TransformSize()
    ir_len = ImpulseLength(schedule, walk_duration, geometry_range, rayleigh_velocity, fs, [(rx, ry)])
    ir_len = min(ir_len, 3 * transform // 4)
    out = np.zeros(n)
    contacts = schedule.Contacts(0, float(n) / fs)
    if not contacts or n <= 0:
        return out
    rq = [_quantized_range(_go_hypot(rx - c.X, ry - c.Y)) for c in contacts]
    uniq = list(dict.fromkeys(rq))  # first use first, as the Go loop meets them
    sens = None
    if sensor is not None:
        h = prop.bank.header
        sens = np.asarray(sensor(np.arange(h.Bins) * fs / h.Samples), dtype=np.complex128)
    irs = prop.bank.ImpulseResponses(uniq, ir_len, fs, sens)
    h_rows = np.stack([irs[uniq.index(r)] for r in rq])
    t = np.arange(n) / fs
    x = np.stack([np.asarray(c.ForceAt(t)[2], dtype=np.float64) for c in contacts])
    conv = stream.NewConvolver(h_rows, min(n, 1024))
    try:
        y = conv.ProcessAll(x)
    finally:
        conv.close()
    for row in y:
        out += row
    return out


def ImpulseLength(schedule, walk_duration: float, geometry_range: float, rayleigh_velocity: float, fs: float,
                  receivers=((0.0, 0.0),)) -> int:
    """The length NewEngineWith gives every impulse response
    (sensing.go:135-140): the furthest travel time plus two seconds of coda,
    the furthest range over every receiver."""
    far, any_ = 0.0, False
    for rx, ry in receivers:
        _, f, ok = RangeSpan(schedule, rx, ry, 0, walk_duration)
        if ok:
            far, any_ = max(far, f), True
    if not any_:
        far = geometry_range
    return int(math.ceil((far / rayleigh_velocity + 2) * fs))


def NewEngineWith(schedule, chunk: int, prop: Optional[FromBank], *, sample_rate: float, walk_duration: float,
                  geometry_range: float, rayleigh_velocity: float,
                  sensor: Optional[Callable[[np.ndarray], np.ndarray]] = None,
                  receivers: Sequence[Sequence[float]] = ((0.0, 0.0),)) -> Engine:
    """sensing.NewEngineWith (sensing.go:117-170) with a bank's propagation.

    sensor maps an array of frequencies (Hz) to the sensor's complex response
    (None: 1).  receivers are (x, y) positions in metres; the reference's
    engine listens at the origin."""
    if chunk <= 0:
        raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, f"sensing: chunk size must be positive, got {chunk}")
    if schedule is None:
        raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, "sensing: a schedule is required")
    if prop is None:
        raise fdtd.FdtdError(fdtd.GF_ERR_INVALID, "sensing: a propagation model is required")
    if isinstance(sensor, Geophone):
        sensor.Validate()
    rec = np.ascontiguousarray(np.atleast_2d(np.asarray(receivers, dtype=np.float64)))
    ir_len = ImpulseLength(schedule, walk_duration, geometry_range, rayleigh_velocity, sample_rate,
                           [tuple(r) for r in rec])
    if isinstance(prop, Analytic):
        n = 1
        while n < 4 * ir_len:  # dsp.NextPow2(4 length), the transform the analytic model leaves to the caller
            n <<= 1
        bins, samples = n // 2 + 1, n
    else:
        bins, samples = prop.bank.header.Bins, prop.bank.header.Samples
    sens = None
    if sensor is not None:
        f = np.arange(bins) * sample_rate / samples
        g = np.asarray(sensor(f), dtype=np.complex128)
        sens = np.ascontiguousarray(np.stack([g.real, g.imag], axis=-1))
    # The longest force record one voice holds: admitted up to a chunk before
    # its heel strike, sampled until two samples past toe-off.
    longest = max([c.Profile.Duration() for c in schedule.Contacts(0, walk_duration + 2)] or [0.0])
    cap = int(math.ceil(longest * sample_rate)) + chunk + 8
    handle = C.c_void_p()
    lib = fdtd.load_library()
    if isinstance(prop, Analytic):
        hs = prop.gf.Soil.to_c()
        fdtd._check(lib.gf_engine_create_analytic(C.byref(hs), prop.gf.RefFreq, rec.shape[0], fdtd._dptr(rec),
                                                  int(chunk), sample_rate, ir_len,
                                                  fdtd._dptr(sens) if sens is not None else None, cap,
                                                  C.byref(handle)))
    else:
        fdtd._check(lib.gf_engine_create(prop.bank._h, rec.shape[0], fdtd._dptr(rec), int(chunk), sample_rate,
                                         ir_len, fdtd._dptr(sens) if sens is not None else None, cap,
                                         C.byref(handle)))
    e = Engine(handle, prop, schedule, int(chunk), sample_rate, rec)
    try:
        e._warm(0, walk_duration)
    except fdtd.FdtdError:
        e.close()
        raise
    return e


def NewEngine(schedule, chunk: int, soil: HalfSpace, *, sample_rate: float, walk_duration: float,
              geometry_range: float, sensor: Optional[Callable[[np.ndarray], np.ndarray]] = None,
              receivers: Sequence[Sequence[float]] = ((0.0, 0.0),)) -> Engine:
    """sensing.NewEngine (sensing.go:110-113): the analytic half-space
    propagation of the resolved soil."""
    prop = Analytic(HalfSpaceGF(soil))
    return NewEngineWith(schedule, chunk, prop, sample_rate=sample_rate, walk_duration=walk_duration,
                         geometry_range=geometry_range, rayleigh_velocity=soil.RayleighVelocity(), sensor=sensor,
                         receivers=receivers)
