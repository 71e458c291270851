"""The named workloads: models, sources and receivers of the benchmark
configurations, built through the solver's own API.

    c1(variant)   homogeneous half-space, 520 x 270 cells at 0.1 m, 10,000 steps
    c2(variant)   two layers over a half-space, 2048 x 1024 cells at 0.05 m
    c3(...)       seeded soil models for batched bank generation

variant "R" is the physics the reference can express (one scalar standard
linear solid, Qp = Qs, one Q for the whole model); "X" is the extension
(per-layer Q, Qp != Qs, three mechanisms fitted over 5-150 Hz).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np

from . import fdtd as fd


@dataclass
class Workload:
    name: str
    model: fd.Model
    force: np.ndarray
    ranges: List[float]
    steps: int

    def shot(self) -> fd.Shot:
        return fd.Shot(self.model, self.force, self.ranges, self.steps)


def _courant_for(model: fd.Model, dt: float) -> float:
    """The Courant factor that makes New's time step equal dt."""
    base = model.Grid()[0] / (model.Courant if model.Courant > 0 else 0.5)
    return dt / base


def c1(variant: str = "R", steps: int = 10_000) -> Workload:
    """Config 1: Vp 400, Vs 200, rho 1800; 520 x 270 cells at h = 0.1 m with a
    20-cell C-PML (MaxRange 49.9 m, Depth 24.9 m); vertical Ricker force of
    800 N at 40 Hz on the axis; dt = 1e-4 s; v_z receivers every metre from 1
    to 49 m (the reference rejects 50 m: it is not inside MaxRange)."""
    if variant == "R":
        stack = [fd.Layer(0, 400, 200, 1800, 30)]
        model = fd.Model(Stack=stack, Relax=fd.Fit(30, 30), RefFreq=30)
    elif variant == "X":
        stack = [fd.Layer(0, 400, 200, 1800, 30, 60)]
        model = fd.Model(Stack=stack, RefFreq=30, GslsMechanisms=3, GslsFmin=5, GslsFmax=150)
    else:
        raise ValueError(f"unknown variant {variant!r}")
    model.MaxRange, model.Depth, model.Spacing = 49.9, 24.9, 0.1
    model.PMLCells, model.DominantFreq = 20, 40
    model.Courant = _courant_for(model, 1e-4)
    dt = model.Grid()[0]
    force = fd.Ricker(800, 40, dt, steps)
    ranges = [float(r) for r in range(1, 50)]
    return Workload(f"c1-{variant}", model, force, ranges, steps)


# ----------------------------------------------------------------- GRF source
# A restatement of internal/grf/grf.go's Stance (grf.go:34-103) as far as the
# vertical force series goes: ProfileAt (grf.go:280-299), transientAt
# (grf.go:230-244), tukey (grf.go:303-314), Sample (grf.go:324-341),
# ImpulseRatio and BalancedValley (grf.go:360-416).  It produces the source
# time function of config 2; it is input generation, not the solver.

_GRAVITY = 9.80665
_HUMP_WIDTH, _TRANSIENT_PEAK, _TRANSIENT_RISE = 0.11, 0.35, 0.012
_TAPER, _P1, _P2, _DUTY = 0.098, 0.25, 0.75, 0.60
_AP_PEAK = 0.20
_CREST, _CREST_PHASE = 0.63790, 0.20090


@dataclass
class Stance:
    Mass: float
    Duration: float = 0.62
    FirstPeak: float = 1.15
    SecondPeak: float = 1.12
    MidstanceValley: float = 0.75
    TransientRise: float = 0.0
    APPeak: float = 0.0

    def Validate(self) -> None:
        """Stance.Validate (grf.go:145-163) for the fields this mirror carries."""
        if self.Mass <= 0:
            raise fd.FdtdError(fd.GF_ERR_INVALID, f"grf: mass must be positive, got {fd._go_g(self.Mass)} kg")
        if self.Duration <= 0:
            raise fd.FdtdError(fd.GF_ERR_INVALID,
                               f"grf: stance duration must be positive, got {fd._go_g(self.Duration)} s")
        if self.FirstPeak <= 0 or self.SecondPeak <= 0:
            raise fd.FdtdError(fd.GF_ERR_INVALID, f"grf: peaks must be positive, got {fd._go_g(self.FirstPeak)} "
                               f"and {fd._go_g(self.SecondPeak)} BW")
        if self.MidstanceValley <= 0:
            raise fd.FdtdError(fd.GF_ERR_INVALID,
                               f"grf: midstance valley must be positive, got {fd._go_g(self.MidstanceValley)} BW")
        if self.MidstanceValley >= self.FirstPeak or self.MidstanceValley >= self.SecondPeak:
            raise fd.FdtdError(fd.GF_ERR_INVALID, f"grf: valley {fd._go_g(self.MidstanceValley)} BW must lie below "
                               f"both peaks ({fd._go_g(self.FirstPeak)}, {fd._go_g(self.SecondPeak)} BW), or the "
                               "profile is not a double hump")

    def anterior_posterior_at(self, tau):
        """Stance.AnteriorPosteriorAt (grf.go:264-270): the fore-aft shear in
        body weights, negative (braking) in the first half of stance."""
        tau = np.asarray(tau, dtype=np.float64)
        a = self.APPeak if self.APPeak > 0 else (0.0 if self.APPeak < 0 else _AP_PEAK)
        r = _TAPER
        tk = np.where(tau < r, 0.5 * (1 - np.cos(math.pi * tau / r)),
                      np.where(tau > 1 - r, 0.5 * (1 - np.cos(math.pi * (1 - tau) / r)), 1.0))
        out = tk * -a * np.sin(2 * math.pi * tau)
        return np.where((tau <= 0) | (tau >= 1) | (a == 0), 0.0, out)

    def _rise(self):
        return self.TransientRise if self.TransientRise > 0 else _TRANSIENT_RISE

    def transient_at(self, tau):
        tau = np.asarray(tau, dtype=np.float64)
        period = self._rise() / _CREST_PHASE / self.Duration
        decay = period / 2
        out = _TRANSIENT_PEAK / _CREST * np.exp(-tau / decay) * np.sin(2 * math.pi * tau / period)
        return np.where(tau <= 0, 0.0, out)

    def profile_at(self, tau):
        tau = np.asarray(tau, dtype=np.float64)
        w = _HUMP_WIDTH
        g1 = np.exp(-(((tau - _P1) / w) ** 2))
        g2 = np.exp(-(((tau - _P2) / w) ** 2))
        body = self.MidstanceValley + (self.FirstPeak - self.MidstanceValley) * g1 + (
            self.SecondPeak - self.MidstanceValley) * g2
        r = _TAPER
        tk = np.where(tau < r, 0.5 * (1 - np.cos(math.pi * tau / r)),
                      np.where(tau > 1 - r, 0.5 * (1 - np.cos(math.pi * (1 - tau) / r)), 1.0))
        out = tk * body + self.transient_at(tau)
        return np.where((tau <= 0) | (tau >= 1), 0.0, out)

    def body_weight(self):
        return self.Mass * _GRAVITY

    def impulse_ratio(self):
        steps = 20000
        tau = (np.arange(steps) + 0.5) / steps
        impulse = self.profile_at(tau).sum() / steps * self.Duration * self.body_weight()
        cycle = self.Duration / _DUTY
        return 2 * impulse / (self.body_weight() * cycle)

    def balanced_valley(self):
        def f(v):
            return Stance(self.Mass, self.Duration, self.FirstPeak, self.SecondPeak, v,
                          self.TransientRise).impulse_ratio() - 1

        lo, hi = 0.05, 1.5
        if f(lo) > 0 or f(hi) < 0:
            raise ValueError("grf: no valley balances the momentum of this stance")
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if f(mid) < 0:
                lo = mid
            else:
                hi = mid
        return 0.5 * (lo + hi)

    def sample(self, fs, lead, tail):
        total = lead + self.Duration + tail
        n = int(round(total * fs))
        t = np.arange(n) / fs - lead
        return self.profile_at(t / self.Duration) * self.body_weight()


def c2_stance():
    """Config 2's footstep: 80 kg, peaks 1.2 and 1.1 body weights, a 25 ms
    heel transient, the valley solved from the momentum balance."""
    s = Stance(Mass=80, FirstPeak=1.2, SecondPeak=1.1, TransientRise=0.025)
    s.MidstanceValley = s.balanced_valley()
    return s


def c2(variant: str = "R", steps: int = 80_000, lead: float = 0.0) -> Workload:
    """Config 2: layers {1.5 m: Vs 150, 4 m: Vs 250} over a Vs 400 half-space,
    Vp = 2 Vs, rho 1700/1850/2000, Qs 20/40/80; 2048 x 1024 cells at 0.05 m
    with a 32-cell C-PML (MaxRange 100.75 m, Depth 49.55 m); dt = 2.5e-5 s
    (the stated 5e-5 s is above the scheme's stability limit for Vp 800 m/s
    at this spacing); 80,000 steps = 2 s; GRF footstep source; 100 receivers
    at 0.5-50 m.  "R": one scalar SLS at the half-space Q of 80 (what the
    reference can express); "X": per-layer Q, Qp = 2 Qs, 3 mechanisms over
    5-150 Hz."""
    vs = [150.0, 250.0, 400.0]
    rho = [1700.0, 1850.0, 2000.0]
    qs = [20.0, 40.0, 80.0]
    th = [1.5, 4.0, 0.0]
    if variant == "R":
        stack = [fd.Layer(th[i], 2 * vs[i], vs[i], rho[i], qs[i]) for i in range(3)]
        model = fd.Model(Stack=stack, Relax=fd.Fit(80, 30), RefFreq=30)
    elif variant == "X":
        stack = [fd.Layer(th[i], 2 * vs[i], vs[i], rho[i], qs[i], 2 * qs[i]) for i in range(3)]
        model = fd.Model(Stack=stack, RefFreq=30, GslsMechanisms=3, GslsFmin=5, GslsFmax=150)
    else:
        raise ValueError(f"unknown variant {variant!r}")
    model.MaxRange, model.Depth, model.Spacing, model.PMLCells = 100.75, 49.55, 0.05, 32
    model.Courant = _courant_for(model, 2.5e-5)
    dt = model.Grid()[0]
    force = c2_stance().sample(1.0 / dt, lead, 0.0)
    ranges = [0.5 * k for k in range(1, 101)]
    return Workload(f"c2-{variant}", model, force, ranges, steps)


# ----------------------------------------------------- batched bank generation


def c3(n_models: int = 256, seed: int = 0, seconds: float = 1.0, variant: str = "X",
       first: int = 0) -> List[Workload]:
    """Config 3: `n_models` seeded soils (models first..first+n_models-1 of the
    stream keyed by `seed`), each with its own cell size, grid extents and
    time step (Courant 0.5, the reference default), run for `seconds`; a
    40 Hz Ricker source; receivers on the bank's range grid (1-50 m).  "X":
    per-layer Q, Qp = 2 Qs, three mechanisms over 5-150 Hz; "R": one scalar
    SLS at the half-space's Qs (what the reference can express)."""
    out = []
    for i in range(first, first + n_models):
        soil = fd.SampleSoil(seed, i)
        if variant == "X":
            model = fd.Model(Stack=soil.Stack, RefFreq=30, GslsMechanisms=3, GslsFmin=5, GslsFmax=150)
        elif variant == "R":
            stack = [fd.Layer(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs) for l in soil.Stack]
            model = fd.Model(Stack=stack, Relax=fd.Fit(soil.Stack[-1].Qs, 30), RefFreq=30)
        else:
            raise ValueError(f"unknown variant {variant!r}")
        model.MaxRange, model.Depth, model.Spacing, model.PMLCells = (
            soil.MaxRange, soil.Depth, soil.Spacing, soil.PMLCells)
        model.DominantFreq = 40
        dt = model.Grid()[0]
        steps = int(math.ceil(seconds / dt))
        force = fd.Ricker(1000, 40, dt, steps)
        out.append(Workload(f"c3-{variant}-{i}", model, force, [], steps))
    return out


def c3_bank_spec(count: int = 50) -> fd.BankSpec:
    """Bank grid of config 3: 1-50 m, 512 complex bins to 250 Hz (rate 500 Hz,
    1024 samples -> 513 bins; bin 0 and bins above 100 Hz, the band the grid
    resolves, left zero)."""
    return fd.BankSpec(SampleRateHz=500.0, Samples=1024, MaxFreqHz=100.0, MinM=1.0, MaxM=50.0, Count=count)
