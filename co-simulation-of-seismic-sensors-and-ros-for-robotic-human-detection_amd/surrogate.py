"""The Green's-function surrogate: a network standing in for the bank lookup.

It maps a layered soil (10 numbers, gf_soil_sample_layered) and a range to
the bank spectrum G(f) that the FDTD would compute for them (vertical surface
displacement per unit vertical point force, m/N, 512 bins at 500/1022 Hz
spacing, zero above 100 Hz).  It is trained on spectra the batched FDTD
itself generates, runs on the library's BF16 MFMA kernels, and is served
behind an error gate: a fraction of the soils in each request is re-solved by
the batched FDTD, and every sampled soil whose relative L2 error exceeds the
bound is answered by its FDTD solution instead.  No reference counterpart.

Normalization.  The network learns G divided by an analytic prior,
    T(f) = G(f) * 2 pi mu1 r * exp(+i 2 pi f r / c1),
mu1 = rho1 Vs1^2 the top layer's shear modulus, c1 = 0.92 Vs1 about its
Rayleigh speed: the static Boussinesq amplitude ~ 1/(2 pi mu r) and the
travel-time phase of the surface wave come out, leaving an O(1), slowly
rotating target.  A per-query scale and a per-bin unit phase factor leave the
relative L2 unchanged, so the loss is the same in both domains.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import fdtd as fd

FEATURES = 15
BINS = 512
SAMPLE_RATE_HZ = 500.0
SAMPLES = 1022                       # 512 bins, 0 .. 250 Hz
BIN_HZ = SAMPLE_RATE_HZ / SAMPLES
MAX_FREQ_HZ = 100.0
# The band in which an FDTD bank is determined: below ~5 Hz the 40 Hz Ricker
# has under 4 % of its peak spectrum, and dividing by it lets the smallest
# late-time residue of a 1 s run dominate G; above 100 Hz the grid does not
# resolve the field (and the bank is zero there).
BAND_HZ = (5.0, 100.0)
RANGES = (1.0, 50.0, 50)            # bank range grid (min, max, count)
SECONDS = 1.0
SOURCE_PEAK_N, SOURCE_F0_HZ = 1000.0, 40.0

# Sampling intervals of gf_soil_sample_layered's 10 numbers.
PARAM_RANGES = np.array([
    (120, 200), (200, 320), (320, 480),   # Vs1..3, m/s
    (10, 40), (20, 60), (40, 100),        # Qs1..3
    (0.5, 3.0), (2.0, 6.0),               # thicknesses, m
    (1.7, 2.5),                           # Vp/Vs
    (1600, 1800),                         # rho0, kg/m^3
], dtype=np.float64)
RAYLEIGH_FRACTION = 0.92


def bank_spec() -> fd.BankSpec:
    return fd.BankSpec(SAMPLE_RATE_HZ, SAMPLES, MAX_FREQ_HZ, RANGES[0], RANGES[1], RANGES[2])


TRAVEL_MAX_S = RANGES[1] / PARAM_RANGES[0, 0]   # longest r / Vs in the family


def grid_spacing(params: np.ndarray) -> np.ndarray:
    """The family's cell size, h = 0.9 Vs1 / (100 Hz * 10), as the sampler
    computes it."""
    return 0.9 * np.atleast_2d(np.asarray(params, dtype=np.float64))[:, 0] / (100.0 * 10.0)


def grid_thicknesses(params: np.ndarray) -> np.ndarray:
    """[n][2]: the layer thicknesses the grid realizes.  A grid node at depth
    j h belongs to the layer whose top-to-bottom span holds it, an interface
    node going to the layer below (internal/fdtd/model.go:200-209), so layer 1
    holds the nodes with j h < h1 and layers 1-2 those with j h < h1 + h2;
    each realized thickness is its node count times h."""
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    h = grid_spacing(p)

    def nodes_above(z):
        n = np.ceil(z / h)
        n = np.where((n - 1) * h >= z, n - 1, n)  # the count of j >= 0 with j h < z, exactly
        return np.where(n * h < z, n + 1, n)

    n1 = nodes_above(p[:, 6])
    n12 = nodes_above(p[:, 6] + p[:, 7])
    return np.column_stack([n1 * h, (n12 - n1) * h])


def features(params: np.ndarray, r: np.ndarray) -> np.ndarray:
    """[n][15] float32, each mapped to about [-1, 1]: the 10 soil numbers
    (with the two thicknesses as the grid realizes them, which is what the
    FDTD spectra depend on: the drawn values move a layer boundary by up to
    half a cell), log r, r, and the shear travel times r / Vs1, r / Vs2,
    r / Vs3 (the phase of the spectrum is close to linear in these)."""
    p = np.array(np.atleast_2d(np.asarray(params, dtype=np.float64)), copy=True)
    p[:, 6:8] = grid_thicknesses(p)
    r = np.asarray(r, dtype=np.float64)
    lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
    x = np.empty((p.shape[0], FEATURES))
    x[:, :10] = 2 * (p - lo) / (hi - lo) - 1
    x[:, 10] = 2 * np.log(r / RANGES[0]) / math.log(RANGES[1] / RANGES[0]) - 1
    x[:, 11] = 2 * r / RANGES[1] - 1
    x[:, 12:15] = 2 * (r[:, None] / p[:, 0:3]) / TRAVEL_MAX_S - 1
    return x.astype(np.float32)


def prior(params: np.ndarray, r: np.ndarray):
    """(scale, delay) per query: G = scale * T * exp(-i 2 pi f delay)."""
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    r = np.asarray(r, dtype=np.float64)
    vs1, rho1 = p[:, 0], p[:, 9]
    scale = 1.0 / (2 * math.pi * rho1 * vs1 * vs1 * r)
    delay = r / (RAYLEIGH_FRACTION * vs1)
    return scale, delay


def band_mask() -> np.ndarray:
    f = np.arange(BINS) * BIN_HZ
    return (f >= BAND_HZ[0] - 1e-9) & (f <= BAND_HZ[1] + 1e-9)


def band_limit(G: np.ndarray) -> np.ndarray:
    """Spectra with the bins outside BAND_HZ set to zero."""
    out = np.array(G, copy=True)
    out[..., ~band_mask()] = 0
    return out


def normalize(G: np.ndarray, scale: np.ndarray, delay: np.ndarray) -> np.ndarray:
    """Complex spectra [n][512] to the network's target [n][1024] float32
    (zero outside BAND_HZ)."""
    f = np.arange(BINS) * BIN_HZ
    T = band_limit(G) / scale[:, None] * np.exp(2j * math.pi * f[None, :] * delay[:, None])
    out = np.empty((G.shape[0], 2 * BINS), dtype=np.float32)
    out[:, 0::2] = T.real
    out[:, 1::2] = T.imag
    return out


def polar_targets(t: np.ndarray):
    """Cartesian normalized targets [n][1024] (from `normalize`) to the polar
    form the network can also learn: per bin (log|T|, unwrapped phase), and
    weights w = |T|^2 / sum |T|^2 on both entries of in-band bins (zero out
    of band), so that sum w ((a - log|T|)^2 + (phi - arg T)^2) is the relative
    L2 of e^{a + i phi} against T to first order.  The phase is unwrapped
    across the band from its first bin; log|T| is floored at 1e-3 of the
    query's largest |T| (where the weight is negligible anyway)."""
    T = t[:, 0::2].astype(np.float64) + 1j * t[:, 1::2]
    m = band_mask()
    mag = np.abs(T)
    floor = 1e-3 * np.maximum(mag.max(axis=1, keepdims=True), 1e-30)
    a = np.log(np.maximum(mag, floor))
    ph = np.zeros_like(a)
    ph[:, m] = np.unwrap(np.angle(T[:, m]), axis=1)
    w = np.where(m[None, :], mag ** 2, 0.0)
    w /= np.maximum(w.sum(axis=1, keepdims=True), 1e-300)
    out = np.empty_like(t, dtype=np.float32)
    out[:, 0::2] = np.where(m[None, :], a, 0.0)
    out[:, 1::2] = ph
    wt = np.empty_like(out)
    wt[:, 0::2] = w
    wt[:, 1::2] = w
    return out, wt


def rel_l2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-row |a - b| / |b| over complex (or interleaved real) spectra."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    num = np.sqrt(np.sum(np.abs(a - b) ** 2, axis=1))
    den = np.sqrt(np.sum(np.abs(b) ** 2, axis=1))
    return num / np.where(den > 0, den, 1.0)


# -------------------------------------------------------------- the network


class Network:
    """A gf_sur handle: the surrogate network on the GPU."""

    def __init__(self, max_batch: int = 8192, seed: int = 0, band=BAND_HZ, polar: bool = False):
        self._lib = fd.load_library()
        h = C.c_void_p()
        cfg = fd.GfSurConfig(max_batch, seed, band[0] if band else 0.0, band[1] if band else 0.0, int(polar))
        self.polar = polar
        fd._check(self._lib.gf_sur_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.max_batch = max_batch

    def close(self):
        if self._h:
            self._lib.gf_sur_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Layers(self) -> List[fd.GfSurLayer]:
        n = C.c_int32()
        fd._check(self._lib.gf_sur_layers(self._h, C.byref(n)))
        out = []
        for l in range(n.value):
            y = fd.GfSurLayer()
            fd._check(self._lib.gf_sur_layer(self._h, l, C.byref(y)))
            out.append(y)
        return out

    def ParamCount(self) -> int:
        n = C.c_int64()
        fd._check(self._lib.gf_sur_param_count(self._h, C.byref(n)))
        return n.value

    def Params(self) -> np.ndarray:
        out = np.zeros(self.ParamCount(), dtype=np.float32)
        fd._check(self._lib.gf_sur_get_params(self._h, _pf(out)))
        return out

    def SetParams(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.ParamCount():
            raise ValueError("parameter count mismatch")
        fd._check(self._lib.gf_sur_set_params(self._h, _pf(p)))

    def Forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((len(x), 2 * BINS), dtype=np.float32)
        fd._check(self._lib.gf_sur_forward(self._h, _pf(x), len(x), _pf(out)))
        return out

    def Predict(self, x: np.ndarray, scale: np.ndarray, delay: np.ndarray) -> np.ndarray:
        """Physical spectra [n][512] complex64."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        s = np.ascontiguousarray(scale, dtype=np.float64)
        d = np.ascontiguousarray(delay, dtype=np.float64)
        out = np.zeros((len(x), 2 * BINS), dtype=np.float32)
        fd._check(self._lib.gf_sur_predict(self._h, _pf(x), _pd(s), _pd(d), BIN_HZ, len(x), _pf(out)))
        return out.view(np.complex64)

    def LossGrad(self, x: np.ndarray, t: np.ndarray, w: Optional[np.ndarray] = None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        keep = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        wp = None if keep is None else _pf(keep)
        loss = C.c_double()
        g = np.zeros(self.ParamCount(), dtype=np.float32)
        fd._check(self._lib.gf_sur_loss_grad(self._h, _pf(x), _pf(t), wp, len(x), C.byref(loss), _pf(g)))
        return loss.value, g

    def Train(self, x: np.ndarray, t: np.ndarray, steps: int, batch: int = 4096, lr: float = 1e-3,
              lr_min: float = 2e-5, seed: int = 0, w: Optional[np.ndarray] = None) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        keep = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        opts = fd.GfSurTrainOpts(steps, batch, lr, lr_min, 0.0, 0.0, 0.0, seed)
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        fd._check(self._lib.gf_sur_train(self._h, _pf(x), _pf(t), None if keep is None else _pf(keep), len(x),
                                         C.byref(opts), _pf(losses)))
        return losses[:steps]

    def Stats(self) -> Dict[str, float]:
        s = fd.GfSurStats()
        fd._check(self._lib.gf_sur_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in fd.GfSurStats._fields_}


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------- FDTD data


@dataclass
class SoilSet:
    """Soils of the layered family and, once solved, their FDTD banks."""

    seed: int
    indices: np.ndarray                  # soil indices in the family's stream
    params: np.ndarray                   # [S][10]
    soils: list
    spectra: Optional[np.ndarray] = None  # [S][R][512] complex64
    ranges: Optional[np.ndarray] = None   # [S][R] receiver ranges (grid columns)
    fdtd_seconds: float = 0.0
    stats: dict = field(default_factory=dict)


def soil_set(seed: int, first: int, count: int) -> SoilSet:
    idx = np.arange(first, first + count, dtype=np.int64)
    soils, params = [], []
    for i in idx:
        s, p = fd.SampleLayeredSoil(seed, int(i))
        soils.append(s)
        params.append(p)
    return SoilSet(seed, idx, np.asarray(params), soils)


def shots_for(soils: Sequence, seconds: float = SECONDS) -> List[fd.Shot]:
    out = []
    for s in soils:
        m = fd.Model(Stack=s.Stack, RefFreq=30, GslsMechanisms=3, GslsFmin=5, GslsFmax=150)
        m.MaxRange, m.Depth, m.Spacing, m.PMLCells = s.MaxRange, s.Depth, s.Spacing, s.PMLCells
        m.DominantFreq = SOURCE_F0_HZ
        dt = m.Grid()[0]
        steps = int(math.ceil(seconds / dt))
        out.append(fd.Shot(m, fd.Ricker(SOURCE_PEAK_N, SOURCE_F0_HZ, dt, steps), [], steps))
    return out


def receiver_ranges(soils: Sequence) -> np.ndarray:
    """[S][R]: the ranges the bank grid's receivers sit at in each soil's
    grid, as the batch snaps them (Sim.Column, internal/fdtd/fdtd.go:142-146:
    column round(r / h), half away from zero, clamped to [0, nr - 2]), without
    solving anything."""
    i = np.arange(RANGES[2], dtype=np.float64)
    grid = RANGES[0] + (RANGES[1] - RANGES[0]) * i / (RANGES[2] - 1)   # the bank's range grid, as bank.cpp forms it
    out = np.empty((len(soils), len(grid)))
    for i, s in enumerate(soils):
        nr = shots_for([s], seconds=1e-3)[0].Model.Grid()[1]
        c = np.clip(np.floor(grid / s.Spacing + 0.5), 0, nr - 2)
        out[i] = c * s.Spacing
    return out


def save_weights(path: str, net: "Network", norm_limit: float, meta: dict) -> None:
    """The network's FP32 master weights, its magnitude-guard limit and a JSON
    record of how they were trained and how they score."""
    import json
    import os

    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params=net.Params(), norm_limit=np.float64(norm_limit),
                        meta=np.array(json.dumps(meta)))


def load_weights(path: str, max_batch: int = 8192):
    """(Network with the saved weights, norm_limit, meta dict)."""
    import json

    z = np.load(path, allow_pickle=False)
    net = Network(max_batch=max_batch, seed=0)
    net.SetParams(z["params"])
    return net, float(z["norm_limit"]), json.loads(str(z["meta"]))


def solve(soils: Sequence, seconds: float = SECONDS, chunk: int = 2048,
          progress: Optional[Callable[[int, int, float], None]] = None):
    """Batched FDTD banks for soils, `chunk` soils per batch: (spectra
    [S][R][512] complex64, ranges [S][R], wall seconds, summed batch
    statistics).  progress(done, total, seconds) after each batch.  The
    batch size changes no result; 2,048 soils of this family hold about 21 GB
    of state, and one batch of 1,311 ran 3 % faster than batches of 256."""
    t0 = time.perf_counter()
    parts, snaps = [], []
    stats = {"kernel_ms": 0.0, "step_launches": 0, "algorithmic_bytes": 0.0}
    for a in range(0, len(soils), chunk):
        spectra, snapped, res = fd.GenerateBanks(shots_for(soils[a:a + chunk], seconds), bank_spec())
        bad = [r for r in res if isinstance(r, fd.FdtdError)]
        if bad:
            raise bad[0]
        parts.append(spectra.astype(np.complex64))
        snaps.append(np.asarray(snapped))
        st = fd.BatchStats()
        for k in stats:
            stats[k] += st[k]
        if progress:
            progress(min(a + chunk, len(soils)), len(soils), time.perf_counter() - t0)
    wall = time.perf_counter() - t0
    spectra = np.concatenate(parts)
    return spectra, np.concatenate(snaps).reshape(spectra.shape[0], spectra.shape[1]), wall, stats


def solve_set(ss: SoilSet, seconds: float = SECONDS, progress=None) -> SoilSet:
    ss.spectra, ss.ranges, ss.fdtd_seconds, ss.stats = solve(ss.soils, seconds, progress=progress)
    return ss


QC_PEAK_OVER_MEDIAN = 30.0


def soil_qc(ss: SoilSet) -> np.ndarray:
    """[S] bool: soils whose every bank spectrum looks physical.  The late-time
    growth a C-PML can develop in a soil with a thick soft top layer reaches
    every receiver of that soil and shows as an in-band |T| (the normalized
    target) peaking far above its median; such a bank is not a fit target."""
    S_, R = ss.ranges.shape
    p = np.repeat(ss.params, R, axis=0)
    r = ss.ranges.reshape(-1)
    scale, delay = prior(p, r)
    t = normalize(ss.spectra.reshape(S_ * R, BINS), scale, delay)
    mag = np.abs(t[:, 0::2] + 1j * t[:, 1::2])[:, band_mask()]
    ok = mag.max(axis=1) <= QC_PEAK_OVER_MEDIAN * np.median(mag, axis=1)
    return ok.reshape(S_, R).all(axis=1)


def subset(ss: SoilSet, keep: np.ndarray) -> SoilSet:
    k = np.flatnonzero(keep)
    return SoilSet(ss.seed, ss.indices[k], ss.params[k], [ss.soils[i] for i in k],
                   None if ss.spectra is None else ss.spectra[k], None if ss.ranges is None else ss.ranges[k],
                   ss.fdtd_seconds, ss.stats)


def norm_bound(t: np.ndarray, factor: float = 4.0) -> float:
    """The service's magnitude guard for a network trained on targets t: a
    prediction whose normalized norm exceeds `factor` times the largest
    training target norm is outside what the network has learned."""
    return float(factor * np.max(np.linalg.norm(t, axis=1)))


def dataset(ss: SoilSet):
    """Flattened (x [n][15], targets [n][1024], scale, delay, G [n][512]) over
    every (soil, range) of a solved set; G band-limited to BAND_HZ."""
    S, R = ss.ranges.shape
    p = np.repeat(ss.params, R, axis=0)
    r = ss.ranges.reshape(-1)
    x = features(p, r)
    scale, delay = prior(p, r)
    G = band_limit(ss.spectra.reshape(S * R, BINS))
    return x, normalize(G, scale, delay), scale, delay, G


# ------------------------------------------------------------ the service


@dataclass
class GateReport:
    queries: int
    soils: int
    sampled_soils: List[int]
    failed_soils: List[int]
    sampled_errors: np.ndarray
    surrogate_seconds: float
    gate_seconds: float
    bound: float
    guarded_soils: List[int] = field(default_factory=list)
    escalated: bool = False               # the whole request was answered by the FDTD
    fdtd_soils: List[int] = field(default_factory=list)   # soils answered by their FDTD solution

    def vouch_bound(self, confidence: float = 0.95) -> float:
        """With no sampled soil failing, the largest fraction of failing
        soils in the request that the sample rules out at `confidence`:
        a request in which a fraction p of the soils fail passes k sampled
        soils unnoticed with probability (1 - p)^k."""
        k = len(self.sampled_soils)
        return 1.0 if k == 0 else 1.0 - (1.0 - confidence) ** (1.0 / k)


def serve(net: Network, params: np.ndarray, ranges: np.ndarray, soils: Sequence,
          gate_fraction: float = 0.01, bound: float = 0.02, seed: int = 0,
          solver: Optional[Callable] = None, predictions: Optional[np.ndarray] = None,
          norm_limit: Optional[float] = None, escalate: bool = True):
    """Answer every (soil, range) query of `soils` ([S] soils with their 10
    numbers `params` [S][10] and receiver ranges [S][R]).

    The network predicts all S*R spectra; then ceil(gate_fraction * S) soils,
    drawn with `seed`, are re-solved by the batched FDTD (`solver(soils)` ->
    (spectra [k][R][512], ...), default `solve`).  Every sampled soil with a
    query whose relative L2 against its FDTD solution (both limited to
    BAND_HZ) exceeds `bound` is answered by that solution.  With `escalate`
    (the default) one such soil is enough to show that the network is out of
    its bound on this request, so every soil of the request is then answered
    by its FDTD solution; otherwise the unsampled soils keep their
    predictions.  With `norm_limit` (see norm_bound), every soil with a
    prediction whose normalized norm exceeds it is re-solved and answered by
    the FDTD as well.  `predictions` ([S*R][512]) replaces the network's
    output, for tests of the gate itself.

    What the gate guarantees: a query answered by the network belongs to a
    request in which no sampled soil exceeded the bound; at 95 % confidence
    fewer than GateReport.vouch_bound() of that request's soils do.

    Returns (spectra [S][R][512] complex64, GateReport)."""
    S, R = ranges.shape
    p = np.repeat(np.asarray(params), R, axis=0)
    r = ranges.reshape(-1)
    t0 = time.perf_counter()
    if predictions is None:
        x = features(p, r)
        scale, delay = prior(p, r)
        pred = net.Predict(x, scale, delay)
    else:
        pred = np.asarray(predictions, dtype=np.complex64)
    t1 = time.perf_counter()
    out = pred.reshape(S, R, BINS).copy()
    k = min(S, int(math.ceil(gate_fraction * S)))
    rng = np.random.default_rng(seed)
    sampled = sorted(rng.choice(S, size=k, replace=False).tolist()) if k > 0 else []
    guarded: List[int] = []
    if norm_limit is not None:
        scale, _ = prior(p, r)
        nrm = (np.linalg.norm(pred.reshape(S * R, BINS), axis=1) / scale).reshape(S, R)
        guarded = sorted(set(np.flatnonzero(np.any(nrm > norm_limit, axis=1)).tolist()) - set(sampled))
    solve_ = solver or solve
    failed: List[int] = []
    fdtd: List[int] = []
    errs = np.zeros((0,))
    known = {}                     # soil -> its FDTD solution, band-limited
    resolve = sampled + guarded
    if resolve:
        truth = band_limit(solve_([soils[i] for i in resolve])[0])
        known = {s_: truth[j] for j, s_ in enumerate(resolve)}
        if sampled:
            kk = len(sampled)
            e = rel_l2(band_limit(out[sampled]).reshape(kk * R, BINS), truth[:kk].reshape(kk * R, BINS)).reshape(kk, R)
            errs = e.reshape(-1)
            failed = [s_ for j, s_ in enumerate(sampled) if np.any(e[j] > bound)]
        fdtd = sorted(set(failed) | set(guarded))
    escalated = bool(escalate and failed)
    if escalated:
        rest = [s_ for s_ in range(S) if s_ not in known]
        if rest:
            truth = band_limit(solve_([soils[i] for i in rest])[0])
            known.update({s_: truth[j] for j, s_ in enumerate(rest)})
        fdtd = list(range(S))
    for s_ in fdtd:
        out[s_] = known[s_]
    t2 = time.perf_counter()
    return out, GateReport(S * R, S, sampled, failed, errs, t1 - t0, t2 - t1, bound, guarded, escalated, fdtd)
