"""Python face of the frequency-wavenumber solver in libgeofdtd: the
reference's fk package API over the C ABI (gf_fk_* in include/geofdtd.h).

The names and argument meanings follow internal/fk of the reference:

    Medium                    fk.Medium (fk.go:35-53): Stack, RefFreq,
                              DefaultQ, Relax (None = Kjartansson constant Q)
    Medium.SurfaceResponse    fk.go:276-307
    Medium.StaticCoefficient  fk.go:347-353
    Medium.VerticalDisplacement       fk.go:378-431
    Medium.GridFor            fk.go:456-489
    Medium.VerticalDisplacementMulti  fk.go:500-550 (one frequency, or many
                              in one GPU launch)
    Medium.checkLayerPhase    fk.go:253-268
    Integration, Grid, MaxLayerPhase
    Synthesise                synthesise in cmd/geofdtd/main.go:168-200
    FillBank                  cmd/geobank's build loop (main.go:100-175)

Every solve runs on the GPU in float64; errors the Go code returns raise
fdtd.FdtdError with the same text.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

from . import fdtd
from .fdtd import FdtdError, Layer, SLS  # noqa: F401  (re-exported for callers)

MaxLayerPhase = 16.0


@dataclass
class Integration:
    """fk.Integration (fk.go:309-331): 0 selects the defaults."""

    Samples: int = 0
    KMaxFactor: float = 0.0

    def to_c(self) -> fdtd.GfIntegration:
        return fdtd.GfIntegration(int(self.Samples), float(self.KMaxFactor))


@dataclass
class Grid:
    """fk.Grid (fk.go:439-443)."""

    KMax: float
    Samples: int


def _freqs(f) -> np.ndarray:
    return np.ascontiguousarray(np.atleast_1d(np.asarray(f, dtype=np.float64)))


@dataclass
class Medium:
    """fk.Medium (fk.go:35-53)."""

    Stack: List[Layer] = field(default_factory=list)
    RefFreq: float = 0.0
    DefaultQ: float = 0.0
    Relax: Optional[SLS] = None

    def to_c(self):
        layers = (fdtd.GfLayer * max(1, len(self.Stack)))()
        for i, l in enumerate(self.Stack):
            layers[i] = fdtd.GfLayer(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs, l.Qp)
        relax = self.Relax or SLS()
        m = fdtd.GfMedium(layers, len(self.Stack), self.RefFreq, self.DefaultQ, 1 if self.Relax is not None else 0,
                          fdtd.GfSls(relax.TauSigma, relax.TauEps))
        return m, layers

    # -- host-side rules (no GPU) ------------------------------------------

    def checkLayerPhase(self, freq: float) -> None:
        lib = fdtd.load_library()
        m, keep = self.to_c()
        fdtd._check(lib.gf_fk_check_layer_phase(C.byref(m), freq))

    def GridFor(self, ranges: Sequence[float], freq: float, opt: Integration = Integration()) -> Grid:
        lib = fdtd.load_library()
        m, keep = self.to_c()
        r = np.ascontiguousarray(ranges, dtype=np.float64)
        g = fdtd.GfFkGrid()
        fdtd._check(lib.gf_fk_grid_for(C.byref(m), fdtd._dptr(r) if len(r) else None, len(r), freq, opt.to_c(),
                                       C.byref(g)))
        return Grid(g.kmax, g.samples)

    # -- GPU ------------------------------------------------------------------

    def SurfaceResponses(self, k, freq) -> np.ndarray:
        """SurfaceResponse at every (k[i], freq[i]) (broadcast), one launch."""
        lib = fdtd.load_library()
        m, keep = self.to_c()
        kk, ff = np.broadcast_arrays(np.asarray(k, dtype=np.float64), np.asarray(freq, dtype=np.float64))
        shape = kk.shape
        kk = np.ascontiguousarray(kk.ravel())
        ff = np.ascontiguousarray(ff.ravel())
        out = np.zeros((max(kk.size, 1), 2))
        fdtd._check(lib.gf_fk_surface_response(C.byref(m), fdtd._dptr(kk), fdtd._dptr(ff), kk.size,
                                               fdtd._dptr(out)))
        return (out[: kk.size, 0] + 1j * out[: kk.size, 1]).reshape(shape)

    def SurfaceResponse(self, k: float, freq: float) -> complex:
        return complex(self.SurfaceResponses(k, freq))

    def StaticCoefficient(self, freq: float, kFar: float) -> complex:
        lib = fdtd.load_library()
        m, keep = self.to_c()
        out = np.zeros(2)
        fdtd._check(lib.gf_fk_static_coefficient(C.byref(m), freq, kFar, fdtd._dptr(out)))
        return complex(out[0], out[1])

    def VerticalDisplacement(self, r: float, freq: float, opt: Integration = Integration()) -> complex:
        lib = fdtd.load_library()
        m, keep = self.to_c()
        out = np.zeros(2)
        fdtd._check(lib.gf_fk_vertical_displacement(C.byref(m), r, freq, opt.to_c(), fdtd._dptr(out)))
        return complex(out[0], out[1])

    def VerticalDisplacementMulti(self, ranges: Sequence[float], freq: Union[float, Sequence[float]],
                                  opt: Integration = Integration()) -> np.ndarray:
        """[len(ranges)] complex for one frequency; [len(freq)][len(ranges)]
        for a sequence of frequencies (all solved in one launch)."""
        lib = fdtd.load_library()
        m, keep = self.to_c()
        r = np.ascontiguousarray(ranges, dtype=np.float64)
        f = _freqs(freq)
        out = np.zeros((max(f.size, 1), max(r.size, 1), 2))
        fdtd._check(lib.gf_fk_vertical_displacement_multi(C.byref(m), fdtd._dptr(r) if r.size else None, r.size,
                                                          fdtd._dptr(f), f.size, opt.to_c(), fdtd._dptr(out)))
        z = out[: f.size, : r.size, 0] + 1j * out[: f.size, : r.size, 1]
        return z[0] if np.ndim(freq) == 0 else z


def Synthesise(stack: Sequence[Layer], sls: SLS, refFreq: float, ranges: Sequence[float], peak: float, f0: float,
               band: float, dt: float, t0: float, n: int, samples: int = 0) -> np.ndarray:
    """synthesise (cmd/geofdtd/main.go:168-200): the f-k vertical velocity
    at each range for a Ricker force, on the instants t0 + i dt; [ranges][n]."""
    lib = fdtd.load_library()
    med = Medium(Stack=list(stack), Relax=sls, RefFreq=refFreq)
    m, keep = med.to_c()
    r = np.ascontiguousarray(ranges, dtype=np.float64)
    out = np.zeros((max(r.size, 1), max(n, 1)))
    fdtd._check(lib.gf_fk_synthesise(C.byref(m), fdtd._dptr(r) if r.size else None, r.size, peak, f0, band, dt, t0,
                                     n, samples, fdtd._dptr(out)))
    return out[: r.size, :n]


def FillBank(medium: Medium, h: "fdtd.BankHeader", max_freq: float) -> np.ndarray:
    """cmd/geobank's build (main.go:100-175): the bank payload for header h by
    wavenumber integration, bins 1..max_freq, as complex [Count][Bins]
    (float32 values, as bank.Set stores them)."""
    lib = fdtd.load_library()
    m, keep = medium.to_c()
    ch, keep_h = h.to_c()
    out = np.zeros((h.Count, h.Bins, 2), dtype=np.float32)
    fdtd._check(lib.gf_fk_bank_fill(C.byref(m), C.byref(ch), max_freq, out.ctypes.data_as(C.POINTER(C.c_float))))
    return out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64)


def Stats() -> dict:
    """Counters of the last f-k integration on this thread."""
    lib = fdtd.load_library()
    s = fdtd.GfFkStats()
    fdtd._check(lib.gf_fk_stats(C.byref(s)))
    return {k: getattr(s, k) for k, _ in fdtd.GfFkStats._fields_}
