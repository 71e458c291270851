"""Python face of libgeofdtd: the reference's fdtd package API over the C ABI.

The names and argument meanings follow internal/fdtd of the reference
(Ryan-Hancock/Co-Simulation-of-Seismic-Sensors-and-ROS-for-Robotic-Human-Detection):

    Model, Layer, SLS        fdtd.Model (model.go:37-76), layer.Layer, visco.SLS
    New(model) -> Sim        fdtd.New (fdtd.go:45-97)
    Sim.Dt/Cells/Time/Drive/Column/Step/SurfaceVelocity/RadialVelocity/MaxSpeed
                             fdtd.go:118-360
    Run(Shot) -> Result      fdtd.Run (run.go:88-127)
    Ricker, RickerSpectrum   run.go:19-46
    Fit                      visco.Fit (visco.go:55-68)

Errors the Go code returns as `error` values raise FdtdError with the same
message.  Every call goes through libgeofdtd.so and the GPU; there is no CPU
fallback, and a missing library or device raises.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GEOFDTD_LIBRARY selects another build of the same library (A/B timing).
LIB_PATH = os.environ.get("GEOFDTD_LIBRARY") or os.path.join(_HERE, "lib", "libgeofdtd.so")

GF_OK, GF_ERR_INVALID, GF_ERR_DIVERGED, GF_ERR_DEVICE, GF_ERR_ARGUMENT = 0, 1, 2, 3, 4


class FdtdError(RuntimeError):
    """An error value from the solver; .status is the C ABI status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


# ---------------------------------------------------------------- ABI structs


class GfLayer(C.Structure):
    _fields_ = [
        ("thickness_m", C.c_double),
        ("vp_mps", C.c_double),
        ("vs_mps", C.c_double),
        ("density_kgm3", C.c_double),
        ("qs", C.c_double),
        ("qp", C.c_double),
    ]


class GfSls(C.Structure):
    _fields_ = [("tau_sigma", C.c_double), ("tau_eps", C.c_double)]


class GfModel(C.Structure):
    _fields_ = [
        ("stack", C.POINTER(GfLayer)),
        ("n_layers", C.c_int32),
        ("relax", GfSls),
        ("ref_freq_hz", C.c_double),
        ("max_range_m", C.c_double),
        ("depth_m", C.c_double),
        ("spacing_m", C.c_double),
        ("pml_cells", C.c_int32),
        ("pml_reflection", C.c_double),
        ("dominant_freq_hz", C.c_double),
        ("courant", C.c_double),
        ("gsls_mechanisms", C.c_int32),
        ("gsls_fmin_hz", C.c_double),
        ("gsls_fmax_hz", C.c_double),
    ]


class GfShot(C.Structure):
    _fields_ = [
        ("model", C.POINTER(GfModel)),
        ("force", C.POINTER(C.c_double)),
        ("n_force", C.c_int64),
        ("ranges", C.POINTER(C.c_double)),
        ("n_ranges", C.c_int32),
        ("steps", C.c_int32),
    ]


class GfResult(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("t0", C.c_double),
        ("nr", C.c_int32),
        ("nz", C.c_int32),
        ("peak", C.c_double),
        ("traces", C.POINTER(C.c_double)),
        ("snapped", C.POINTER(C.c_double)),
        ("status", C.c_int32),
        ("error", C.c_char * 244),
    ]


class GfBatchStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("step_launches", C.c_int64),
        ("algorithmic_bytes", C.c_double),
        ("tile_rows", C.c_int32),
        ("models", C.c_int32),
        ("waves", C.c_int32),
        ("steps_per_launch", C.c_int32),
        ("budget_mb", C.c_double),
        ("pass_bytes", C.c_double),
    ]


class GfBankHeader(C.Structure):
    _fields_ = [
        ("solver", C.c_char_p),
        ("config_hash", C.c_char_p),
        ("created", C.c_char_p),
        ("notes", C.c_char_p),
        ("medium", C.POINTER(GfLayer)),
        ("n_medium", C.c_int32),
        ("sample_rate_hz", C.c_double),
        ("samples", C.c_int32),
        ("r_min_m", C.c_double),
        ("r_max_m", C.c_double),
        ("count", C.c_int32),
        ("component", C.c_char_p),
        ("units", C.c_char_p),
    ]


class GfBankSpec(C.Structure):
    _fields_ = [
        ("sample_rate_hz", C.c_double),
        ("samples", C.c_int32),
        ("max_freq_hz", C.c_double),
        ("r_min_m", C.c_double),
        ("r_max_m", C.c_double),
        ("count", C.c_int32),
    ]


class GfSoil(C.Structure):
    _fields_ = [
        ("layers", GfLayer * 4),
        ("n_layers", C.c_int32),
        ("spacing_m", C.c_double),
        ("max_range_m", C.c_double),
        ("depth_m", C.c_double),
        ("pml_cells", C.c_int32),
    ]


class GfSurConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("seed", C.c_uint64), ("band_lo_hz", C.c_double),
                ("band_hi_hz", C.c_double), ("polar", C.c_int32)]


class GfSurLayer(C.Structure):
    _fields_ = [
        ("k_in", C.c_int32),
        ("n_out", C.c_int32),
        ("gelu", C.c_int32),
        ("rows_per_query", C.c_int32),
        ("bias_channels", C.c_int32),
        ("w_offset", C.c_int64),
        ("b_offset", C.c_int64),
    ]


class GfSurTrainOpts(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_double),
        ("lr_min", C.c_double),
        ("beta1", C.c_double),
        ("beta2", C.c_double),
        ("eps", C.c_double),
        ("seed", C.c_uint64),
    ]


class GfSurStats(C.Structure):
    _fields_ = [
        ("forward_ms", C.c_double),
        ("forward_queries", C.c_int64),
        ("train_ms", C.c_double),
        ("train_steps", C.c_int64),
        ("train_queries", C.c_int64),
        ("flops_per_query", C.c_double),
    ]


class GfHalfspace(C.Structure):
    """gf_halfspace (soil.HalfSpace)."""

    _fields_ = [("vp_mps", C.c_double), ("vs_mps", C.c_double), ("density_kgm3", C.c_double), ("qs", C.c_double),
                ("qp", C.c_double)]


class GfMedium(C.Structure):
    _fields_ = [
        ("stack", C.POINTER(GfLayer)),
        ("n_layers", C.c_int32),
        ("ref_freq_hz", C.c_double),
        ("default_q", C.c_double),
        ("has_relax", C.c_int32),
        ("relax", GfSls),
    ]


class GfIntegration(C.Structure):
    _fields_ = [("samples", C.c_int32), ("kmax_factor", C.c_double)]


class GfFkGrid(C.Structure):
    _fields_ = [("kmax", C.c_double), ("samples", C.c_int32)]


class GfFkStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("prologue_ms", C.c_double),
        ("solve_ms", C.c_double),
        ("hankel_ms", C.c_double),
        ("fold_ms", C.c_double),
        ("samples", C.c_int64),
        ("weights", C.c_int64),
        ("frequencies", C.c_int32),
        ("launches", C.c_int32),
    ]


# Every symbol include/geofdtd.h declares, with its ctypes signature.
_D, _I, _I32, _I64 = C.c_double, C.c_int, C.c_int32, C.c_int64
_PD, _PI32, _PI64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
SIGNATURES = {
    "gf_last_error": (C.c_char_p, []),
    "gf_abi_version": (_I, []),
    "gf_visco_fit": (_I, [_D, _D, C.POINTER(GfSls)]),
    "gf_spacing_for": (_I, [C.POINTER(GfModel), _D, _D, _PD]),
    "gf_ricker": (_I, [_D, _D, _D, _I32, _PD]),
    "gf_ricker_spectrum": (_I, [_D, _D, _D, _PD, _PD]),
    "gf_model_grid": (_I, [C.POINTER(GfModel), _PD, _PI32, _PI32]),
    "gf_stack_validate": (_I, [C.POINTER(GfLayer), _I32]),
    "gf_sim_reset": (_I, [C.c_void_p]),
    "gf_sim_create": (_I, [C.POINTER(GfModel), C.POINTER(C.c_void_p)]),
    "gf_sim_destroy": (None, [C.c_void_p]),
    "gf_sim_info": (_I, [C.c_void_p, _PD, _PI32, _PI32]),
    "gf_sim_time": (_I, [C.c_void_p, _PD]),
    "gf_sim_drive": (_I, [C.c_void_p, _PD, _I64]),
    "gf_sim_column": (_I, [C.c_void_p, _D, _PI32, _PD]),
    "gf_sim_step": (_I, [C.c_void_p]),
    "gf_sim_step_n": (_I, [C.c_void_p, _I32, _PI32, _I32, _PD]),
    "gf_sim_surface_velocity": (_I, [C.c_void_p, _I32, _PD]),
    "gf_sim_radial_velocity": (_I, [C.c_void_p, _I32, _PD]),
    "gf_sim_max_speed": (_I, [C.c_void_p, _PD]),
    "gf_run": (_I, [C.POINTER(GfShot), C.POINTER(GfResult)]),
    "gf_run_batch": (_I, [C.POINTER(GfShot), _I32, C.POINTER(GfResult), _PI32]),
    "gf_sim_set_receivers": (_I, [C.c_void_p, _PI32, _I32, _I32]),
    "gf_sim_advance": (_I, [C.c_void_p, _I32, _I32]),
    "gf_sim_sync_timing": (_I, [C.c_void_p, _PD, _PI64]),
    "gf_sim_read_traces": (_I, [C.c_void_p, _PD, _PI32]),
    "gf_sim_step_bytes": (_I, [C.c_void_p, _PD]),
    "gf_sim_autotune": (_I, [C.c_void_p, _I32, _PI32]),
    "gf_sim_set_tile_rows": (_I, [C.c_void_p, _I32]),
    "gf_sim_set_launch": (_I, [C.c_void_p, _I32, _I32]),
    "gf_sim_launch_shape": (_I, [C.c_void_p, _PI32, _PI32, _PI32]),
    "gf_bank_write": (_I, [C.c_char_p, C.POINTER(GfBankHeader), C.POINTER(C.c_float)]),
    "gf_bank_header_json": (_I, [C.POINTER(GfBankHeader), C.c_char_p, _I32]),
    "gf_bank_generate": (_I, [C.POINTER(GfShot), _I32, C.POINTER(GfBankSpec), C.POINTER(C.c_float), _PD,
                              C.POINTER(GfResult)]),
    "gf_soil_sample": (_I, [C.c_uint64, C.c_uint32, C.POINTER(GfSoil)]),
    "gf_batch_stats": (_I, [C.POINTER(GfBatchStats)]),
    "gf_batch_create": (_I, [C.POINTER(GfShot), _I32, C.POINTER(GfBankSpec), C.POINTER(C.c_void_p)]),
    "gf_batch_advance": (_I, [C.c_void_p, _I32, _I32]),
    "gf_batch_reset": (_I, [C.c_void_p]),
    "gf_batch_results": (_I, [C.c_void_p, C.POINTER(GfResult), _PI32]),
    "gf_batch_spectra": (_I, [C.c_void_p, C.POINTER(C.c_float), _PD, C.POINTER(GfResult)]),
    "gf_batch_destroy": (None, [C.c_void_p]),
    "gf_spectrum": (_I, [C.POINTER(C.c_float), _I32, _I32, _D, _D, _PD, _I64, _PD, _I32, _PD]),
    "gf_philox4x32": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "gf_soil_sample_layered": (_I, [C.c_uint64, C.c_uint32, C.POINTER(GfSoil), _PD]),
    "gf_sur_create": (_I, [C.POINTER(GfSurConfig), C.POINTER(C.c_void_p)]),
    "gf_sur_destroy": (None, [C.c_void_p]),
    "gf_sur_layers": (_I, [C.c_void_p, _PI32]),
    "gf_sur_layer": (_I, [C.c_void_p, _I32, C.POINTER(GfSurLayer)]),
    "gf_sur_param_count": (_I, [C.c_void_p, _PI64]),
    "gf_sur_get_params": (_I, [C.c_void_p, C.POINTER(C.c_float)]),
    "gf_sur_set_params": (_I, [C.c_void_p, C.POINTER(C.c_float)]),
    "gf_sur_forward": (_I, [C.c_void_p, C.POINTER(C.c_float), _I32, C.POINTER(C.c_float)]),
    "gf_sur_predict": (_I, [C.c_void_p, C.POINTER(C.c_float), _PD, _PD, _D, _I32, C.POINTER(C.c_float)]),
    "gf_sur_loss_grad": (_I, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), _I32,
                              _PD, C.POINTER(C.c_float)]),
    "gf_sur_train": (_I, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), _I64,
                          C.POINTER(GfSurTrainOpts), C.POINTER(C.c_float)]),
    "gf_sur_stats": (_I, [C.c_void_p, C.POINTER(GfSurStats)]),
    "gf_gemm_bf16": (_I, [_I32, _I32, _I32, C.POINTER(C.c_uint16), _I64, _I32, C.POINTER(C.c_uint16), _I64, _I32,
                          _I32, C.POINTER(C.c_float)]),
    "gf_gemm_bf16_noalloc": (_I, [_I32, _I32, _I32, C.c_void_p, _I64, _I32, C.c_void_p, _I64, _I32, _I32, C.c_void_p,
                                  C.c_void_p]),
    "gf_fk_surface_response": (_I, [C.POINTER(GfMedium), _PD, _PD, _I64, _PD]),
    "gf_fk_static_coefficient": (_I, [C.POINTER(GfMedium), _D, _D, _PD]),
    "gf_fk_grid_for": (_I, [C.POINTER(GfMedium), _PD, _I32, _D, GfIntegration, C.POINTER(GfFkGrid)]),
    "gf_fk_check_layer_phase": (_I, [C.POINTER(GfMedium), _D]),
    "gf_fk_vertical_displacement": (_I, [C.POINTER(GfMedium), _D, _D, GfIntegration, _PD]),
    "gf_fk_vertical_displacement_multi": (_I, [C.POINTER(GfMedium), _PD, _I32, _PD, _I32, GfIntegration, _PD]),
    "gf_fk_synthesise": (_I, [C.POINTER(GfMedium), _PD, _I32, _D, _D, _D, _D, _D, _I32, _I32, _PD]),
    "gf_fk_stats": (_I, [C.POINTER(GfFkStats)]),
    "gf_fk_bank_fill": (_I, [C.POINTER(GfMedium), C.POINTER(GfBankHeader), _D, C.POINTER(C.c_float)]),
    "gf_bank_dev_open": (_I, [C.POINTER(GfBankHeader), C.POINTER(C.c_float), C.POINTER(C.c_void_p)]),
    "gf_bank_dev_close": (None, [C.c_void_p]),
    "gf_bank_dev_response": (_I, [C.c_void_p, _PD, _PI32, _I64, _PD]),
    "gf_bank_dev_spectra": (_I, [C.c_void_p, _PD, _I32, _PD]),
    "gf_bank_dev_stats": (_I, [C.c_void_p, _PD, _PD]),
    "gf_bank_dev_impulse": (_I, [C.c_void_p, _PD, _I32, _I32, _D, _PD, _PD]),
    "gf_range_nyquist": (_D, [_D, _D]),
    "gf_convolver_create": (_I, [_PD, _I32, _I32, _I32, C.POINTER(C.c_void_p)]),
    "gf_convolver_destroy": (None, [C.c_void_p]),
    "gf_convolver_info": (_I, [C.c_void_p, _PI32, _PI32, _PI32]),
    "gf_convolver_process": (_I, [C.c_void_p, _PD, _I32, _PD, _I32]),
    "gf_convolver_reset": (_I, [C.c_void_p]),
    "gf_convolver_process_all": (_I, [C.c_void_p, _PD, _I64, _PD]),
    "gf_convolver_stats": (_I, [C.c_void_p, _PD]),
    "gf_engine_create": (_I, [C.c_void_p, _I32, _PD, _I32, _D, _I32, _PD, _I32, C.POINTER(C.c_void_p)]),
    "gf_engine_destroy": (None, [C.c_void_p]),
    "gf_engine_warm": (_I, [C.c_void_p, _PD, _I32]),
    "gf_engine_admit": (_I, [C.c_void_p, C.c_char_p, _D, _D, _PD, _I32, _D]),
    "gf_engine_next": (_I, [C.c_void_p, C.POINTER(C.c_float), _I32]),
    "gf_engine_info": (_I, [C.c_void_p, _PI64, _PI32, _PI32, _PI32]),
    "gf_engine_stats": (_I, [C.c_void_p, _PD, _PD, _PI64, _PD, _PI64]),
    "gf_rayleigh_velocity": (_I, [C.POINTER(GfHalfspace), _PD]),
    "gf_engine_create_analytic": (_I, [C.POINTER(GfHalfspace), _D, _I32, _PD, _I32, _D, _I32, _PD, _I32,
                                       C.POINTER(C.c_void_p)]),
    "gf_engine_admit_xyz": (_I, [C.c_void_p, C.c_char_p, _D, _D, _PD, _PD, _PD, _I32, _D]),
    "gf_engine_transform": (_I, [C.c_void_p, _PI32]),
    "gf_bank_check_range_sampling": (_I, [C.POINTER(GfBankHeader), _D]),
}

_lib: Optional[C.CDLL] = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libgeofdtd.so (built by __graft_entry__.build()).  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FdtdError(GF_ERR_DEVICE, f"libgeofdtd.so is not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != GF_OK:
        msg = load_library().gf_last_error().decode()
        raise FdtdError(rc, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_PD)


# ------------------------------------------------------------ Go-shaped types


@dataclass
class Layer:
    """layer.Layer (internal/layer/layer.go:24-36); Qp is an extension (0 = Qs)."""

    Thickness: float = 0.0
    Vp: float = 0.0
    Vs: float = 0.0
    Density: float = 0.0
    Qs: float = 0.0
    Qp: float = 0.0


def ValidateStack(stack: Sequence[Layer]) -> None:
    """layer.Stack.Validate (internal/layer/layer.go:53-73): raises FdtdError
    with the reference's message for a physically impossible stack."""
    lib = load_library()
    layers = (GfLayer * max(1, len(stack)))()
    for i, l in enumerate(stack):
        layers[i] = GfLayer(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs, l.Qp)
    _check(lib.gf_stack_validate(layers, len(stack)))


@dataclass
class SLS:
    """visco.SLS (internal/visco/visco.go:42-47)."""

    TauSigma: float = 0.0
    TauEps: float = 0.0


@dataclass
class Model:
    """fdtd.Model (internal/fdtd/model.go:37-76).  Zero values take the
    reference's defaults.  GslsMechanisms > 0 selects the per-layer,
    multi-mechanism attenuation extension."""

    Stack: List[Layer] = field(default_factory=list)
    Relax: SLS = field(default_factory=SLS)
    RefFreq: float = 0.0
    MaxRange: float = 0.0
    Depth: float = 0.0
    Spacing: float = 0.0
    PMLCells: int = 0
    PMLReflection: float = 0.0
    DominantFreq: float = 0.0
    Courant: float = 0.0
    GslsMechanisms: int = 0
    GslsFmin: float = 0.0
    GslsFmax: float = 0.0

    def to_c(self):
        """(GfModel, keep-alive array of layers)."""
        layers = (GfLayer * max(1, len(self.Stack)))()
        for i, l in enumerate(self.Stack):
            layers[i] = GfLayer(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs, l.Qp)
        m = GfModel(
            C.cast(layers, C.POINTER(GfLayer)),
            len(self.Stack),
            GfSls(self.Relax.TauSigma, self.Relax.TauEps),
            self.RefFreq,
            self.MaxRange,
            self.Depth,
            self.Spacing,
            self.PMLCells,
            self.PMLReflection,
            self.DominantFreq,
            self.Courant,
            self.GslsMechanisms,
            self.GslsFmin,
            self.GslsFmax,
        )
        return m, layers

    def Grid(self):
        """(dt, nr, nz) that New would use, computed on the host only."""
        lib = load_library()
        m, keep = self.to_c()
        dt, nr, nz = C.c_double(), C.c_int32(), C.c_int32()
        _check(lib.gf_model_grid(C.byref(m), C.byref(dt), C.byref(nr), C.byref(nz)))
        return dt.value, nr.value, nz.value

    def SpacingFor(self, max_freq: float, cells_per_wavelength: float) -> float:
        """Model.SpacingFor (model.go:121-136)."""
        lib = load_library()
        m, keep = self.to_c()
        out = C.c_double()
        _check(lib.gf_spacing_for(C.byref(m), max_freq, cells_per_wavelength, C.byref(out)))
        return out.value


def Fit(q: float, f0: float) -> SLS:
    """visco.Fit (visco.go:55-68)."""
    lib = load_library()
    s = GfSls()
    _check(lib.gf_visco_fit(q, f0, C.byref(s)))
    return SLS(s.tau_sigma, s.tau_eps)


def Ricker(peak: float, f0: float, dt: float, n: int) -> np.ndarray:
    """fdtd.Ricker (run.go:19-31)."""
    lib = load_library()
    out = np.zeros(max(n, 1), dtype=np.float64)
    _check(lib.gf_ricker(peak, f0, dt, n, _dptr(out)))
    return out[:n]


def RickerSpectrum(peak: float, f0: float, f: float) -> complex:
    """fdtd.RickerSpectrum (run.go:41-46)."""
    lib = load_library()
    re, im = C.c_double(), C.c_double()
    _check(lib.gf_ricker_spectrum(peak, f0, f, C.byref(re), C.byref(im)))
    return complex(re.value, im.value)


class Sim:
    """*fdtd.Sim (fdtd.go:20-42) on the GPU.  Use New(model)."""

    def __init__(self, handle: C.c_void_p, model: Model):
        self._h = handle
        self._model = model

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load_library().gf_sim_destroy(h)
            self._h = None

    def close(self) -> None:
        self.__del__()

    def Dt(self) -> float:
        dt = C.c_double()
        _check(load_library().gf_sim_info(self._h, C.byref(dt), None, None))
        return dt.value

    def Cells(self):
        nr, nz = C.c_int32(), C.c_int32()
        _check(load_library().gf_sim_info(self._h, None, C.byref(nr), C.byref(nz)))
        return nr.value, nz.value

    def Time(self) -> float:
        t = C.c_double()
        _check(load_library().gf_sim_time(self._h, C.byref(t)))
        return t.value

    def Drive(self, force: Sequence[float]) -> None:
        f = np.ascontiguousarray(force, dtype=np.float64)
        _check(load_library().gf_sim_drive(self._h, _dptr(f), len(f)))

    def Column(self, r: float):
        c, snapped = C.c_int32(), C.c_double()
        _check(load_library().gf_sim_column(self._h, r, C.byref(c), C.byref(snapped)))
        return c.value, snapped.value

    def Step(self) -> None:
        _check(load_library().gf_sim_step(self._h))

    def StepN(self, n: int, cols: Sequence[int] = ()) -> np.ndarray:
        """n Steps, recording SurfaceVelocity at cols after each: [len(cols), n]."""
        cols_a = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros((len(cols_a), max(n, 0)), dtype=np.float64)
        _check(
            load_library().gf_sim_step_n(
                self._h, n, cols_a.ctypes.data_as(_PI32), len(cols_a), _dptr(out) if len(cols_a) else None
            )
        )
        return out

    def SurfaceVelocity(self, i: int) -> float:
        v = C.c_double()
        _check(load_library().gf_sim_surface_velocity(self._h, i, C.byref(v)))
        return v.value

    def RadialVelocity(self, i: int) -> float:
        v = C.c_double()
        _check(load_library().gf_sim_radial_velocity(self._h, i, C.byref(v)))
        return v.value

    def MaxSpeed(self) -> float:
        v = C.c_double()
        _check(load_library().gf_sim_max_speed(self._h, C.byref(v)))
        return v.value

    def Reset(self) -> None:
        """Back to the state New leaves (step 0, zero fields); keeps set-up."""
        _check(load_library().gf_sim_reset(self._h))

    # device-resident benchmarking
    def SetReceivers(self, cols: Sequence[int], capacity_steps: int) -> None:
        c = np.ascontiguousarray(cols, dtype=np.int32)
        _check(load_library().gf_sim_set_receivers(self._h, c.ctypes.data_as(_PI32), len(c), capacity_steps))

    def Advance(self, n: int, use_graph: bool = True) -> None:
        _check(load_library().gf_sim_advance(self._h, n, 1 if use_graph else 0))

    def SyncTiming(self):
        ms, launches = C.c_double(), C.c_int64()
        _check(load_library().gf_sim_sync_timing(self._h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def ReadTraces(self, ncols: int) -> np.ndarray:
        n = C.c_int32()
        _check(load_library().gf_sim_read_traces(self._h, None, C.byref(n)))
        out = np.zeros((ncols, n.value), dtype=np.float64)
        _check(load_library().gf_sim_read_traces(self._h, _dptr(out), C.byref(n)))
        return out

    def Autotune(self, trial_steps: int = 10) -> int:
        """Pick the tile height by timing; resets to step 0.  Returns it."""
        th = C.c_int32()
        _check(load_library().gf_sim_autotune(self._h, trial_steps, C.byref(th)))
        return th.value

    def SetTileRows(self, rows: int) -> None:
        _check(load_library().gf_sim_set_tile_rows(self._h, rows))

    def SetLaunch(self, steps_per_launch: int, rows: int) -> None:
        """Time steps per kernel launch (1, or 2: the temporally blocked
        kernel) and its tile height."""
        _check(load_library().gf_sim_set_launch(self._h, steps_per_launch, rows))

    def LaunchShape(self) -> dict:
        sp, th, w = C.c_int32(), C.c_int32(), C.c_int32()
        _check(load_library().gf_sim_launch_shape(self._h, C.byref(sp), C.byref(th), C.byref(w)))
        return {"steps_per_launch": sp.value, "tile_rows": th.value, "waves": w.value}

    def StepBytes(self) -> float:
        b = C.c_double()
        _check(load_library().gf_sim_step_bytes(self._h, C.byref(b)))
        return b.value


def New(m: Model) -> Sim:
    """fdtd.New (fdtd.go:45-97)."""
    lib = load_library()
    cm, keep = m.to_c()
    h = C.c_void_p()
    _check(lib.gf_sim_create(C.byref(cm), C.byref(h)))
    return Sim(h, m)


@dataclass
class Shot:
    """fdtd.Shot (run.go:49-61)."""

    Model: Model
    Force: Sequence[float]
    Ranges: Sequence[float]
    Steps: int
    Progress: Optional[Callable[[int, int], None]] = None


@dataclass
class Trace:
    """fdtd.Trace (run.go:63-70)."""

    Range: float
    Vertical: np.ndarray


@dataclass
class Result:
    """fdtd.Result (run.go:72-86)."""

    Dt: float
    T0: float
    Traces: List[Trace]
    Cells: tuple
    Peak: float


def BatchStats() -> dict:
    """Timing of this thread's last RunBatch / GenerateBanks."""
    s = GfBatchStats()
    _check(load_library().gf_batch_stats(C.byref(s)))
    return {"kernel_ms": s.kernel_ms, "step_launches": s.step_launches, "algorithmic_bytes": s.algorithmic_bytes,
            "tile_rows": s.tile_rows, "models": s.models, "waves": s.waves, "steps_per_launch": s.steps_per_launch,
            "budget_mb": s.budget_mb, "pass_bytes": s.pass_bytes}


def _shot_c(sh: Shot):
    cm, keep_l = sh.Model.to_c()
    force = np.ascontiguousarray(sh.Force, dtype=np.float64)
    ranges = np.ascontiguousarray(sh.Ranges, dtype=np.float64)
    c = GfShot(
        C.pointer(cm),
        _dptr(force) if len(force) else None,
        len(force),
        _dptr(ranges) if len(ranges) else None,
        len(ranges),
        sh.Steps,
    )
    return c, (cm, keep_l, force, ranges)


def RunBatch(shots: Sequence[Shot]) -> List[object]:
    """Many shots advanced together on the GPU; each entry is a Result or the
    FdtdError that shot's fdtd.Run would have returned."""
    lib = load_library()
    n = len(shots)
    cshots = (GfShot * n)()
    cres = (GfResult * n)()
    keep, outs = [], []
    for i, sh in enumerate(shots):
        c, k = _shot_c(sh)
        cshots[i] = c
        keep.append(k)
        tr = np.zeros((max(len(sh.Ranges), 1), max(sh.Steps, 1)), dtype=np.float64)
        sn = np.zeros(max(len(sh.Ranges), 1), dtype=np.float64)
        outs.append((tr, sn))
        cres[i].traces = _dptr(tr)
        cres[i].snapped = _dptr(sn)
    statuses = np.zeros(n, dtype=np.int32)
    rc = lib.gf_run_batch(cshots, n, cres, statuses.ctypes.data_as(_PI32))
    if rc != GF_OK and not statuses.any():
        _check(rc)
    results: List[object] = []
    for i, sh in enumerate(shots):
        if statuses[i] != GF_OK:
            results.append(FdtdError(int(statuses[i]), cres[i].error.decode()))
            continue
        tr, sn = outs[i]
        r = cres[i]
        traces = [Trace(float(sn[k]), tr[k, : sh.Steps].copy()) for k in range(len(sh.Ranges))]
        results.append(Result(r.dt, r.t0, traces, (r.nr, r.nz), r.peak))
    return results


def Run(sh: Shot) -> Result:
    """fdtd.Run (run.go:88-127).  Progress, when set, is called every 500
    steps as the reference does; the steps run in blocks between calls."""
    if sh.Progress is None:
        r = RunBatch([sh])[0]
        if isinstance(r, FdtdError):
            raise r
        return r
    return _run_with_progress(sh)


def _run_with_progress(sh: Shot) -> Result:
    """Run's loop (run.go:104-126) driven from the host: steps run on the GPU
    in blocks that end where the reference reports progress (every 500 steps)
    or probes for divergence (every 256)."""
    s = New(sh.Model)
    try:
        if sh.Steps <= 0:
            raise FdtdError(GF_ERR_INVALID, "fdtd: a shot needs a positive number of steps")
        s.Drive(sh.Force)
        cols, snapped = [], []
        for r in sh.Ranges:
            c, sn = s.Column(r)
            if sn >= sh.Model.MaxRange:
                raise FdtdError(
                    GF_ERR_INVALID,
                    f"fdtd: receiver at {_go_g(r)} m is outside the interior ({_go_g(sh.Model.MaxRange)} m)",
                )
            cols.append(c)
            snapped.append(sn)
        out = np.zeros((len(cols), sh.Steps), dtype=np.float64)
        peak = 0.0
        n = 0
        while n < sh.Steps:
            # next step index that triggers an event
            nxt = n
            while nxt < sh.Steps - 1 and nxt % 500 != 0 and nxt % 256 != 0:
                nxt += 1
            block = nxt - n + 1
            out[:, n : nxt + 1] = s.StepN(block, cols)
            if nxt % 500 == 0:
                sh.Progress(nxt, sh.Steps)
            if nxt % 256 == 0:
                p = s.MaxSpeed()
                peak = p if (np.isnan(p) or p > peak) else peak
                if np.isnan(p) or np.isinf(p):
                    raise FdtdError(GF_ERR_DIVERGED, f"fdtd: diverged at step {nxt} of {sh.Steps}")
            n = nxt + 1
        dt = s.Dt()
        traces = [Trace(snapped[k], out[k]) for k in range(len(cols))]
        return Result(dt, 0.5 * dt, traces, s.Cells(), peak)
    finally:
        s.close()


def _go_g(x: float) -> str:
    """Go's %g: shortest round-trip digits, exponent form only for a decimal
    exponent below -4 or at least 6."""
    if x == 0:
        return "-0" if str(x).startswith("-") else "0"
    p = 1
    while p < 17 and float("%.*e" % (p - 1, x)) != x:
        p += 1
    t = "%.*e" % (p - 1, x)
    e = int(t.split("e")[1])
    if -4 <= e < 6:
        return "%.*f" % (max(p - 1 - e, 0), x)
    return t


# ---------------------------------------------------------------- banks


@dataclass
class BankHeader:
    """bank.Header (internal/bank/format.go:115-131)."""

    Solver: str
    Medium: List[Layer]
    SampleRateHz: float
    Samples: int
    MinM: float
    MaxM: float
    Count: int
    Component: str = "vertical surface displacement per unit vertical surface point force"
    Units: str = "m/N"
    ConfigHash: str = ""
    Created: str = ""
    Notes: str = ""

    def to_c(self):
        layers = (GfLayer * max(1, len(self.Medium)))()
        for i, l in enumerate(self.Medium):
            layers[i] = GfLayer(l.Thickness, l.Vp, l.Vs, l.Density, l.Qs, l.Qp)
        h = GfBankHeader(
            self.Solver.encode(), self.ConfigHash.encode(), self.Created.encode(), self.Notes.encode(),
            C.cast(layers, C.POINTER(GfLayer)) if self.Medium else None, len(self.Medium),
            self.SampleRateHz, self.Samples, self.MinM, self.MaxM, self.Count,
            self.Component.encode(), self.Units.encode(),
        )
        return h, layers

    @property
    def Bins(self) -> int:
        return self.Samples // 2 + 1


def WriteBank(path: str, h: BankHeader, data: np.ndarray) -> None:
    """bank.Write: data is complex [Count, Samples/2+1] (stored as float32)."""
    d = np.ascontiguousarray(np.stack([np.real(data), np.imag(data)], axis=-1), dtype=np.float32)
    if d.shape != (h.Count, h.Bins, 2):
        raise FdtdError(GF_ERR_ARGUMENT, f"data is {data.shape}, header implies {(h.Count, h.Bins)}")
    ch, keep = h.to_c()
    _check(load_library().gf_bank_write(path.encode(), C.byref(ch), d.ctypes.data_as(C.POINTER(C.c_float))))


def BankHeaderJSON(h: BankHeader) -> str:
    ch, keep = h.to_c()
    buf = C.create_string_buffer(1 << 16)
    _check(load_library().gf_bank_header_json(C.byref(ch), buf, len(buf)))
    return buf.value.decode()


@dataclass
class BankSpec:
    SampleRateHz: float = 500.0
    Samples: int = 1024
    MaxFreqHz: float = 100.0
    MinM: float = 1.0
    MaxM: float = 50.0
    Count: int = 50


def GenerateBanks(shots: Sequence[Shot], spec: BankSpec):
    """FDTD-filled banks: (spectra complex [n, Count, Samples/2+1], snapped
    ranges [n, Count], per-shot Result-or-FdtdError list)."""
    lib = load_library()
    n = len(shots)
    cshots = (GfShot * n)()
    keep = []
    for i, sh in enumerate(shots):
        c, k = _shot_c(Shot(sh.Model, sh.Force, [], sh.Steps))
        cshots[i] = c
        keep.append(k)
    bins = spec.Samples // 2 + 1
    out = np.zeros((n, spec.Count, bins, 2), dtype=np.float32)
    snapped = np.zeros((n, spec.Count), dtype=np.float64)
    cres = (GfResult * n)()
    cs = GfBankSpec(spec.SampleRateHz, spec.Samples, spec.MaxFreqHz, spec.MinM, spec.MaxM, spec.Count)
    rc = lib.gf_bank_generate(cshots, n, C.byref(cs), out.ctypes.data_as(C.POINTER(C.c_float)), _dptr(snapped), cres)
    status = [cres[i].status for i in range(n)]
    if rc != GF_OK and not any(status):
        _check(rc)
    results: List[object] = []
    for i in range(n):
        if status[i] != GF_OK:
            results.append(FdtdError(status[i], cres[i].error.decode()))
        else:
            results.append(Result(cres[i].dt, cres[i].t0, [], (cres[i].nr, cres[i].nz), cres[i].peak))
    return out[..., 0] + 1j * out[..., 1], snapped, results


class Batch:
    """Many shots as one handle, run in pieces: the bank generator's job loop
    (cmd/geobank/main.go:128-167) split into set-up (constructor: every shot
    through fdtd.New's checks, launches planned and tuned), Advance(num, den)
    (every chunk of models to that fraction of its longest run), Reset() and,
    once the batch has run to its end, Results() (what RunBatch returns) or
    Spectra() (what GenerateBanks returns, when built with a BankSpec)."""

    def __init__(self, shots: Sequence[Shot], spec: Optional["BankSpec"] = None):
        lib = load_library()
        self._shots = list(shots)
        self._spec = spec
        n = len(self._shots)
        cshots = (GfShot * n)()
        self._keep = []
        for i, sh in enumerate(self._shots):
            c, k = _shot_c(Shot(sh.Model, sh.Force, [] if spec else sh.Ranges, sh.Steps))
            cshots[i] = c
            self._keep.append(k)
        self._cs = None if spec is None else GfBankSpec(spec.SampleRateHz, spec.Samples, spec.MaxFreqHz, spec.MinM,
                                                        spec.MaxM, spec.Count)
        h = C.c_void_p()
        _check(lib.gf_batch_create(cshots, n, C.byref(self._cs) if self._cs else None, C.byref(h)))
        self._h = h

    def Advance(self, num: int = 1, den: int = 1) -> None:
        _check(load_library().gf_batch_advance(self._h, num, den))

    def Reset(self) -> None:
        _check(load_library().gf_batch_reset(self._h))

    def Stats(self) -> dict:
        """Work and device time since the last Reset (the thread's last batch)."""
        return BatchStats()

    def _results(self, cres, status):
        out: List[object] = []
        for i in range(len(self._shots)):
            if status[i] != GF_OK:
                out.append(FdtdError(int(status[i]), cres[i].error.decode()))
            else:
                out.append(Result(cres[i].dt, cres[i].t0, [], (cres[i].nr, cres[i].nz), cres[i].peak))
        return out

    def Results(self) -> List[object]:
        n = len(self._shots)
        cres = (GfResult * n)()
        bufs = []
        for i, sh in enumerate(self._shots):
            tr = np.zeros((max(len(sh.Ranges), 1), max(sh.Steps, 1)), dtype=np.float64)
            sn = np.zeros(max(len(sh.Ranges), 1), dtype=np.float64)
            bufs.append((tr, sn))
            cres[i].traces = _dptr(tr)
            cres[i].snapped = _dptr(sn)
        statuses = np.zeros(n, dtype=np.int32)
        rc = load_library().gf_batch_results(self._h, cres, statuses.ctypes.data_as(_PI32))
        if rc != GF_OK and not statuses.any():
            _check(rc)
        out = self._results(cres, statuses)
        for i, sh in enumerate(self._shots):
            if isinstance(out[i], Result):
                tr, sn = bufs[i]
                out[i].Traces = [Trace(float(sn[k]), tr[k, : sh.Steps].copy()) for k in range(len(sh.Ranges))]
        return out

    def Spectra(self):
        spec = self._spec
        if spec is None:
            raise FdtdError(GF_ERR_ARGUMENT, "geofdtd: the batch was created without a bank spec")
        n = len(self._shots)
        bins = spec.Samples // 2 + 1
        out = np.zeros((n, spec.Count, bins, 2), dtype=np.float32)
        snapped = np.zeros((n, spec.Count), dtype=np.float64)
        cres = (GfResult * n)()
        rc = load_library().gf_batch_spectra(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), _dptr(snapped), cres)
        status = [cres[i].status for i in range(n)]
        if rc != GF_OK and not any(status):
            _check(rc)
        return out[..., 0] + 1j * out[..., 1], snapped, self._results(cres, status)

    def close(self) -> None:
        if self._h:
            load_library().gf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def Spectrum(traces: np.ndarray, dt: float, t0: float, force: Sequence[float], freqs: Sequence[float]) -> np.ndarray:
    """G(f) = V(f) / (i w W(f)) of float32 traces [rows, n] on the device."""
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    if tr.ndim == 1:
        tr = tr[None, :]
    f = np.ascontiguousarray(force, dtype=np.float64)
    fq = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros((tr.shape[0], len(fq), 2), dtype=np.float64)
    _check(load_library().gf_spectrum(tr.ctypes.data_as(C.POINTER(C.c_float)), tr.shape[0], tr.shape[1], dt, t0,
                                      _dptr(f), len(f), _dptr(fq), len(fq), _dptr(out)))
    return out[..., 0] + 1j * out[..., 1]


@dataclass
class Soil:
    Stack: List[Layer]
    Spacing: float
    MaxRange: float
    Depth: float
    PMLCells: int


def SampleSoil(seed: int, index: int) -> Soil:
    """Seeded soil model `index` (Philox4x32-10 stream keyed by seed)."""
    s = GfSoil()
    _check(load_library().gf_soil_sample(seed, index, C.byref(s)))
    stack = [Layer(l.thickness_m, l.vp_mps, l.vs_mps, l.density_kgm3, l.qs, l.qp) for l in s.layers[: s.n_layers]]
    return Soil(stack, s.spacing_m, s.max_range_m, s.depth_m, s.pml_cells)


def SampleLayeredSoil(seed: int, index: int):
    """Soil model `index` of the layered family the surrogate is trained on,
    and the 10 numbers that define it {Vs1, Vs2, Vs3, Qs1, Qs2, Qs3, h1, h2,
    Vp/Vs, rho0}."""
    s = GfSoil()
    p = (C.c_double * 10)()
    _check(load_library().gf_soil_sample_layered(seed, index, C.byref(s), p))
    stack = [Layer(l.thickness_m, l.vp_mps, l.vs_mps, l.density_kgm3, l.qs, l.qp) for l in s.layers[: s.n_layers]]
    return Soil(stack, s.spacing_m, s.max_range_m, s.depth_m, s.pml_cells), list(p)


def Philox4x32(counter: Sequence[int], key: Sequence[int]) -> List[int]:
    c = (C.c_uint32 * 4)(*counter)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load_library().gf_philox4x32(c, k, o)
    return list(o)
