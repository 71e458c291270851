"""Python face of the device bank lookup in libgeofdtd: bank.Bank's
interpolation (internal/bank/interp.go of the reference) over the C ABI.

    Bank(header, data)        bank.New + Set of every value, then Bank.prepare
                              (format.go:167-194, interp.go:64-101), in HBM
    Bank.Response(r, bin)     interp.go:29-54
    Bank.Responses(r, bin)    the same at many points, one launch
    Bank.Spectra(ranges)      every bin at each range, one launch
    Bank.ImpulseResponses     sensing.FromBank (propagation.go:79-112) through
                              dsp.CausalImpulseResponseAt (dsp.go:232-260)
    RangeNyquist, CheckRangeSampling      interp.go:120-141

Errors the Go code returns raise fdtd.FdtdError with the same text.  Every
lookup runs on the GPU; there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import fdtd
from .fdtd import BankHeader  # noqa: F401  (re-exported for callers)


def RangeNyquist(slowest_velocity: float, max_freq: float) -> float:
    return fdtd.load_library().gf_range_nyquist(slowest_velocity, max_freq)


def CheckRangeSampling(h: "fdtd.BankHeader", max_freq: float) -> None:
    ch, keep = h.to_c()
    fdtd._check(fdtd.load_library().gf_bank_check_range_sampling(C.byref(ch), max_freq))


class Bank:
    """A bank resident in HBM with its interpolation tables."""

    def __init__(self, h: "fdtd.BankHeader", data: np.ndarray):
        lib = fdtd.load_library()
        d = np.ascontiguousarray(np.stack([np.real(data), np.imag(data)], axis=-1), dtype=np.float32)
        if d.shape != (h.Count, h.Bins, 2):
            raise fdtd.FdtdError(fdtd.GF_ERR_ARGUMENT, f"data is {data.shape}, header implies {(h.Count, h.Bins)}")
        ch, keep = h.to_c()
        handle = C.c_void_p()
        fdtd._check(lib.gf_bank_dev_open(C.byref(ch), d.ctypes.data_as(C.POINTER(C.c_float)), C.byref(handle)))
        self._h = handle
        self.header = h

    def close(self):
        if getattr(self, "_h", None):
            fdtd.load_library().gf_bank_dev_close(self._h)
            self._h = None

    __del__ = close

    @property
    def Bins(self) -> int:
        return self.header.Bins

    def Responses(self, r, bins) -> np.ndarray:
        rr, bb = np.broadcast_arrays(np.asarray(r, dtype=np.float64), np.asarray(bins, dtype=np.int32))
        shape = rr.shape
        rr = np.ascontiguousarray(rr.ravel())
        bb = np.ascontiguousarray(bb.ravel())
        out = np.zeros((max(rr.size, 1), 2))
        fdtd._check(fdtd.load_library().gf_bank_dev_response(
            self._h, fdtd._dptr(rr), bb.ctypes.data_as(C.POINTER(C.c_int32)), rr.size, fdtd._dptr(out)))
        return (out[: rr.size, 0] + 1j * out[: rr.size, 1]).reshape(shape)

    def Response(self, r: float, bin: int) -> complex:
        return complex(self.Responses(r, bin))

    def Spectra(self, ranges: Sequence[float]) -> np.ndarray:
        rr = np.ascontiguousarray(ranges, dtype=np.float64)
        out = np.zeros((max(rr.size, 1), self.Bins, 2))
        fdtd._check(fdtd.load_library().gf_bank_dev_spectra(self._h, fdtd._dptr(rr), rr.size, fdtd._dptr(out)))
        return out[: rr.size, :, 0] + 1j * out[: rr.size, :, 1]

    def ImpulseResponses(self, ranges: Sequence[float], length: int, fs: float, sensor=None) -> np.ndarray:
        """sensing.FromBank's vertical velocity response through
        dsp.CausalImpulseResponseAt at the bank's transform size, for each
        range: [len(ranges)][length].  sensor: complex per bin, or None."""
        rr = np.ascontiguousarray(ranges, dtype=np.float64)
        sen = None
        if sensor is not None:
            z = np.asarray(sensor, dtype=complex)
            sen = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1))
        out = np.zeros((max(rr.size, 1), max(int(length), 1)))
        fdtd._check(fdtd.load_library().gf_bank_dev_impulse(
            self._h, fdtd._dptr(rr), rr.size, int(length), fs, fdtd._dptr(sen) if sen is not None else None,
            fdtd._dptr(out)))
        return out[: rr.size, : int(length)]

    def Stats(self) -> dict:
        ms, b = C.c_double(), C.c_double()
        fdtd._check(fdtd.load_library().gf_bank_dev_stats(self._h, C.byref(ms), C.byref(b)))
        return {"kernel_ms": ms.value, "bytes": b.value}
