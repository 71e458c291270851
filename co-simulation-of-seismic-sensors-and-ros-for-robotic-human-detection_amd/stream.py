"""Python face of the streaming convolver in libgeofdtd: the reference's
stream package (internal/stream/stream.go) over the C ABI, for one or many
receivers at once.

    NewConvolver(impulse, chunk)   stream.go:72-107; impulse [len] or
                                   [receivers][len]
    Convolver.ChunkSize, Partitions, Process, Reset, ProcessAll
                                   stream.go:109-186

Errors raise fdtd.FdtdError with the Go text.  Every sum runs on the GPU.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import fdtd


class Convolver:
    def __init__(self, handle, n_recv, single):
        self._h = handle
        self.n_recv = n_recv
        self._single = single

    def close(self):
        if getattr(self, "_h", None):
            fdtd.load_library().gf_convolver_destroy(self._h)
            self._h = None

    __del__ = close

    def _info(self):
        c, p, r = C.c_int32(), C.c_int32(), C.c_int32()
        fdtd._check(fdtd.load_library().gf_convolver_info(self._h, C.byref(c), C.byref(p), C.byref(r)))
        return c.value, p.value, r.value

    def ChunkSize(self) -> int:
        return self._info()[0]

    def Partitions(self) -> int:
        return self._info()[1]

    def Process(self, x) -> np.ndarray:
        """One chunk for every receiver: x [chunk] or [receivers][chunk]."""
        a = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        out = np.zeros((a.shape[0], self.ChunkSize()))
        fdtd._check(fdtd.load_library().gf_convolver_process(self._h, fdtd._dptr(a), a.shape[1], fdtd._dptr(out),
                                                             out.shape[1]))
        return out[0] if self._single else out

    def Reset(self) -> None:
        fdtd._check(fdtd.load_library().gf_convolver_reset(self._h))

    def ProcessAll(self, x) -> np.ndarray:
        a = np.ascontiguousarray(np.atleast_2d(np.asarray(x, dtype=np.float64)))
        out = np.zeros_like(a)
        fdtd._check(fdtd.load_library().gf_convolver_process_all(self._h, fdtd._dptr(a), a.shape[1],
                                                                 fdtd._dptr(out)))
        return out[0] if self._single else out

    def KernelMs(self) -> float:
        ms = C.c_double()
        fdtd._check(fdtd.load_library().gf_convolver_stats(self._h, C.byref(ms)))
        return ms.value


def NewConvolver(impulse, chunk: int) -> Convolver:
    h = np.asarray(impulse, dtype=np.float64)
    single = h.ndim <= 1
    h = np.ascontiguousarray(np.atleast_2d(h)) if h.size else np.zeros((1, 0))
    handle = C.c_void_p()
    fdtd._check(fdtd.load_library().gf_convolver_create(fdtd._dptr(h) if h.size else None, h.shape[0], h.shape[1],
                                                        int(chunk), C.byref(handle)))
    return Convolver(handle, h.shape[0], single)
