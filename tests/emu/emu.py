"""ctypes access to the CPU emulation of the step kernel (test infrastructure)."""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def lib(asan=False):
    name = "libemu_asan.so" if asan else "libemu.so"
    path = os.path.join(HERE, "lib", name)
    subprocess.run(["make", "-s", "-C", HERE] + (["SANITIZE=1"] if asan else []), check=True)
    L = C.CDLL(path)
    L.emu_run.restype = C.c_int
    L.emu_last_error.restype = C.c_char_p
    return L


def run(model, force, cols, steps, tile_rows=8, asan=False, span=1):
    """(traces [ncols, steps] float32, probe bits) through the emulated kernel
    of `span` time steps per launch."""
    L = lib(asan)
    cm, keep = model.to_c()
    f = np.ascontiguousarray(force, dtype=np.float64)
    cl = np.ascontiguousarray(cols, dtype=np.int32)
    tr = np.zeros((max(len(cl), 1), steps), dtype=np.float32)
    mb = np.zeros(steps // 256 + 2, dtype=np.uint32)
    rc = L.emu_run(
        C.byref(cm), f.ctypes.data_as(C.c_void_p), C.c_int64(len(f)), cl.ctypes.data_as(C.c_void_p),
        C.c_int32(len(cl)), C.c_int32(steps), C.c_int32(tile_rows), C.c_int32(span), tr.ctypes.data_as(C.c_void_p),
        mb.ctypes.data_as(C.c_void_p),
    )
    if rc:
        raise RuntimeError(L.emu_last_error().decode())
    return tr[: len(cl)], mb
