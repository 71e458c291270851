"""gf_gemm_bf16_noalloc, the GEMM entry point that allocates nothing, against
gf_gemm_bf16 (the entry point it is a variant of) through the C ABI: two
calls in a row, the second smaller, each bit for bit equal to the existing
entry point's result; a request larger than its scratch is refused."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16(rng, shape):
    f = rng.normal(size=shape).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


class Device:
    """Device buffers from the HIP runtime the library itself runs on (the
    test allocates; the entry point under test does not)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)


def test_noalloc_gemm_equals_gf_gemm_bf16(fdtd):
    lib = fdtd.load_library()
    dev = Device()
    fn = lib.gf_gemm_bf16_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64,
                   C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    ref = lib.gf_gemm_bf16
    P16 = C.POINTER(C.c_uint16)
    rng = np.random.default_rng(5)
    # (M, N, K, a_trans, b_nk, splits): a weight-gradient shape with split K,
    # then a smaller forward-layer shape
    for M, N, K, at, bnk, splits in [(512, 512, 4096, 1, 0, 8), (136, 72, 264, 0, 1, 3)]:
        a = bf16(rng, (K, M) if at else (M, K))
        b = bf16(rng, (N, K) if bnk else (K, N))
        lda = M if at else K
        ldb = K if bnk else N
        want = np.zeros((M, N), dtype=np.float32)
        rc = ref(M, N, K, a.ctypes.data_as(P16), lda, at, b.ctypes.data_as(P16), ldb, bnk, splits,
                 want.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0, lib.gf_last_error()
        da, db = dev.put(a), dev.put(b)
        dc = dev.put(np.full((M, N), np.nan, dtype=np.float32))
        print(f"noalloc gemm M={M} N={N} K={K} at={at} b_nk={bnk} splits={splits}", flush=True)
        rc = fn(M, N, K, da, lda, at, db, ldb, bnk, splits, dc, None)
        assert rc == 0, lib.gf_last_error()
        np.testing.assert_array_equal(dev.get(dc, want), want)
    # more split-K partials than the scratch holds: refused before any launch
    dz = dev.put(np.zeros(8, dtype=np.float32))
    rc = fn(4096, 4096, 4096, dz, 4096, 0, dz, 4096, 1, 2, dz, None)
    assert rc != 0 and "scratch" in lib.gf_last_error().decode()
    dev.close()
