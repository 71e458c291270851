// gemm_noalloc.hip — the surrogate's BF16 GEMM (gf_gemm_bf16) for callers
// that cannot allocate device memory: code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc.
//
// The caller passes device pointers for A, B and C and a stream; the split-K
// partial sums, which gf_gemm_bf16 allocates per call, live here in one
// file-scope device array of kScratchFloats floats (16 MiB), held by every
// GPU the library runs on until the process ends.  Every call on a GPU uses
// that GPU's array, so calls that overlap on one GPU race; the kernels write
// every partial they later read, so nothing has to be cleared between calls.
// The entry point calls no hipMalloc, hipFree or allocating library routine.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>
#include <string>

#include "../../include/geofdtd.h"
#include "surrogate_kernels.h"

namespace gfs {
int api_set_err(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *where);
int check_device(std::string &why);
}  // namespace gfs

namespace {
constexpr int64_t kScratchFloats = (int64_t)4 << 20;  // 16 MiB of split-K partials
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

// `scratch` names a host-side stand-in in host code; its device address, one
// per GPU, comes from hipGetSymbolAddress with that GPU current.
std::mutex g_mu;
float *g_scratch[kMaxDevices] = {};

int scratch_on_current_device(float **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return gfs::hip_fail(e, "gf_gemm_bf16_noalloc");
  if (dev < 0 || dev >= kMaxDevices) return gfs::api_set_err(GF_ERR_DEVICE, "geofdtd: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return gfs::hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    g_scratch[dev] = (float *)p;
  }
  *out = g_scratch[dev];
  return GF_OK;
}

}  // namespace

extern "C" int gf_gemm_bf16_noalloc(int32_t M, int32_t N, int32_t K, const uint16_t *A, int64_t lda, int32_t a_trans,
                                    const uint16_t *B, int64_t ldb, int32_t b_nk, int32_t splits, float *C,
                                    void *stream) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || splits < 1)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad gemm arguments");
  if (K % 8 || lda % 8 || ldb % 8 || (a_trans && M % 8) || (!b_nk && N % 8))
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: gemm sizes break the 8-element vector rule");
  if ((!a_trans && lda < K) || (a_trans && lda < M) || (b_nk && ldb < K) || (!b_nk && ldb < N))
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: gemm leading dimension too small");
  if (((uintptr_t)A | (uintptr_t)B) & 15)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: gemm operands must be 16-byte aligned");
  gfsk::GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = N;
  g.k_split = (((K + splits - 1) / splits) + 31) / 32 * 32;
  const int s = (K + g.k_split - 1) / g.k_split;
  g.split_stride = (int64_t)M * N;
  if ((int64_t)s * M * N > kScratchFloats) {
    char buf[160];
    snprintf(buf, sizeof buf, "geofdtd: gemm needs %lld floats of split-K scratch; the no-allocation entry point holds %lld",
             (long long)s * M * N, (long long)kScratchFloats);
    return gfs::api_set_err(GF_ERR_ARGUMENT, buf);
  }
  std::string why;
  int rc = gfs::check_device(why);
  if (rc) return gfs::api_set_err(rc, why);
  float *part = nullptr;
  if ((rc = scratch_on_current_device(&part))) return rc;
  g.A = A;
  g.B = B;
  g.cf = part;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = gfsk::launch_gemm(g, a_trans != 0, b_nk != 0, gfsk::kEpiF32, s, st);
  if (e == hipSuccess) e = gfsk::launch_sum_splits(part, s, (int64_t)M * N, C, st);
  if (e != hipSuccess) return gfs::hip_fail(e, "gf_gemm_bf16_noalloc");
  return gfs::api_set_err(GF_OK, "");
}
