// sur_math_check — runs sur_math.h's device functions on the GPU against
// references and exits non-zero on a mismatch (test infrastructure; the
// tests run it as a child process).
//   to_bf16: bit for bit round-to-nearest-even (software, on the device) over
//     2^24 FP32 patterns: every sign and exponent, with tie, carry and
//     boundary mantissas and hashed ones; NaN inputs must give NaN.
//   gelu, gelu_grad: against the tanh form in FP64 on 2^20 points in
//     [-16, 16]; reports the largest |error| / (1 + |z|).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "sur_math.h"

namespace {

__device__ uint16_t rne_reference(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ uint32_t pattern(uint32_t i) {
  const uint32_t se = (i >> 15) & 0x1ff, m = i & 0x7fff;
  const uint32_t sp[16] = {0, 1, 0x7fff, 0x8000, 0x8001, 0xffff, 0x10000, 0x17fff, 0x18000, 0x18001,
                           0x7fffff, 0x7f8000, 0x7f7fff, 0x7e8000, 0x400000, 0x3f8000};
  uint32_t mant;
  if (m < 16) {
    mant = sp[m];
  } else {
    uint32_t h = m * 2654435761u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    h ^= h >> 15;
    mant = h & 0x7fffff;
  }
  return (se << 23) | mant;
}

__global__ void k_round(uint32_t n, unsigned long long *bad, unsigned long long *nan_bad, uint32_t *first) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t u = pattern(i);
  const float x = __uint_as_float(u);
  const uint16_t got = gfsk::to_bf16(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    if ((got & 0x7fffu) <= 0x7f80u) atomicAdd(nan_bad, 1ull);
  } else if (got != rne_reference(x)) {
    if (atomicAdd(bad, 1ull) == 0) {
      first[0] = u;
      first[1] = rne_reference(x);
      first[2] = got;
    }
  }
}

__global__ void k_gelu(uint32_t n, float *err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = -16.0f + 32.0f * (float)i / (float)n;
  const double zd = z, c = 0.7978845608028654, a = 0.044715;
  const double t = tanh(c * (zd + a * zd * zd * zd));
  const double g = 0.5 * zd * (1.0 + t);
  const double gp = 0.5 * (1.0 + t) + 0.5 * zd * (1.0 - t * t) * c * (1.0 + 3.0 * a * zd * zd);
  err[2 * i] = (float)(fabs((double)gfsk::gelu(z) - g) / (1.0 + fabs(zd)));
  err[2 * i + 1] = (float)(fabs((double)gfsk::gelu_grad(z) - gp) / (1.0 + fabs(zd)));
}

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      printf("%s failed: %s\n", #x, hipGetErrorString(e_));                   \
      return 2;                                                               \
    }                                                                         \
  } while (0)

}  // namespace

int main() {
  const uint32_t n = 1u << 24, ng = 1u << 20;
  unsigned long long *bad = nullptr, *nb = nullptr;
  uint32_t *first = nullptr;
  float *err = nullptr;
  CHECK(hipMalloc(&bad, 8));
  CHECK(hipMalloc(&nb, 8));
  CHECK(hipMalloc(&first, 12));
  CHECK(hipMalloc(&err, sizeof(float) * 2 * ng));
  CHECK(hipMemset(bad, 0, 8));
  CHECK(hipMemset(nb, 0, 8));
  CHECK(hipMemset(first, 0, 12));
  hipLaunchKernelGGL(k_round, dim3(n / 256), dim3(256), 0, 0, n, bad, nb, first);
  CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gelu, dim3(ng / 256), dim3(256), 0, 0, ng, err);
  CHECK(hipGetLastError());
  unsigned long long hb = 0, hn = 0;
  uint32_t hf[3] = {0, 0, 0};
  static float he[2 * (1u << 20)];
  CHECK(hipMemcpy(&hb, bad, 8, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(&hn, nb, 8, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(hf, first, 12, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(he, err, sizeof(float) * 2 * ng, hipMemcpyDeviceToHost));
  float eg = 0, ed = 0;
  for (uint32_t i = 0; i < ng; i++) {
    eg = fmaxf(eg, he[2 * i]);
    ed = fmaxf(ed, he[2 * i + 1]);
  }
  printf("to_bf16: %u patterns, %llu differ from round-to-nearest-even (first %08x: want %04x got %04x), "
         "%llu NaN inputs not NaN\n", n, hb, hf[0], hf[1], hf[2], hn);
  printf("gelu: max |err| / (1 + |z|) %.3e; gelu_grad: %.3e (FP64 tanh form, %u points in [-16, 16])\n", eg, ed, ng);
  hipFree(bad);
  hipFree(nb);
  hipFree(first);
  hipFree(err);
  return (hb == 0 && hn == 0 && eg <= 1e-5f && ed <= 1e-5f) ? 0 : 1;
}
