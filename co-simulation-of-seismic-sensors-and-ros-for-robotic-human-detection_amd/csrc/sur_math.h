// sur_math.h — the surrogate's per-element device functions: BF16 rounding
// and the GELU activation with its derivative.  Shared by every kernel that
// rounds or activates (the per-layer GEMMs, the fused inference kernel) so
// that all of them give the same bits, and by the rounding check that runs
// on the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfsk {

__device__ __forceinline__ uint16_t to_bf16(float x) {
  // round to nearest even, NaN stays NaN: gfx950's v_cvt_pk_bf16_f32
  const __bf16 h = (__bf16)x;
  uint16_t r;
  __builtin_memcpy(&r, &h, 2);
  return r;
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = c (z + a z^3), c = sqrt(2/pi),
// a = 0.044715.  With r = 1 / (1 + e^{2u}), tanh(u) = 1 - 2r, so
// GELU = z - z r and GELU' = (1 - r) + 2 z r (1 - r) c (1 + 3 a z^2).
// e^{2u} and the reciprocal are the hardware's (v_exp_f32, v_rcp_f32, about
// 1 ulp); r runs to 0 and 1 exactly when e^{2u} overflows or vanishes.
constexpr float kGc = 0.7978845608028654f, kGa = 0.044715f;

// e^{2u} = 2^{z (K + K a z^2)}, K = 2 c log2(e): three VALU operations before
// the exp2.
constexpr float kGk = 2.0f * kGc * 1.4426950408889634f;

__device__ __forceinline__ float gelu_r(float z) {
  const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(z * z, kGk * kGa, kGk));
  return __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ __forceinline__ float gelu(float z) { return __builtin_fmaf(-z, gelu_r(z), z); }

__device__ __forceinline__ float gelu_grad(float z) {
  const float r = gelu_r(z), q = 1.0f - r;
  return q + 2.0f * z * r * q * kGc * (1.0f + 3.0f * kGa * z * z);
}

}  // namespace gfsk
