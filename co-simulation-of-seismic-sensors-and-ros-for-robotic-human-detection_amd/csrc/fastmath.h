// fastmath.h — FP64 e^x and sin/cos for device kernels that evaluate them per
// value (bank lookups), written from a published method so that each is a
// short straight-line sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gff {

// e^x and (sin x, cos x) by Cody and Waite's argument reduction (Software
// Manual for the Elementary Functions, 1980) with fused multiply-adds, then
// Taylor series: x = n ln 2 + r with |r| <= ln2 / 2 and e^r to r^13
// (remainder below 5e-18 relative);
// x = q pi/2 + r with pi/2 in three parts, |r| <= pi/4, sin r to r^17 and
// cos r to r^18 (remainders below 1e-19), the quadrant q mod 4 choosing
// signs and roles.  Accurate to a few ulp for |x| below about 1e5, at a
// fraction of the general routines' cost; host-callable, so a CPU check can
// compare them with libm.
constexpr double kInvLn2 = 1.4426950408889634, kLn2Hi = 0.6931471805599453, kLn2Lo = 2.3190468138462996e-17;
constexpr double kTwoOverPi = 0.6366197723675814, kPio2A = 1.5707963267948966, kPio2B = 6.123233995736766e-17,
                 kPio2C = -1.4973849048591698e-33;

// A coefficient in scalar registers: on the device, an empty asm statement
// with an "s" operand keeps the value in an SGPR pair, so each fused
// multiply-add reads it from there instead of first copying the 64-bit
// literal into vector registers (two v_mov per coefficient per wave).
__host__ __device__ __forceinline__ double sc(double c) {
#ifdef __HIP_DEVICE_COMPILE__
  __asm__("" : "+s"(c));
#endif
  return c;
}

__host__ __device__ __forceinline__ double exp_reduced(double x) {
  if (x < -746.0) return 0.0;
  const double n = rint(x * sc(kInvLn2));
  double r = fma(-n, sc(kLn2Hi), x);
  r = fma(-n, sc(kLn2Lo), r);
  // 1/k! for k = 13 down to 0
  double p = sc(1.6059043836821613e-10);
  p = fma(p, r, sc(2.08767569878681e-09));
  p = fma(p, r, sc(2.505210838544172e-08));
  p = fma(p, r, sc(2.755731922398589e-07));
  p = fma(p, r, sc(2.7557319223985893e-06));
  p = fma(p, r, sc(2.48015873015873e-05));
  p = fma(p, r, sc(0.0001984126984126984));
  p = fma(p, r, sc(0.001388888888888889));
  p = fma(p, r, sc(0.008333333333333333));
  p = fma(p, r, sc(0.041666666666666664));
  p = fma(p, r, sc(0.16666666666666666));
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

__host__ __device__ __forceinline__ void sincos_reduced(double x, double *sn, double *cs) {
  const double q = rint(x * sc(kTwoOverPi));
  double r = fma(-q, sc(kPio2A), x);
  r = fma(-q, sc(kPio2B), r);
  r = fma(-q, sc(kPio2C), r);
  const double r2 = r * r;
  // sin r = r + r^3 (-1/3! + r^2/5! - ... + r^14/17!)
  double ps = sc(2.8114572543455206e-15);
  ps = fma(ps, r2, sc(-7.647163731819816e-13));
  ps = fma(ps, r2, sc(1.6059043836821613e-10));
  ps = fma(ps, r2, sc(-2.505210838544172e-08));
  ps = fma(ps, r2, sc(2.7557319223985893e-06));
  ps = fma(ps, r2, sc(-0.0001984126984126984));
  ps = fma(ps, r2, sc(0.008333333333333333));
  ps = fma(ps, r2, sc(-0.16666666666666666));
  const double sr = fma(r * r2, ps, r);
  // cos r = 1 + r^2 (-1/2! + r^2/4! - ... - r^16/18!)
  double pc = sc(-1.5619206968586225e-16);
  pc = fma(pc, r2, sc(4.779477332387385e-14));
  pc = fma(pc, r2, sc(-1.1470745597729725e-11));
  pc = fma(pc, r2, sc(2.08767569878681e-09));
  pc = fma(pc, r2, sc(-2.755731922398589e-07));
  pc = fma(pc, r2, sc(2.48015873015873e-05));
  pc = fma(pc, r2, sc(-0.001388888888888889));
  pc = fma(pc, r2, sc(0.041666666666666664));
  pc = fma(pc, r2, -0.5);
  const double cr = fma(r2, pc, 1.0);
  // quadrant q mod 4: (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr),
  // chosen with selects rather than branches
  const int qi = (int)((long long)q & 3);
  const double s = (qi & 1) ? cr : sr, c = (qi & 1) ? sr : cr;
  *sn = (qi & 2) ? -s : s;
  *cs = ((qi + 1) & 2) ? -c : c;
}

}  // namespace gff
