// fk.hip — the frequency-wavenumber solver (Go package internal/fk of the
// reference) on gfx950, in float64.
//
// The work of fk.Medium.VerticalDisplacementMulti (fk.go:500-550) is a sum,
// per frequency, over a few thousand wavenumber samples of
//     w_i * J0(k_i r) * dk * (k_i * u(k_i) - C)
// where u(k) is the layered boundary-value solve of SurfaceResponse
// (fk.go:276-307): one 4x4 complex propagator per layer (eigenvectors,
// exponentials, an inverse by pivoted elimination, two products) and a 4x4
// pivoted solve.  Every sample is independent, and none depends on range.
//
// Layout of the work:
//   k_fk_prologue   one thread per frequency: each layer's complex constants
//                   (propsOf, fk.go:114-125) and the static coefficient
//                   C = kfar * u(kfar) (fk.go:347-353).
//   k_fk_solve      one thread per wavenumber sample, in workgroups of 256
//                   consecutive samples of one frequency (all frequencies of
//                   a call in one launch): the layered solve, k u - C stored.
//   k_fk_hankel     the same workgroups: the 256 stored values into LDS, then
//                   the Hankel sums for every range (threads split as ranges
//                   x sample slices), one partial sum per range.  Apart from
//                   the solve so that the J0 evaluations run at full
//                   occupancy while the solve uses every register.
//   k_fk_fold       one wavefront per (frequency, range): C/r - dk/2 C plus
//                   the workgroup partials, summed in a fixed order.
// The solve is FP64-VALU work with register-resident 4x4 matrices; HBM
// traffic is a few bytes per sample (the partials), so the bound is the FP64
// vector rate.  Row exchanges in the eliminations are selects over
// compile-time indices, so the matrices never leave registers.
//
// Complex arithmetic follows Go's definitions: products (ac-bd)+(ad+bc)i,
// quotients by Smith's algorithm (runtime/complex.go), square roots as
// math/cmplx.Sqrt and magnitudes as math.Hypot, so that pivots are chosen by
// the same comparisons as in the Go code.  The device compiler fuses
// multiply-adds, so results differ from Go's in the last bits.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "fk_kernels.h"

namespace gfs {
int hip_fail(hipError_t e, const char *where);
}

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

namespace gfw {
namespace {

constexpr double kPi = 3.141592653589793;

struct dcx {
  double re, im;
};

__device__ __forceinline__ dcx mk(double r, double i) { return dcx{r, i}; }
__device__ __forceinline__ dcx cadd(dcx a, dcx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ dcx csub(dcx a, dcx b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ dcx cneg(dcx a) { return mk(-a.re, -a.im); }
__device__ __forceinline__ dcx cmul(dcx a, dcx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// A real scale as Go multiplies complex(s, 0) * z.
__device__ __forceinline__ dcx rmul(double s, dcx z) { return mk(s * z.re - 0.0 * z.im, s * z.im + 0.0 * z.re); }

// Smith's quotient, as Go's runtime complex128div.
__device__ __forceinline__ dcx cdiv(dcx n, dcx m) {
  double e, f;
  if (fabs(m.re) >= fabs(m.im)) {
    double ratio = m.im / m.re;
    double denom = m.re + ratio * m.im;
    e = (n.re + n.im * ratio) / denom;
    f = (n.im - n.re * ratio) / denom;
  } else {
    double ratio = m.re / m.im;
    double denom = m.im + ratio * m.re;
    e = (n.re * ratio + n.im) / denom;
    f = (n.im * ratio - n.re) / denom;
  }
  return mk(e, f);
}

// math.Hypot for finite arguments.
__device__ __forceinline__ double hypot_go(double p, double q) {
  p = fabs(p);
  q = fabs(q);
  if (isinf(p) || isinf(q)) return INFINITY;
  if (p < q) {
    double t = p;
    p = q;
    q = t;
  }
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

// math/cmplx.Sqrt.
__device__ dcx csqrt_go(dcx x) {
  if (x.im == 0) {
    if (x.re == 0) return mk(0, x.im);
    if (x.re < 0) return mk(0, copysign(sqrt(-x.re), x.im));
    return mk(sqrt(x.re), x.im);
  }
  if (x.re == 0) {
    if (x.im < 0) {
      double t = sqrt(-0.5 * x.im);
      return mk(t, -t);
    }
    double t = sqrt(0.5 * x.im);
    return mk(t, t);
  }
  double a = x.re, b = x.im, scale;
  if (fabs(a) > 4 || fabs(b) > 4) {
    a *= 0.25;
    b *= 0.25;
    scale = 2;
  } else {
    a *= 1.8014398509481984e16;
    b *= 1.8014398509481984e16;
    scale = 7.450580596923828125e-9;
  }
  double r = hypot_go(a, b), t;
  if (a > 0) {
    t = sqrt(0.5 * r + 0.5 * a);
    r = scale * fabs((0.5 * b) / t);
    t *= scale;
  } else {
    r = sqrt(0.5 * r - 0.5 * a);
    t = scale * fabs((0.5 * b) / r);
    r *= scale;
  }
  return b < 0 ? mk(t, -r) : mk(t, r);
}

__device__ __forceinline__ dcx cexp_go(dcx x) {
  double r = exp(x.re), s, c;
  sincos(x.im, &s, &c);
  return mk(r * c, r * s);
}

// One layer's complex constants at one frequency (props, fk.go:107-112).
struct Props {
  dcx alpha, beta, mu;
  double rho, thick;
};

// visco.SLS.Modulus (visco.go:77-81).
__device__ dcx sls_modulus(double ts, double te, double f) {
  dcx w = mk(0, 2 * kPi * f);
  return cdiv(cadd(mk(1, 0), cmul(w, mk(te, 0))), cadd(mk(1, 0), cmul(w, mk(ts, 0))));
}

// complexVelocity (fk.go:92-104).
__device__ dcx complex_velocity(const Medium &M, double v, double q, double f) {
  if (M.has_relax) {
    // visco.SLS.Velocity: RelaxedVelocity(v, fRef) * Sqrt(Modulus(f)).
    double c0 = v * cdiv(mk(1, 0), csqrt_go(sls_modulus(M.tau_sigma, M.tau_eps, M.ref_freq))).re;
    if (f <= 0) return mk(c0, 0);
    return cmul(mk(c0, 0), csqrt_go(sls_modulus(M.tau_sigma, M.tau_eps, f)));
  }
  if (q <= 0 || f <= 0) return mk(v, 0);
  double gam = atan(1 / q) / kPi;
  double c0 = v * cos(kPi * gam / 2);
  dcx p = mk(pow(f / M.ref_freq, gam), 0);
  return cmul(cmul(mk(c0, 0), p), cexp_go(mk(0, kPi * gam / 2)));
}

__device__ Props props_of(const Medium &M, int l, double f) {
  Props P;
  dcx a = complex_velocity(M, M.vp[l], 2.25 * M.q[l], f);
  dcx b = complex_velocity(M, M.vs[l], M.q[l], f);
  double rho = M.rho[l];
  P.mu = cmul(cmul(mk(rho, 0), b), b);
  P.alpha = a;
  P.beta = b;
  P.rho = rho;
  P.thick = M.thick[l];
  return P;
}

// The eigen decomposition of props.eigen (fk.go:135-167) reduced to the five
// numbers its vectors are built from.  Columns (P down, SV down, P up, SV up):
//   pd = (1, -ra, -a2,  b1)   sd = ( rb, -1, -b1,  c1)
//   pu = (1,  ra,  a2,  b1)   su = (-rb, -1, -b1, -c1)
// with a2 = 2mu/mr*ra, b1 = (2mu - rc2)/mr, c1 = 2mu/mr*rb.  The Go code's
// -2*mu/mr*ra and (rc2 - 2*mu)/mr are these values negated exactly.
struct Eigen {
  dcx ra, rb, a2, b1, c1;
};

__device__ __forceinline__ Eigen eigen(const Props &p, double k, double omega, double mu_ref) {
  Eigen E;
  dcx kk = cmul(mk(k, 0), mk(k, 0));
  dcx w2 = mk(omega * omega, 0);
  dcx ra = csqrt_go(csub(mk(1, 0), cdiv(w2, cmul(cmul(kk, p.alpha), p.alpha))));
  dcx rb = csqrt_go(csub(mk(1, 0), cdiv(w2, cmul(cmul(kk, p.beta), p.beta))));
  if (ra.re < 0) ra = cneg(ra);
  if (rb.re < 0) rb = cneg(rb);
  dcx rc2 = cdiv(cmul(mk(p.rho, 0), w2), kk);
  dcx mr = mk(mu_ref, 0);
  dcx two_mu = cmul(mk(2, 0), p.mu);
  dcx t = cdiv(two_mu, mr);
  E.ra = ra;
  E.rb = rb;
  E.a2 = cmul(t, ra);
  E.b1 = cdiv(csub(two_mu, rc2), mr);
  E.c1 = cmul(t, rb);
  return E;
}

__device__ __forceinline__ void eigvecs(const Eigen &E, dcx (&e)[4][4]) {
  e[0][0] = mk(1, 0);   e[0][1] = E.rb;           e[0][2] = mk(1, 0); e[0][3] = cneg(E.rb);
  e[1][0] = cneg(E.ra); e[1][1] = mk(-1, 0);      e[1][2] = E.ra;     e[1][3] = mk(-1, 0);
  e[2][0] = cneg(E.a2); e[2][1] = cneg(E.b1);     e[2][2] = E.a2;     e[2][3] = cneg(E.b1);
  e[3][0] = E.b1;       e[3][1] = E.c1;           e[3][2] = E.b1;     e[3][3] = cneg(E.c1);
}

// cmat.solve (cmat.go:54-96) for NR right-hand sides at once: the
// elimination of `m` does not depend on the right-hand side, so this is the
// Go solve repeated per column.  Solutions overwrite x.  False when a pivot
// column is exactly zero.
template <int NR>
__device__ __forceinline__ bool solve(dcx (&m)[4][4], dcx (&x)[4][NR]) {
#pragma unroll
  for (int col = 0; col < 4; col++) {
    double best = 0;
    int br = col;
#pragma unroll
    for (int r = col; r < 4; r++) {
      double v = hypot_go(m[r][col].re, m[r][col].im);
      if (v > best) {
        best = v;
        br = r;
      }
    }
    if (best == 0) return false;
#pragma unroll
    for (int r = col + 1; r < 4; r++) {
      if (br == r) {
#pragma unroll
        for (int c = col; c < 4; c++) {
          dcx t = m[col][c];
          m[col][c] = m[r][c];
          m[r][c] = t;
        }
#pragma unroll
        for (int q = 0; q < NR; q++) {
          dcx t = x[col][q];
          x[col][q] = x[r][q];
          x[r][q] = t;
        }
      }
    }
    dcx p = m[col][col];
#pragma unroll
    for (int r = col + 1; r < 4; r++) {
      dcx f = cdiv(m[r][col], p);
#pragma unroll
      for (int c = col; c < 4; c++) m[r][c] = csub(m[r][c], cmul(f, m[col][c]));
#pragma unroll
      for (int q = 0; q < NR; q++) x[r][q] = csub(x[r][q], cmul(f, x[col][q]));
    }
  }
#pragma unroll
  for (int q = 0; q < NR; q++) {
#pragma unroll
    for (int i = 3; i >= 0; i--) {
      dcx s = x[i][q];
#pragma unroll
      for (int j = i + 1; j < 4; j++) s = csub(s, cmul(m[i][j], x[j][q]));
      x[i][q] = cdiv(s, m[i][i]);
    }
  }
  return true;
}

// SurfaceResponse (fk.go:276-307) for a validated medium: 0, or 1 when a
// layer propagator is singular, 2 when the surface system is.
__device__ int surface_response(const Props *P, int nl, double k, double omega, double mu_ref, dcx &u) {
  dcx T[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) T[i][j] = mk(i == j ? 1 : 0, 0);

  for (int l = 0; l < nl - 1; l++) {
    const Props p = P[l];
    // props.propagator (fk.go:176-196): E diag(exp((lam - max Re lam) k h)) E^-1.
    Eigen E = eigen(p, k, omega, mu_ref);
    dcx zeta = mk(k * p.thick, 0);
    double max_re = 0;
    max_re = fmax(max_re, -E.ra.re);
    max_re = fmax(max_re, -E.rb.re);
    max_re = fmax(max_re, E.ra.re);
    max_re = fmax(max_re, E.rb.re);
    dcx d[4];
    d[0] = cexp_go(cmul(csub(cneg(E.ra), mk(max_re, 0)), zeta));
    d[1] = cexp_go(cmul(csub(cneg(E.rb), mk(max_re, 0)), zeta));
    d[2] = cexp_go(cmul(csub(E.ra, mk(max_re, 0)), zeta));
    d[3] = cexp_go(cmul(csub(E.rb, mk(max_re, 0)), zeta));

    dcx m[4][4], inv[4][4];
    eigvecs(E, m);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) inv[i][j] = mk(i == j ? 1 : 0, 0);
    if (!solve<4>(m, inv)) return 1;
    // (E D) E^-1, row by row into m, then T <- P T, row by row into inv.
    eigvecs(E, m);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      dcx ed[4];
#pragma unroll
      for (int c = 0; c < 4; c++) ed[c] = cmul(m[i][c], d[c]);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        dcx s = mk(0, 0);
#pragma unroll
        for (int c = 0; c < 4; c++) s = cadd(s, cmul(ed[c], inv[c][j]));
        m[i][j] = s;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        dcx s = mk(0, 0);
#pragma unroll
        for (int c = 0; c < 4; c++) s = cadd(s, cmul(m[i][c], T[c][j]));
        inv[i][j] = s;
      }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) T[i][j] = inv[i][j];
  }

  // The half-space's decaying solutions are eigenvector columns 0 and 1.
  const Props h = P[nl - 1];
  Eigen H = eigen(h, k, omega, mu_ref);
  dcx b4 = cdiv(mk(0, -1), mk(2 * kPi * mu_ref * k, 0));
  dcx a[4][4], x[4][1];
  dcx down0[4] = {mk(1, 0), cneg(H.ra), cneg(H.a2), H.b1};
  dcx down1[4] = {H.rb, mk(-1, 0), cneg(H.b1), H.c1};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a[i][0] = T[i][0];
    a[i][1] = T[i][1];
    a[i][2] = cneg(down0[i]);
    a[i][3] = cneg(down1[i]);
    x[i][0] = cmul(cneg(b4), T[i][3]);
  }
  if (!solve<1>(a, x)) return 2;
  u = cmul(mk(0, -1), x[1][0]);
  return 0;
}

// ----------------------------------------------------------------- kernels

// Per frequency: layer constants, then (when kfar > 0) u(kfar) and C.
__global__ void k_fk_prologue(Medium M, const Job *__restrict__ jobs, int32_t njobs, Props *__restrict__ props,
                              dcx *__restrict__ ufar, dcx *__restrict__ cstat, int32_t *__restrict__ far_status) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const Job J = jobs[j];
  Props *P = props + (int64_t)j * M.nl;
  for (int l = 0; l < M.nl; l++) P[l] = props_of(M, l, J.f);
  dcx u = mk(0, 0);
  int st = 0;
  if (J.kfar > 0) st = surface_response(P, M.nl, J.kfar, 2 * kPi * J.f, M.mu_ref, u);
  ufar[j] = u;
  cstat[j] = rmul(J.kfar, u);
  far_status[j] = st;
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Each thread solves one wavenumber sample and stores k u - C; samples past
// a job's n store 0.  rem is laid out like the workgroups: 256 per workgroup.
// Workgroups of k_fk_solve resident per CU, which sets its register budget:
// 1 keeps every matrix in registers (one wavefront per SIMD); 2 halves the
// budget and spills about 90 registers to scratch, and is faster (measured
// on the bank of a three-layer medium, 163 bins x 50 ranges: 2.06 ms
// against 2.65 ms), because two wavefronts hide
// the latency of the FP64 divide and square-root sequences.
#ifndef GF_FK_SOLVE_GROUPS
#define GF_FK_SOLVE_GROUPS 2
#endif

__global__ __launch_bounds__(kBlock, GF_FK_SOLVE_GROUPS) void k_fk_solve(Medium M, const Job *__restrict__ jobs,
                                                     const int32_t *__restrict__ blk_job,
                                                     const Props *__restrict__ props, const dcx *__restrict__ cstat,
                                                     dcx *__restrict__ rem_out, unsigned long long *__restrict__ err) {
  const int t = threadIdx.x;
  const int j = blk_job[blockIdx.x];
  const Job J = jobs[j];
  const int i = (blockIdx.x - J.first_block) * kBlock + 1 + t;
  dcx rem = mk(0, 0);
  if (i <= J.n) {
    const double k = (double)i * J.dk;
    dcx u;
    int st = surface_response(props + (int64_t)j * M.nl, M.nl, k, 2 * kPi * J.f, M.mu_ref, u);
    if (st) atomicMin(&err[j], (unsigned long long)i * 2 + (unsigned long long)(st - 1));
    rem = csub(rmul(k, u), cstat[j]);
  }
  rem_out[(int64_t)blockIdx.x * kBlock + t] = rem;
}

// The Hankel sums of one workgroup's 256 samples for every range, from LDS:
// threads split as (range within a group of G = min(256, ranges left),
// slice s < S = 256 / G); slice s takes samples s, s+S, ... in order, then
// the slices are added in order.  One partial per (workgroup, range).  Kept
// apart from the solve so that it runs at full occupancy: the solve needs
// every register a lane can have, the J0 evaluations need latency hiding.
// Workgroups of k_fk_hankel resident per CU; 1 leaves the register count to
// the compiler (130, three wavefronts per SIMD).  4 (128 registers) measured
// no faster.
#ifndef GF_FK_HANKEL_GROUPS
#define GF_FK_HANKEL_GROUPS 1
#endif

__global__ __launch_bounds__(kBlock, GF_FK_HANKEL_GROUPS) void k_fk_hankel(const Job *__restrict__ jobs,
                                                      const int32_t *__restrict__ blk_job,
                                                      const dcx *__restrict__ rem_in,
                                                      const double *__restrict__ ranges, int32_t nranges,
                                                      dcx *__restrict__ partial) {
  __shared__ dcx s_rem[kBlock];
  __shared__ double s_k[kBlock], s_w[kBlock];
  __shared__ dcx s_part[kBlock];
  const int t = threadIdx.x;
  const int j = blk_job[blockIdx.x];
  const Job J = jobs[j];
  const int first = (blockIdx.x - J.first_block) * kBlock + 1;
  const int i = first + t;
  s_rem[t] = rem_in[(int64_t)blockIdx.x * kBlock + t];
  s_k[t] = (double)i * J.dk;
  s_w[t] = i > J.n ? 0.0 : i == J.n ? 0.5 : 1.0;
  __syncthreads();
  const int cnt = min(kBlock, J.n - first + 1);
  for (int g0 = 0; g0 < nranges; g0 += kBlock) {
    const int G = min(kBlock, nranges - g0);
    const int S = kBlock / G;
    const int jr = t % G, s = t / G;
    dcx acc = mk(0, 0);
    if (s < S) {
      const double r = ranges[g0 + jr];
      for (int ii = s; ii < cnt; ii += S) {
        double sw = s_w[ii] * j0(s_k[ii] * r) * J.dk;
        acc = cadd(acc, rmul(sw, s_rem[ii]));
      }
    }
    s_part[t] = acc;
    __syncthreads();
    if (t < G) {
      dcx tot = s_part[t];
      for (int q = 1; q < S; q++) tot = cadd(tot, s_part[q * G + t]);
      partial[(int64_t)blockIdx.x * nranges + g0 + t] = tot;
    }
    __syncthreads();
  }
}

// out[j][r] = C/r - (dk/2) C + the sum of job j's workgroup partials, one
// wavefront per (job, range): lane l adds partials l, l+64, l+128, ... in
// order, then the 64 lane sums are added pairwise at distances 32, 16, ...,
// 1 (lane l takes lane l+d), a fixed order, so results do not depend on
// timing.  Low frequencies have thousands of workgroups; one thread per sum
// left those chains serial.
__global__ __launch_bounds__(256) void k_fk_fold(const Job *__restrict__ jobs, int32_t njobs,
                                                 const dcx *__restrict__ cstat, const double *__restrict__ ranges,
                                                 int32_t nranges, const dcx *__restrict__ partial,
                                                 dcx *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t idx = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (idx >= (int64_t)njobs * nranges) return;  // whole wavefronts only
  const int j = (int)(idx / nranges), r = (int)(idx % nranges);
  const Job J = jobs[j];
  const int nb = (J.n + kBlock - 1) / kBlock;
  dcx s = mk(0, 0);
  for (int b = lane; b < nb; b += 64) s = cadd(s, partial[(int64_t)(J.first_block + b) * nranges + r]);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    double re = __shfl_down(s.re, d, 64), im = __shfl_down(s.im, d, 64);
    if (lane < d) s = cadd(s, mk(re, im));
  }
  if (lane == 0) {
    const dcx C = cstat[j];
    out[idx] = cadd(csub(cdiv(C, mk(ranges[r], 0)), rmul(0.5 * J.dk, C)), s);
  }
}

// synthesise's spectra (cmd/geofdtd/main.go:186-193):
// X[r][b] = u * RickerSpectrum(f) * (i 2 pi f) * exp(-i 2 pi f t0) * fs.
__global__ void k_fk_synth_coeff(const Job *__restrict__ jobs, int32_t nbins, const dcx *__restrict__ u,
                                 int32_t nranges, double peak, double f0, double t0, double fs,
                                 dcx *__restrict__ coeff) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nbins * nranges) return;
  int b = (int)(idx / nranges), r = (int)(idx % nranges);
  double f = jobs[b].f;
  // fdtd.RickerSpectrum (run.go:41-46).
  double fr = f / f0;
  double mag = 2 / sqrt(kPi) * fr * fr * exp(-fr * fr) / f0;
  double phase = -2 * kPi * f * 1.5 / f0;
  double sp, cp;
  sincos(phase, &sp, &cp);
  dcx w = cmul(mk(peak * mag, 0), mk(cp, sp));
  dcx jw = mk(0, 2 * kPi * f);
  dcx shift = cexp_go(mk(0, -2 * kPi * f * t0));
  dcx v = cmul(cmul(cmul(cmul(u[idx], w), jw), shift), mk(fs, 0));
  coeff[(int64_t)r * nbins + b] = v;
}

// dsp.IRFFT (dsp.go:42-50) of coefficients that are zero outside bins
// 1..nbins, first n samples: x_t = (2/N) sum_b Re(X_b e^{2 pi i b t / N}),
// with the Nyquist bin (if present) counted once and its imaginary part
// ignored.  The phase b t mod N is reduced exactly in integers.
__global__ void k_fk_irdft(const dcx *__restrict__ coeff, int32_t nbins, int32_t nranges, int32_t nfft, int32_t n,
                           double *__restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nranges * n) return;
  int r = (int)(idx / n), t = (int)(idx % n);
  const dcx *X = coeff + (int64_t)r * nbins;
  double s = 0;
  const int64_t N = nfft;
  for (int b = 1; b <= nbins; b++) {
    dcx x = X[b - 1];
    if (2 * b == nfft) {
      s += (t & 1) ? -x.re : x.re;
      continue;
    }
    int64_t ph = ((int64_t)b * t) % N;
    double sn, cs;
    sincospi(2.0 * (double)ph / (double)N, &sn, &cs);
    s += 2 * (x.re * cs - x.im * sn);
  }
  out[idx] = s * (1 / (double)nfft);
}

// --------------------------------------------------------------- workspace

struct Workspace {
  hipStream_t st = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  void *buf = nullptr;
  size_t cap = 0;
};

thread_local Workspace g_ws;

int ensure(size_t bytes) {
  Workspace &W = g_ws;
  if (!W.st) {
    HIPCHK(hipStreamCreate(&W.st));
    for (auto &e : W.ev) HIPCHK(hipEventCreate(&e));
  }
  if (W.cap < bytes) {
    if (W.buf) {
      HIPCHK(hipStreamSynchronize(W.st));
      HIPCHK(hipFree(W.buf));
      W.buf = nullptr;
      W.cap = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    HIPCHK(hipMalloc(&W.buf, want));
    W.cap = want;
  }
  return 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves the workspace: returns byte offsets of consecutive regions.
struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  }
};

// Device pointers of one multi() call.
struct MultiBufs {
  Job *jobs;
  int32_t *blk_job;
  Props *props;
  dcx *ufar, *cstat, *partial, *out, *rem;
  int32_t *far;
  unsigned long long *err;
  double *ranges;
  dcx *coeff;
  double *trace;
};

int launch_multi(const Medium &M, const Job *jobs, int32_t njobs, const double *ranges, int32_t nranges,
                 int64_t extra_coeff, int64_t extra_trace, MultiBufs &B, JobStatus *status, Timing *tm) {
  int64_t blocks = 0;
  std::vector<Job> J(jobs, jobs + njobs);
  for (auto &x : J) {
    x.first_block = (int32_t)blocks;
    blocks += (x.n + kBlock - 1) / kBlock;
  }
  if (blocks > INT32_MAX / 2) return gfs::hip_fail(hipErrorInvalidValue, "fk: too many wavenumber samples");
  std::vector<int32_t> blk((size_t)std::max<int64_t>(blocks, 1));
  for (int32_t j = 0; j < njobs; j++) {
    int nb = (J[j].n + kBlock - 1) / kBlock;
    for (int b = 0; b < nb; b++) blk[(size_t)J[j].first_block + b] = j;
  }
  Carve cv;
  size_t o_jobs = cv.take(sizeof(Job) * njobs);
  size_t o_blk = cv.take(sizeof(int32_t) * blk.size());
  size_t o_props = cv.take(sizeof(Props) * (size_t)njobs * M.nl);
  size_t o_ufar = cv.take(sizeof(dcx) * njobs);
  size_t o_c = cv.take(sizeof(dcx) * njobs);
  size_t o_far = cv.take(sizeof(int32_t) * njobs);
  size_t o_err = cv.take(sizeof(unsigned long long) * njobs);
  size_t o_r = cv.take(sizeof(double) * nranges);
  size_t o_part = cv.take(sizeof(dcx) * (size_t)std::max<int64_t>(blocks, 1) * nranges);
  size_t o_rem = cv.take(sizeof(dcx) * (size_t)std::max<int64_t>(blocks, 1) * kBlock);
  size_t o_out = cv.take(sizeof(dcx) * (size_t)njobs * nranges);
  size_t o_coeff = cv.take(sizeof(dcx) * (size_t)std::max<int64_t>(extra_coeff, 1));
  size_t o_trace = cv.take(sizeof(double) * (size_t)std::max<int64_t>(extra_trace, 1));
  if (int rc = ensure(cv.off)) return rc;
  char *base = (char *)g_ws.buf;
  B.jobs = (Job *)(base + o_jobs);
  B.blk_job = (int32_t *)(base + o_blk);
  B.props = (Props *)(base + o_props);
  B.ufar = (dcx *)(base + o_ufar);
  B.cstat = (dcx *)(base + o_c);
  B.far = (int32_t *)(base + o_far);
  B.err = (unsigned long long *)(base + o_err);
  B.ranges = (double *)(base + o_r);
  B.partial = (dcx *)(base + o_part);
  B.rem = (dcx *)(base + o_rem);
  B.out = (dcx *)(base + o_out);
  B.coeff = (dcx *)(base + o_coeff);
  B.trace = (double *)(base + o_trace);
  hipStream_t st = g_ws.st;
  HIPCHK(hipMemcpyAsync(B.jobs, J.data(), sizeof(Job) * njobs, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.blk_job, blk.data(), sizeof(int32_t) * blk.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.ranges, ranges, sizeof(double) * nranges, hipMemcpyHostToDevice, st));
  k_fill_u64<<<(njobs + 255) / 256, 256, 0, st>>>((uint64_t *)B.err, njobs, ~0ull);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g_ws.ev[0], st));
  k_fk_prologue<<<(njobs + 63) / 64, 64, 0, st>>>(M, B.jobs, njobs, B.props, B.ufar, B.cstat, B.far);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g_ws.ev[1], st));
  if (blocks > 0) {
    k_fk_solve<<<(unsigned)blocks, kBlock, 0, st>>>(M, B.jobs, B.blk_job, B.props, B.cstat, B.rem, B.err);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(g_ws.ev[2], st));
  if (blocks > 0) {
    k_fk_hankel<<<(unsigned)blocks, kBlock, 0, st>>>(B.jobs, B.blk_job, B.rem, B.ranges, nranges, B.partial);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(g_ws.ev[3], st));
  int64_t nout = (int64_t)njobs * nranges;  // one wavefront each, four per workgroup
  k_fk_fold<<<(unsigned)((nout + 3) / 4), 256, 0, st>>>(B.jobs, njobs, B.cstat, B.ranges, nranges, B.partial, B.out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g_ws.ev[4], st));
  std::vector<int32_t> far(njobs);
  std::vector<unsigned long long> err(njobs);
  HIPCHK(hipMemcpyAsync(far.data(), B.far, sizeof(int32_t) * njobs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(err.data(), B.err, sizeof(unsigned long long) * njobs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int32_t j = 0; j < njobs; j++) {
    status[j].far = far[j];
    status[j].sample = err[j];
  }
  if (tm) {
    float a = 0, b = 0, c = 0, d = 0;
    HIPCHK(hipEventElapsedTime(&a, g_ws.ev[0], g_ws.ev[1]));
    HIPCHK(hipEventElapsedTime(&b, g_ws.ev[1], g_ws.ev[2]));
    HIPCHK(hipEventElapsedTime(&c, g_ws.ev[2], g_ws.ev[3]));
    HIPCHK(hipEventElapsedTime(&d, g_ws.ev[3], g_ws.ev[4]));
    tm->prologue_ms = a;
    tm->solve_ms = b;
    tm->hankel_ms = c;
    tm->fold_ms = d;
    tm->launches = 3 + 2 * (blocks > 0);
  }
  return 0;
}

}  // namespace

int surface_points(const Medium &M, const double *k, const double *f, int64_t n, double *out, int32_t *status) {
  if (n <= 0) return 0;
  if (n > INT32_MAX / 4) return gfs::hip_fail(hipErrorInvalidValue, "fk: too many points");
  std::vector<Job> J((size_t)n);
  for (int64_t i = 0; i < n; i++) J[i] = Job{f[i], 0, k[i], 0, 0};
  Carve cv;
  size_t o_jobs = cv.take(sizeof(Job) * n);
  size_t o_props = cv.take(sizeof(Props) * (size_t)n * M.nl);
  size_t o_u = cv.take(sizeof(dcx) * n);
  size_t o_c = cv.take(sizeof(dcx) * n);
  size_t o_far = cv.take(sizeof(int32_t) * n);
  if (int rc = ensure(cv.off)) return rc;
  char *base = (char *)g_ws.buf;
  hipStream_t st = g_ws.st;
  Job *dj = (Job *)(base + o_jobs);
  HIPCHK(hipMemcpyAsync(dj, J.data(), sizeof(Job) * n, hipMemcpyHostToDevice, st));
  k_fk_prologue<<<(unsigned)((n + 63) / 64), 64, 0, st>>>(M, dj, (int32_t)n, (Props *)(base + o_props),
                                                          (dcx *)(base + o_u), (dcx *)(base + o_c),
                                                          (int32_t *)(base + o_far));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, base + o_u, sizeof(dcx) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(status, base + o_far, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int multi(const Medium &M, const Job *jobs, int32_t njobs, const double *ranges, int32_t nranges, double *out,
          JobStatus *status, Timing *t) {
  if (njobs <= 0) return 0;
  MultiBufs B;
  if (int rc = launch_multi(M, jobs, njobs, ranges, nranges, 0, 0, B, status, t)) return rc;
  if (out) {
    HIPCHK(hipMemcpyAsync(out, B.out, sizeof(dcx) * (size_t)njobs * nranges, hipMemcpyDeviceToHost, g_ws.st));
    HIPCHK(hipStreamSynchronize(g_ws.st));
  }
  return 0;
}

int synthesise(const Medium &M, const Job *jobs, int32_t nbins, const double *ranges, int32_t nranges, double peak,
               double f0, double t0, double fs, int32_t nfft, int32_t n, double *out, JobStatus *status, Timing *t) {
  if (nranges <= 0 || n <= 0) return 0;
  if (nbins <= 0) {
    std::fill(out, out + (size_t)nranges * n, 0.0);
    return 0;
  }
  MultiBufs B;
  if (int rc = launch_multi(M, jobs, nbins, ranges, nranges, (int64_t)nbins * nranges, (int64_t)nranges * n, B,
                            status, t))
    return rc;
  for (int32_t j = 0; j < nbins; j++)
    if (status[j].far || status[j].sample != ~0ull) return 0;  // the caller reports it
  hipStream_t st = g_ws.st;
  int64_t nc = (int64_t)nbins * nranges;
  k_fk_synth_coeff<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>(B.jobs, nbins, B.out, nranges, peak, f0, t0, fs,
                                                                 B.coeff);
  HIPCHK(hipGetLastError());
  int64_t nt = (int64_t)nranges * n;
  k_fk_irdft<<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(B.coeff, nbins, nranges, nfft, n, B.trace);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, B.trace, sizeof(double) * nt, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (t) t->launches += 2;
  return 0;
}

}  // namespace gfw
