// fk_kernels.h — device side of the frequency-wavenumber solver: the medium
// and per-frequency job descriptions the host builds, and the launchers.
#pragma once

#include <stdint.h>

namespace gfw {

constexpr int kMaxLayers = 16;
constexpr int kBlock = 256;  // wavenumber samples per quadrature workgroup

// fk.Medium resolved on the host: per-layer constants with qOf applied, the
// reference modulus, and the relaxation model.
struct Medium {
  int32_t nl;
  int32_t has_relax;
  double ref_freq;
  double mu_ref;
  double tau_sigma, tau_eps;
  double thick[kMaxLayers], vp[kMaxLayers], vs[kMaxLayers], rho[kMaxLayers], q[kMaxLayers];
};

// One frequency of a VerticalDisplacementMulti call: its grid (dk, n), the
// wavenumber of its static coefficient, and where its workgroups start.
struct Job {
  double f;
  double dk;
  double kfar;
  int32_t n;
  int32_t first_block;
};

// What a job's Go call would have reported.  far: 0, or 1/2 when the static
// coefficient's solve hit a singular propagator / surface system.  sample:
// 2*i + (kind-1) for the first sample i whose solve was singular, else ~0.
struct JobStatus {
  int32_t far;
  uint64_t sample;
};

struct Timing {
  double prologue_ms, solve_ms, hankel_ms, fold_ms;
  int32_t launches;
};

// SurfaceResponse at (k[i], f[i]); status[i] as JobStatus::far.
int surface_points(const Medium &M, const double *k, const double *f, int64_t n, double *out, int32_t *status);

// VerticalDisplacementMulti for every job: out [njobs][nranges][2] on the
// host (may be null when only the statuses are wanted).
int multi(const Medium &M, const Job *jobs, int32_t njobs, const double *ranges, int32_t nranges, double *out,
          JobStatus *status, Timing *t);

// The synthesise step after multi() over bins 1..nbins of an nfft-point
// transform: coefficients, inverse real DFT, first n samples per range.
int synthesise(const Medium &M, const Job *jobs, int32_t nbins, const double *ranges, int32_t nranges, double peak,
               double f0, double t0, double fs, int32_t nfft, int32_t n, double *out, JobStatus *status, Timing *t);

}  // namespace gfw
