// spectrum.hip — FDTD traces to bank spectra on the device.
//
// One workgroup per (row, frequency): 256 lanes stride through the samples,
// each carrying its own twiddle e^{-i w (t0 + n dt)} by complex recurrence
// (exact phase at the lane's first sample, then a fixed rotation of
// 256 dt per iteration), then an LDS tree sum.  FP64 throughout: a 1 s trace
// at 40 kHz is ~20k samples, and the bank's low bins need the sum exact to
// well below the FP32 payload's precision.  This is not the hot path (it is a
// few milliseconds per bank); it keeps traces in HBM instead of shipping them
// to the host.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace gfk {

constexpr int kDftThreads = 256;

template <typename T>
__global__ __launch_bounds__(kDftThreads) void k_dft(const T *__restrict__ x, int64_t ld, int32_t n, double dt,
                                                     double t0, const double *__restrict__ freqs, int32_t nbins,
                                                     double *__restrict__ out) {
  const int row = blockIdx.x / nbins, b = blockIdx.x % nbins;
  const int tid = threadIdx.x;
  const double w = -2.0 * 3.141592653589793 * freqs[b];
  double s0, c0, ss, cs;
  sincos(w * (t0 + (double)tid * dt), &s0, &c0);
  sincos(w * (double)kDftThreads * dt, &ss, &cs);
  double re = 0, im = 0, tr = c0, ti = s0;
  const T *xr = x + (size_t)row * ld;
  for (int k = tid; k < n; k += kDftThreads) {
    const double v = (double)xr[k];
    re += v * tr;
    im += v * ti;
    const double nr = tr * cs - ti * ss;
    ti = tr * ss + ti * cs;
    tr = nr;
  }
  __shared__ double sr[kDftThreads], si[kDftThreads];
  sr[tid] = re;
  si[tid] = im;
  __syncthreads();
  for (int off = kDftThreads / 2; off > 0; off >>= 1) {
    if (tid < off) {
      sr[tid] += sr[tid + off];
      si[tid] += si[tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[2 * ((size_t)row * nbins + b)] = dt * sr[0];
    out[2 * ((size_t)row * nbins + b) + 1] = dt * si[0];
  }
}

hipError_t launch_dft_f32(const float *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st) {
  if (rows <= 0 || nbins <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dft<float>, dim3(rows * nbins), dim3(kDftThreads), 0, st, x, ld, n, dt, t0, freqs, nbins, out);
  return hipGetLastError();
}

hipError_t launch_dft_f64(const double *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st) {
  if (rows <= 0 || nbins <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dft<double>, dim3(rows * nbins), dim3(kDftThreads), 0, st, x, ld, n, dt, t0, freqs, nbins,
                     out);
  return hipGetLastError();
}

}  // namespace gfk
