// stream.hip — stream.Convolver (internal/stream/stream.go of the reference)
// for many receivers on gfx950.
//
// The reference convolves a chunk at a time by uniformly partitioned
// overlap-add (FFTs of 2C points, a frequency-domain delay line, a carried
// tail), so that chunked output equals the whole-record convolution truncated
// to the input length.  Here the same linear map is computed directly: each
// receiver keeps its past input in a ring in HBM (M >= len + chunk - 1
// samples, a power of two, zero at the start, so samples before the first
// chunk read as zeros) and each output sample is
//     y(s) = sum_{j < len} h[j] x(s - j),
// summed over j in order.  A streaming chunk is a few dozen samples, so the
// GPU's cost per Process is launch latency; the direct form needs one kernel
// per chunk (write the chunk into the ring, then the sums) and is exact for any
// chunk size, where a transform pipeline would need several launches and a
// mixed-radix FFT for chunk sizes like 7 or 20.
//
//   k_ring_write  one thread per (receiver, sample): the newest min(n, M)
//                 samples into the ring (older ones would be overwritten).
//   k_fir_chunk   a streaming chunk of at most 64 samples: 1024-thread
//                 workgroups of a few outputs of one receiver, lanes split
//                 over outputs and tap slices, so the serial chain is a small
//                 fraction of the response.
//   k_fir<RING>   one workgroup per (receiver, 2048 outputs): taps in tiles of
//                 256 staged in LDS with the 2303 input samples they touch,
//                 each thread accumulating eight consecutive outputs over the
//                 tile's taps in order from a sliding register window.
//                 RING reads the ring (Process); otherwise a linear buffer of
//                 the last len - 1 samples followed by the new input
//                 (ProcessAll over a whole record in one launch).
// Work: len multiply-adds per output sample, FP64; for long records the
// LDS-staged form runs at about half the FP64 FMA peak.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"

namespace gfs {
int hip_fail(hipError_t e, const char *where);
int api_set_err(int code, const std::string &msg);
int check_device(std::string &why);
}  // namespace gfs

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

struct gf_convolver {
  int32_t nrecv = 0, len = 0, chunk = 0, parts = 0;
  int64_t ring = 0;  // ring length M (power of two)
  int64_t pos = 0;   // samples consumed so far: the next input is x(pos)
  double *h = nullptr;     // [nrecv][len]
  double *x = nullptr;     // [nrecv][ring]
  double *in = nullptr;    // staging [nrecv][cap_in]
  double *out = nullptr;   // [nrecv][cap_out]
  double *pin_in = nullptr, *pin_out = nullptr;  // pinned host chunks [nrecv][chunk]
  size_t cap_in = 0, cap_out = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0;
};

namespace {

constexpr int kT = 256;  // outputs per workgroup and taps per LDS tile

__global__ void k_fill(double *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ring[r][(s0 + i) mod M] = src[r][i] for the last min(n, M) of the n samples.
__global__ void k_ring_write(double *__restrict__ ring, int64_t M, int64_t s0, const double *__restrict__ src,
                             int64_t n, int64_t ld) {
  const int r = blockIdx.y;
  const int64_t first = n > M ? n - M : 0;
  const int64_t i = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ring[(int64_t)r * M + ((s0 + i) & (M - 1))] = src[(int64_t)r * ld + i];
}

// out[r][t] = sum_{j < len} h[r][j] x_r(s0 + t - j), t < n.  RING: x_r(s) =
// src[r][s mod M]; else x_r(s) = src[r][s - s0 + len - 1] (a linear buffer
// that starts len - 1 samples before s0), with row stride ld.
//
// A workgroup makes kOut = 8 x 256 consecutive outputs; thread u owns outputs
// 8u..8u+7.  Per tile of kT taps the inputs they meet (kOut + kT - 1) are
// staged in LDS; the thread keeps a sliding window of eight inputs in
// registers, so each tap costs one broadcast tap read, one input read and
// eight multiply-adds, with no test in the tap loop (it unrolls into 16-byte
// LDS reads and runs of multiply-adds).  Each output still sums its taps in
// order.
constexpr int kR = 8;
constexpr int kOut = kR * kT;

template <bool RING>
__global__ __launch_bounds__(kT) void k_fir(const double *__restrict__ h, int32_t len, const double *__restrict__ src,
                                            int64_t ld, int64_t M, int64_t s0, int64_t n,
                                            double *__restrict__ out, int64_t ldo) {
  __shared__ double sh[kT];
  // sx[1 + m] holds input m; sx[0] is a spare slot, so the window's last
  // slide (whose value is never used) reads inside the array without a test.
  __shared__ double sx[1 + kOut + kT - 1];
  const int r = blockIdx.y;
  const int u = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * kOut;
  const double *hr = h + (int64_t)r * len;
  const double *xr = src + (int64_t)r * ld;
  double acc[kR];
#pragma unroll
  for (int i = 0; i < kR; i++) acc[i] = 0;
  for (int j0 = 0; j0 < len; j0 += kT) {
    // Taps j0..j0+kT-1 and the inputs they meet: s = s0 + t0 - j0 - (kT-1) + m.
    const int jt = min(kT, len - j0);
    sh[u] = u < jt ? hr[j0 + u] : 0.0;
    for (int m = u; m < kOut + kT - 1; m += kT) {
      const int64_t s = s0 + t0 - j0 - (kT - 1) + m;
      double v;
      if (RING) {
        v = xr[s & (M - 1)];
      } else {
        const int64_t k = s - s0 + len - 1;  // >= 0 for every tap in range
        v = k >= 0 && k < ld ? xr[k] : 0.0;
      }
      sx[1 + m] = v;
    }
    __syncthreads();
    // x(s0 + t0 + t - j) for j = j0 + q sits at sx[t - q + kT - 1]; for the
    // thread's outputs t = kR u + i that is w[i] = sx[kR u + i - q + kT - 1].
    const int b = kR * u + kT - 1;
    double w[kR];
#pragma unroll
    for (int i = 0; i < kR; i++) w[i] = sx[1 + b + i];
    for (int q = 0; q < jt; q++) {
      const double hq = sh[q];
#pragma unroll
      for (int i = 0; i < kR; i++) acc[i] = fma(hq, w[i], acc[i]);
      // Slide the window one tap: w[i] <- sx[b + i - (q + 1)].
#pragma unroll
      for (int i = kR - 1; i > 0; i--) w[i] = w[i - 1];
      w[0] = sx[b - q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < kR; i++) {
    const int64_t t = t0 + kR * u + i;
    if (t < n) out[(int64_t)r * ldo + t] = acc[i];
  }
}

// The latency form for one streaming chunk (n <= 64 outputs): 1024-thread
// workgroups, each taking np outputs of one receiver, lanes split as np
// outputs x S = 1024 / np tap slices; np is chosen on the host so that the
// grid covers the 256 CUs (64 receivers x 20 samples: 5 outputs a workgroup).
// Slice q sums taps q, q + S, q + 2S, ... in order straight from the ring (it
// sits in L2); the slice sums are then added in a fixed order (groups of 32,
// then the groups).  The chain is latency-bound, so the serial length per
// lane, len / S taps, sets the time.
//
// Every workgroup first writes its receiver's new chunk into the ring (one
// launch per chunk, not two; workgroups of one receiver write identical
// values); __syncthreads makes those writes visible to the workgroup's own
// reads, and no workgroup reads a new sample it has not itself written.
__global__ __launch_bounds__(1024) void k_fir_chunk(const double *__restrict__ h, int32_t len, double *ring, int64_t M,
                                                    int64_t s0, const double *__restrict__ in, int32_t n, int32_t np,
                                                    double *__restrict__ out) {
  __shared__ double part[1024];
  __shared__ double grp[1024 / 32 + 64];
  const int r = blockIdx.x;
  const int t = threadIdx.x % np, q = threadIdx.x / np;
  const int S = 1024 / np;
  const int i = blockIdx.y * np + t;
  const double *hr = h + (int64_t)r * len;
  double *xr = ring + (int64_t)r * M;
  // every workgroup of the receiver writes the same chunk: identical values
  if (threadIdx.x < n) xr[(s0 + threadIdx.x) & (M - 1)] = in[(int64_t)r * n + threadIdx.x];
  __syncthreads();
  double acc = 0;
  if (i < n && q < S) {
#pragma unroll 8
    for (int j = q; j < len; j += S) acc = fma(hr[j], xr[(s0 + i - j) & (M - 1)], acc);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  const int G = (S + 31) / 32;
  if (q < G && i < n) {
    double g = 0;
    for (int k = q * 32; k < q * 32 + 32 && k < S; k++) g += part[k * np + t];
    grp[q * np + t] = g;
  }
  __syncthreads();
  if (q == 0 && i < n) {
    double tot = 0;
    for (int k = 0; k < G; k++) tot += grp[k * np + t];
    out[(int64_t)r * n + i] = tot;
  }
}

int fail(int code, const std::string &m) { return gfs::api_set_err(code, m); }
int ok() { return gfs::api_set_err(GF_OK, ""); }

int grow(gf_convolver *c, size_t in_elems, size_t out_elems) {
  if (c->cap_in < in_elems) {
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->in) HIPCHK(hipFree(c->in));
    c->in = nullptr;
    HIPCHK(hipMalloc(&c->in, sizeof(double) * in_elems));
    c->cap_in = in_elems;
  }
  if (c->cap_out < out_elems) {
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->out) HIPCHK(hipFree(c->out));
    c->out = nullptr;
    HIPCHK(hipMalloc(&c->out, sizeof(double) * out_elems));
    c->cap_out = out_elems;
  }
  return GF_OK;
}

void release(gf_convolver *c) {
  if (c->st) hipStreamSynchronize(c->st);
  hipFree(c->h);
  hipFree(c->x);
  hipFree(c->in);
  hipFree(c->out);
  if (c->pin_in) hipHostFree(c->pin_in);
  if (c->pin_out) hipHostFree(c->pin_out);
  if (c->e0) hipEventDestroy(c->e0);
  if (c->e1) hipEventDestroy(c->e1);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
}

int zero_ring(gf_convolver *c) {
  const int64_t n = (int64_t)c->nrecv * c->ring;
  k_fill<<<(unsigned)((n + 255) / 256), 256, 0, c->st>>>(c->x, n);
  HIPCHK(hipGetLastError());
  c->pos = 0;
  return GF_OK;
}

}  // namespace

extern "C" {

int gf_convolver_create(const double *impulse, int32_t n_recv, int32_t len, int32_t chunk, gf_convolver **out) {
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  *out = nullptr;
  if (chunk <= 0) return fail(GF_ERR_INVALID, "stream: chunk size must be positive, got " + std::to_string(chunk));
  if (len <= 0 || !impulse) return fail(GF_ERR_INVALID, "stream: impulse response must not be empty");
  if (n_recv <= 0 || n_recv > 65535) return fail(GF_ERR_ARGUMENT, "geofdtd: receivers must be in [1, 65535]");
  std::string why;
  if (int rc = gfs::check_device(why)) return fail(rc, why);
  gf_convolver *c = new gf_convolver;
  c->nrecv = n_recv;
  c->len = len;
  c->chunk = chunk;
  c->parts = (int32_t)(((int64_t)len + chunk - 1) / chunk);
  int64_t M = 1;
  while (M < (int64_t)len - 1 + chunk) M <<= 1;
  c->ring = M;
  auto bail = [&](hipError_t e, const char *where) {
    release(c);
    return gfs::hip_fail(e, where);
  };
  hipError_t e;
  if ((e = hipStreamCreate(&c->st)) != hipSuccess) return bail(e, "hipStreamCreate");
  if ((e = hipEventCreate(&c->e0)) != hipSuccess || (e = hipEventCreate(&c->e1)) != hipSuccess)
    return bail(e, "hipEventCreate");
  if ((e = hipMalloc(&c->h, sizeof(double) * (size_t)n_recv * len)) != hipSuccess ||
      (e = hipMalloc(&c->x, sizeof(double) * (size_t)n_recv * M)) != hipSuccess)
    return bail(e, "hipMalloc");
  if ((e = hipMemcpyAsync(c->h, impulse, sizeof(double) * (size_t)n_recv * len, hipMemcpyHostToDevice, c->st)) !=
      hipSuccess)
    return bail(e, "hipMemcpyAsync");
  if (zero_ring(c) != GF_OK) {
    release(c);
    return GF_ERR_DEVICE;
  }
  if ((e = hipStreamSynchronize(c->st)) != hipSuccess) return bail(e, "hipStreamSynchronize");
  *out = c;
  return ok();
}

void gf_convolver_destroy(gf_convolver *c) {
  if (c) release(c);
}

int gf_convolver_info(const gf_convolver *c, int32_t *chunk, int32_t *partitions, int32_t *n_recv) {
  if (!c) return fail(GF_ERR_ARGUMENT, "geofdtd: null convolver");
  if (chunk) *chunk = c->chunk;
  if (partitions) *partitions = c->parts;
  if (n_recv) *n_recv = c->nrecv;
  return ok();
}

int gf_convolver_process(gf_convolver *c, const double *in, int32_t len_in, double *out, int32_t len_out) {
  if (!c) return fail(GF_ERR_ARGUMENT, "geofdtd: null convolver");
  if (len_in != c->chunk || len_out != c->chunk)
    return fail(GF_ERR_INVALID, "stream: chunk length mismatch: in " + std::to_string(len_in) + ", out " +
                                    std::to_string(len_out) + ", want " + std::to_string(c->chunk));
  if (!in || !out) return fail(GF_ERR_ARGUMENT, "geofdtd: null chunk");
  const int64_t C = c->chunk;
  if (int rc = grow(c, (size_t)c->nrecv * C, (size_t)c->nrecv * C)) return rc;
  // Chunks go through pinned host buffers: a pageable copy costs a staging
  // round trip that would dominate a 20-sample chunk's latency.
  const size_t bytes = sizeof(double) * c->nrecv * C;
  if (!c->pin_in) {
    HIPCHK(hipHostMalloc((void **)&c->pin_in, bytes, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&c->pin_out, bytes, hipHostMallocDefault));
  }
  memcpy(c->pin_in, in, bytes);
  HIPCHK(hipMemcpyAsync(c->in, c->pin_in, bytes, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipEventRecord(c->e0, c->st));
  if (C <= 64) {
    // outputs per workgroup: few enough that the grid covers the 256 CUs
    const int64_t want = (C * c->nrecv + 255) / 256;
    const int32_t np = (int32_t)std::max<int64_t>(1, std::min<int64_t>(want, C));
    dim3 g(c->nrecv, (unsigned)((C + np - 1) / np));
    k_fir_chunk<<<g, 1024, 0, c->st>>>(c->h, c->len, c->x, c->ring, c->pos, c->in, (int32_t)C, np, c->out);
  } else {
    dim3 gw((unsigned)((C + 255) / 256), c->nrecv);
    k_ring_write<<<gw, 256, 0, c->st>>>(c->x, c->ring, c->pos, c->in, C, C);
    HIPCHK(hipGetLastError());
    dim3 gf((unsigned)((C + kOut - 1) / kOut), c->nrecv);
    k_fir<true><<<gf, kT, 0, c->st>>>(c->h, c->len, c->x, c->ring, c->ring, c->pos, C, c->out, C);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->e1, c->st));
  HIPCHK(hipMemcpyAsync(c->pin_out, c->out, bytes, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  memcpy(out, c->pin_out, bytes);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->last_ms = ms;
  c->pos += C;
  return ok();
}

int gf_convolver_reset(gf_convolver *c) {
  if (!c) return fail(GF_ERR_ARGUMENT, "geofdtd: null convolver");
  if (int rc = zero_ring(c)) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  return ok();
}

int gf_convolver_process_all(gf_convolver *c, const double *in, int64_t n, double *out) {
  if (!c) return fail(GF_ERR_ARGUMENT, "geofdtd: null convolver");
  if (n < 0 || (n > 0 && (!in || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad record arrays");
  if (n == 0) return ok();
  const int64_t C = c->chunk, L = c->len;
  const int64_t npad = (n + C - 1) / C * C;
  const int64_t ld = L - 1 + npad;  // linear buffer: len - 1 past samples, then the padded record
  if ((double)c->nrecv * ld > 4e9 || (double)c->nrecv * npad > 4e9)
    return fail(GF_ERR_ARGUMENT, "geofdtd: record too long for one ProcessAll call");
  if (int rc = grow(c, (size_t)c->nrecv * ld, (size_t)c->nrecv * npad)) return rc;
  // The linear buffer: past input x(pos - len + 1 .. pos - 1) from the ring,
  // then the record, zero-padded to whole chunks.
  std::vector<double> host((size_t)c->nrecv * ld, 0.0);
  if (L > 1 && c->pos > 0) {
    std::vector<double> ringh((size_t)c->nrecv * c->ring);
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(ringh.data(), c->x, sizeof(double) * ringh.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrecv; r++)
      for (int64_t k = 0; k < L - 1; k++) {
        const int64_t s = c->pos - (L - 1) + k;
        host[(size_t)r * ld + k] = s < 0 ? 0.0 : ringh[(size_t)r * c->ring + (s & (c->ring - 1))];
      }
  }
  for (int r = 0; r < c->nrecv; r++) memcpy(&host[(size_t)r * ld + (L - 1)], in + (size_t)r * n, sizeof(double) * n);
  HIPCHK(hipMemcpyAsync(c->in, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipEventRecord(c->e0, c->st));
  dim3 gf((unsigned)((npad + kOut - 1) / kOut), c->nrecv);
  k_fir<false><<<gf, kT, 0, c->st>>>(c->h, (int32_t)L, c->in, ld, c->ring, c->pos, npad, c->out, npad);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->e1, c->st));
  // The ring as after these chunks: the newest samples of the padded record.
  dim3 gw((unsigned)((std::min<int64_t>(npad, c->ring) + 255) / 256), c->nrecv);
  k_ring_write<<<gw, 256, 0, c->st>>>(c->x, c->ring, c->pos, c->in + (L - 1), npad, ld);
  HIPCHK(hipGetLastError());
  for (int r = 0; r < c->nrecv; r++)
    HIPCHK(hipMemcpyAsync(out + (size_t)r * n, c->out + (size_t)r * npad, sizeof(double) * n, hipMemcpyDeviceToHost,
                          c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->last_ms = ms;
  c->pos += npad;
  return ok();
}

int gf_convolver_stats(const gf_convolver *c, double *kernel_ms) {
  if (!c || !kernel_ms) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *kernel_ms = c->last_ms;
  return ok();
}

}  // extern "C"
