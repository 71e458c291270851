// bank_lookup.hip — bank.Bank's interpolation (internal/bank/interp.go of the
// reference) on gfx950: a bank resident in HBM and batched lookups.
//
// Layout: the payload [count][bins] float2 (range-major, as the file stores
// it) and two FP64 tables of the same shape, log|G| and the phase unwrapped
// along range (Bank.prepare, interp.go:64-101).  Zero values carry log|G| =
// -1e308 and stay zero on lookup.
//
//   k_bank_prepare  one thread per bin walks the range axis (the unwrapping is
//                   a running sum along range); consecutive threads read
//                   consecutive bins of one range row, so every step is one
//                   coalesced row read.
//   k_bank_points   one thread per (range, bin) query (Bank.Response).
//   k_bank_spectra  a workgroup per query range (more when the queries are
//                   few): the bracket of the query range, then for each bin
//                   two rows of each table read at consecutive bins and one
//                   complex value written.  Queries run in range order and
//                   each XCD takes one contiguous stretch of them, so a table
//                   row is fetched by about one L2 rather than by every one.
// A value reads 32 B of tables (mostly from L2 once queries run in range
// order) and writes 16 B; the writes bound it.
//
// Arithmetic is Go's: cmplx.Abs as math.Hypot, math.Round half away from
// zero, cmplx.Rect as r cos + i r sin; log and atan2 are the device
// library's, exp and sincos on lookup those of fastmath.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <rocfft/rocfft.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "model.h"
#include "fastmath.h"

namespace gfs {
int hip_fail(hipError_t e, const char *where);
int api_set_err(int code, const std::string &msg);
int check_device(std::string &why);
int bank_validate_header(const gf_bank_header &h, std::string &err);
}  // namespace gfs

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

struct gf_bank_dev {
  double r_min = 0, r_max = 0, rate = 0;
  int32_t count = 0, bins = 0;
  float2 *data = nullptr;
  double *logmag = nullptr, *phase = nullptr;
  double *q_r = nullptr;
  int32_t *q_bin = nullptr;
  double2 *out = nullptr;
  double2 *sensor = nullptr;  // a caller's sensor response per bin (impulse responses)
  size_t q_cap = 0, out_cap = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0, last_bytes = 0;
};

namespace {

constexpr double kNegInf = -1e308;
constexpr double kPi = 3.141592653589793;

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

__global__ void k_bank_prepare(int32_t count, int32_t bins, const float2 *__restrict__ data,
                               double *__restrict__ logmag, double *__restrict__ phase) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= bins) return;
  double prev = 0, unwrapped = 0;
  for (int i = 0; i < count; i++) {
    const int64_t idx = (int64_t)i * bins + k;
    const float2 v = data[idx];
    const double re = v.x, im = v.y;
    const double mag = hypot_go(re, im);
    if (mag == 0) {
      logmag[idx] = kNegInf;
      phase[idx] = unwrapped;
      continue;
    }
    logmag[idx] = log(mag);
    const double w = atan2(im, re);
    if (i == 0) {
      unwrapped = w;
    } else {
      double d = w - prev;
      d -= 2 * kPi * round(d / (2 * kPi));
      unwrapped += d;
    }
    prev = w;
    phase[idx] = unwrapped;
  }
}

// Bank.bracket (format.go:295-309) for a range already checked to lie in the
// bank: the lower row and the fraction towards the next.
struct Bracket {
  int lo;
  double frac;
};

__device__ __forceinline__ Bracket bracket(double r, double r_min, double r_max, int32_t count) {
  Bracket b{0, 0.0};
  if (count > 1) {
    const double spacing = (r_max - r_min) / (double)(count - 1);
    const double x = (r - r_min) / spacing;
    b.lo = (int)floor(x);
    if (b.lo >= count - 1) {
      b.lo = count - 2;
      b.frac = 1;
    } else {
      b.frac = x - (double)b.lo;
    }
  }
  return b;
}

// The interpolation of Bank.Response (interp.go:38-53) at one bin.
__device__ __forceinline__ double2 interpolate(const Bracket &b, int bin, int32_t count, int32_t bins,
                                               const double *__restrict__ logmag, const double *__restrict__ phase) {
  const int64_t i0 = (int64_t)b.lo * bins + bin;
  const int64_t i1 = count == 1 ? i0 : (int64_t)(b.lo + 1) * bins + bin;
  const double l0 = logmag[i0], l1 = logmag[i1];
  if (l0 == kNegInf || l1 == kNegInf) return make_double2(0, 0);
  const double p0 = phase[i0], p1 = phase[i1];
  const double lm = l0 + b.frac * (l1 - l0);
  const double ph = p0 + b.frac * (p1 - p0);
  const double m = gff::exp_reduced(lm);
  double s, c;
  gff::sincos_reduced(ph, &s, &c);
  return make_double2(m * c, m * s);
}

__device__ __forceinline__ double2 lookup(double r, int bin, double r_min, double r_max, int32_t count, int32_t bins,
                                          const double *__restrict__ logmag, const double *__restrict__ phase) {
  return interpolate(bracket(r, r_min, r_max, count), bin, count, bins, logmag, phase);
}

__global__ void k_bank_points(const double *__restrict__ r, const int32_t *__restrict__ bin, int64_t n,
                              double r_min, double r_max, int32_t count, int32_t bins,
                              const double *__restrict__ logmag, const double *__restrict__ phase,
                              double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = lookup(r[i], bin[i], r_min, r_max, count, bins, logmag, phase);
}

// Queries arrive sorted by range (slot q holds caller query dest[q]), so
// neighbouring workgroups read the same two table rows and each XCD's L2
// fetches a row pair once for all the queries that bracket into it.
// A thread walks the bins of its query with a stride of the workgroups
// along x times 256, so the bracket and the coefficients held in scalar
// registers are set up once for several values.
// Workgroups are dealt to the 8 XCDs round-robin by launch order.  The
// workgroup that launches b-th takes the b-th of each XCD's contiguous run of
// (bin slice, query) pairs, so an XCD serves one stretch of the range-sorted
// queries and its L2 fetches about 1/8 of the tables instead of all of them.
__device__ __forceinline__ int xcd_run(int b, int nb) {
  const int x = b & 7, j = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + j;
}

__global__ void k_bank_spectra(const double *__restrict__ r, const int32_t *__restrict__ dest, int32_t nq,
                               double r_min, double r_max, int32_t count, int32_t bins,
                               const double *__restrict__ logmag, const double *__restrict__ phase,
                               double2 *__restrict__ out) {
  const int nb = (int)(gridDim.x * gridDim.y);
  const int w = xcd_run((int)(blockIdx.x + blockIdx.y * gridDim.x), nb);
  const int q = w / (int)gridDim.x, bx = w % (int)gridDim.x;
  if (q >= nq) return;
  const Bracket b = bracket(r[q], r_min, r_max, count);
  double2 *row = out + (int64_t)dest[q] * bins;
  for (int k = bx * blockDim.x + threadIdx.x; k < bins; k += gridDim.x * blockDim.x)
    row[k] = interpolate(b, k, count, bins, logmag, phase);
}

// The spectra CausalImpulseResponseAt inverts (dsp.go:245-258) for the
// banked vertical velocity response (propagation.go:89-107): bin k at
// f = k fs / N is (i 2 pi f) G(r, k), times the sensor, tapered by a raised
// cosine above 0.8 of Nyquist; bin 0 is zero (f <= 0 answers 0).
__global__ void k_bank_impulse_coeff(const double *__restrict__ r, int32_t nq, double r_min, double r_max,
                                     int32_t count, int32_t bins, const double *__restrict__ logmag,
                                     const double *__restrict__ phase, const double2 *__restrict__ sensor, double fs,
                                     int32_t nfft, double bank_df, double2 *__restrict__ coeff) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (k >= bins || q >= nq) return;
  const double df = fs / (double)nfft;
  const double f = (double)k * df;
  // The bank bin this frequency lands on (checked on the host); past the
  // bank's last bin the response is 0 (propagation.go:96-98).
  const double bin = round(f / bank_df);
  double2 h = make_double2(0, 0);
  if (f > 0 && bin >= 0 && bin < bins) {
    const double2 u = lookup(r[q], (int)bin, r_min, r_max, count, bins, logmag, phase);
    const double w = 2 * kPi * f;
    h = make_double2(0.0 * u.x - w * u.y, 0.0 * u.y + w * u.x);  // complex(0, w) * u
  }
  if (sensor) {
    const double2 g = sensor[k];
    h = make_double2(h.x * g.x - h.y * g.y, h.x * g.y + h.y * g.x);
  }
  const double nyq = fs / 2, from = 0.8 * nyq;
  if (f > from) {
    const double t = 0.5 * (1 + cos(kPi * (f - from) / (nyq - from)));
    h = make_double2(h.x * t - h.y * 0.0, h.x * 0.0 + h.y * t);  // h * complex(t, 0)
  }
  coeff[(int64_t)q * bins + k] = h;
}

// dsp.IRFFT (dsp.go:42-50) of an nfft-point spectrum, first `length`
// samples: x_t = (1/N)(Re X_0 + Re X_{N/2} (-1)^t + 2 sum_k Re(X_k e^{2 pi i k t / N})),
// the phase k t mod N carried exactly in integers (t < N).  For N up to
// kTwiddleMax the workgroup first tabulates sincospi(2 j / N) for every j in
// LDS, the same values the direct form computes term by term, so each term
// is one table read and two multiply-adds.
constexpr int kTwiddleMax = 4096;  // 64 KB of LDS

__global__ void k_irdft_first(const double2 *__restrict__ coeff, int32_t bins, int32_t nq, int32_t nfft,
                              int32_t length, double *__restrict__ out) {
  extern __shared__ double2 tab[];  // [nfft] when nfft <= kTwiddleMax
  const bool use_tab = nfft <= kTwiddleMax;
  if (use_tab) {
    for (int j = threadIdx.x; j < nfft; j += blockDim.x) {
      double sn, cs;
      sincospi(2.0 * (double)j / (double)nfft, &sn, &cs);
      tab[j] = make_double2(sn, cs);
    }
    __syncthreads();
  }
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (t >= length || q >= nq) return;
  const double2 *X = coeff + (int64_t)q * bins;
  double s = X[0].x;
  int32_t ph = 0;  // k t mod nfft
  // Bins 1..N/2-1 in order, then the Nyquist bin N/2 (the last, bins - 1),
  // so the loop carries no test but its bound.
  const int kn = bins - 1;
  for (int k = 1; k < kn; k++) {
    ph += t;
    if (ph >= nfft) ph -= nfft;
    const double2 x = X[k];
    double sn, cs;
    if (use_tab) {
      const double2 w = tab[ph];
      sn = w.x;
      cs = w.y;
    } else {
      sincospi(2.0 * (double)ph / (double)nfft, &sn, &cs);
    }
    s += 2 * (x.x * cs - x.y * sn);
  }
  if (kn >= 1) s += (t & 1) ? -X[kn].x : X[kn].x;
  out[(int64_t)q * length + t] = s * (1 / (double)nfft);
}

// Inverse transforms longer than kTwiddleMax go through rocFFT (its
// double-precision real inverse transform, O(N log N)) instead of the
// direct sum: the imaginary parts of bins 0 and N/2 are cleared first (the
// direct sum and dsp.IRFFT ignore them), and the first `length` samples of
// each row are kept, scaled by 1/N.
__global__ void k_real_ends(double2 *__restrict__ coeff, int32_t bins, int32_t n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  coeff[(int64_t)j * bins].y = 0;
  coeff[(int64_t)j * bins + bins - 1].y = 0;
}

__global__ void k_take_first(const double *__restrict__ full, int32_t nfft, int32_t length, int32_t n, double scale,
                             double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (t >= length || j >= n) return;
  out[(int64_t)j * length + t] = full[(int64_t)j * nfft + t] * scale;
}

int fail(int code, const std::string &m) { return gfs::api_set_err(code, m); }
int ok() { return gfs::api_set_err(GF_OK, ""); }

// The Go error of Response(r, bin), or GF_OK (interp.go:30-36, format.go:297-299).
int check_query(const gf_bank_dev *b, double r, int32_t bin, std::string &err) {
  if (bin < 0 || bin >= b->bins) {
    err = "bank: bin " + std::to_string(bin) + " out of range [0, " + std::to_string(b->bins) + ")";
    return GF_ERR_INVALID;
  }
  if (r < b->r_min || r > b->r_max) {
    err = "bank: range " + gfm::go_g(r) + " m is outside the bank's " + gfm::go_g(b->r_min) + " to " +
          gfm::go_g(b->r_max) + " m";
    return GF_ERR_INVALID;
  }
  return GF_OK;
}

int reserve(gf_bank_dev *b, size_t nq, size_t nout) {
  if (b->q_cap < nq) {
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->q_r) HIPCHK(hipFree(b->q_r));
    if (b->q_bin) HIPCHK(hipFree(b->q_bin));
    b->q_r = nullptr;
    b->q_bin = nullptr;
    HIPCHK(hipMalloc(&b->q_r, sizeof(double) * nq));
    HIPCHK(hipMalloc(&b->q_bin, sizeof(int32_t) * nq));
    b->q_cap = nq;
  }
  if (b->out_cap < nout) {
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->out) HIPCHK(hipFree(b->out));
    b->out = nullptr;
    HIPCHK(hipMalloc(&b->out, sizeof(double2) * nout));
    b->out_cap = nout;
  }
  return GF_OK;
}

void release(gf_bank_dev *b) {
  if (b->st) hipStreamSynchronize(b->st);
  hipFree(b->data);
  hipFree(b->logmag);
  hipFree(b->phase);
  hipFree(b->q_r);
  hipFree(b->q_bin);
  hipFree(b->out);
  hipFree(b->sensor);
  if (b->e0) hipEventDestroy(b->e0);
  if (b->e1) hipEventDestroy(b->e1);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
}

}  // namespace

namespace gfbank {

hipError_t irdft_launch(double2 *d_coeff, int32_t bins, int32_t n, int32_t length, double *d_out, hipStream_t st);

int32_t bins(const gf_bank_dev *b) { return b->bins; }
double rate(const gf_bank_dev *b) { return b->rate; }

// CausalImpulseResponseAt's checks (dsp.go:233-244) with the bank's
// transform size, then, for each range in order, what the Go loop over bins
// 0..N/2 would reject first: the grid check of propagation.go:94-102 at the
// first f > 0 that lands inside the bank, and Bank.Response's range check
// there (the same for every bin).
int impulse_check(const gf_bank_dev *b, const double *r_m, int32_t n, int32_t length, double fs_hz) {
  const int32_t nfft = 2 * (b->bins - 1);
  if (length <= 0)
    return fail(GF_ERR_INVALID, "dsp: impulse response length must be positive, got " + std::to_string(length));
  if (fs_hz <= 0) return fail(GF_ERR_INVALID, "dsp: sample rate must be positive, got " + gfm::go_g(fs_hz));
  if (nfft < length)
    return fail(GF_ERR_INVALID, "dsp: transform size " + std::to_string(nfft) + " is shorter than the " +
                                    std::to_string(length) + " samples to keep");
  const double bank_df = b->rate / (double)nfft, df = fs_hz / (double)nfft;
  std::string err;
  for (int32_t q = 0; q < n; q++) {
    for (int32_t k = 1; k < b->bins; k++) {
      const double f = (double)k * df;
      const double x = f / bank_df;
      const double bin = round(x);
      if (bin < 0 || bin >= b->bins) continue;  // answered 0
      if (fabs(x - bin) > 1e-6)
        return fail(GF_ERR_INVALID, "sensing: " + gfm::go_g(f) + " Hz does not land on the bank's frequency grid "
                                    "(spacing " + gfm::go_g(bank_df) + " Hz)");
      if (int rc = check_query(b, r_m[q], (int32_t)bin, err)) return fail(rc, err);
      break;
    }
  }
  return GF_OK;
}

// The two kernels of an impulse-response build for n ranges already in HBM
// (d_r), on stream st: spectra into d_coeff [n][bins], the first `length`
// samples of their inverse transforms into d_out [n][length].
hipError_t impulse_launch(const gf_bank_dev *b, const double *d_r, int32_t n, int32_t length, double fs_hz,
                          const double2 *d_sensor, double2 *d_coeff, double *d_out, hipStream_t st) {
  const int32_t nfft = 2 * (b->bins - 1);
  const double bank_df = b->rate / (double)nfft;
  dim3 g1((b->bins + 255) / 256, n);
  k_bank_impulse_coeff<<<g1, 256, 0, st>>>(d_r, n, b->r_min, b->r_max, b->count, b->bins, b->logmag, b->phase,
                                           d_sensor, fs_hz, nfft, bank_df, d_coeff);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return irdft_launch(d_coeff, b->bins, n, length, d_out, st);
}

// The inverse transform on its own, for spectra built elsewhere (the
// analytic propagation of the sensing engine): coeff [n][bins] -> out
// [n][length], nfft = 2 (bins - 1).  coeff is scratch (the transform may
// overwrite it).  Up to kTwiddleMax points, the direct sum with its LDS
// table; beyond, rocFFT in batches of at most 256 MB of output.
hipError_t irdft_launch(double2 *d_coeff, int32_t bins, int32_t n, int32_t length, double *d_out, hipStream_t st) {
  const int32_t nfft = 2 * (bins - 1);
  if (nfft <= kTwiddleMax) {
    dim3 g2((length + 255) / 256, n);
    const size_t lds = sizeof(double2) * (size_t)nfft;
    k_irdft_first<<<g2, 256, lds, st>>>(d_coeff, bins, n, nfft, length, d_out);
    return hipGetLastError();
  }
  hipError_t e;
  k_real_ends<<<(n + 255) / 256, 256, 0, st>>>(d_coeff, bins, n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int32_t batch = (int32_t)std::max<int64_t>(1, std::min<int64_t>(n, (256LL << 20) / ((int64_t)nfft * 8)));
  static std::once_flag setup;
  std::call_once(setup, [] { rocfft_setup(); });
  double *full = nullptr;
  void *work = nullptr;
  size_t work_bytes = 0;
  if ((e = hipMalloc(&full, sizeof(double) * (size_t)batch * nfft)) != hipSuccess) return e;
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  int32_t planned = 0;
  const size_t len = (size_t)nfft;
  if (rocfft_execution_info_create(&info) != rocfft_status_success ||
      rocfft_execution_info_set_stream(info, st) != rocfft_status_success)
    e = hipErrorUnknown;
  for (int32_t j0 = 0; j0 < n && e == hipSuccess; j0 += batch) {
    const int32_t nb = std::min(batch, n - j0);
    if (nb != planned) {
      if (planned) rocfft_plan_destroy(plan);
      planned = 0;
      size_t need = 0;
      if (rocfft_plan_create(&plan, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                             rocfft_precision_double, 1, &len, (size_t)nb, nullptr) != rocfft_status_success ||
          rocfft_plan_get_work_buffer_size(plan, &need) != rocfft_status_success) {
        e = hipErrorUnknown;
        break;
      }
      planned = nb;
      if (need > work_bytes) {
        // the previous transform may still read the old work buffer
        if (work && (e = hipStreamSynchronize(st)) != hipSuccess) break;
        hipFree(work);
        work = nullptr;
        if ((e = hipMalloc(&work, need)) != hipSuccess) break;
        work_bytes = need;
      }
      if (need && rocfft_execution_info_set_work_buffer(info, work, work_bytes) != rocfft_status_success) {
        e = hipErrorUnknown;
        break;
      }
    }
    void *in[1] = {d_coeff + (size_t)j0 * bins};
    void *out[1] = {full};
    if (rocfft_execute(plan, in, out, info) != rocfft_status_success) {
      e = hipErrorUnknown;
      break;
    }
    dim3 g((length + 255) / 256, nb);
    k_take_first<<<g, 256, 0, st>>>(full, nfft, length, nb, 1 / (double)nfft, d_out + (size_t)j0 * length);
    e = hipGetLastError();
  }
  const hipError_t s2 = hipStreamSynchronize(st);
  if (planned) rocfft_plan_destroy(plan);
  if (info) rocfft_execution_info_destroy(info);
  hipFree(work);
  hipFree(full);
  return e != hipSuccess ? e : s2;
}

}  // namespace gfbank

extern "C" {

int gf_bank_dev_open(const gf_bank_header *h, const float *payload, gf_bank_dev **out) {
  if (!h || !payload || !out) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = nullptr;
  std::string err;
  int rc = gfs::bank_validate_header(*h, err);
  if (rc) return fail(rc, err);
  if ((rc = gfs::check_device(err))) return fail(rc, err);
  gf_bank_dev *b = new gf_bank_dev;
  b->r_min = h->r_min_m;
  b->r_max = h->r_max_m;
  b->rate = h->sample_rate_hz;
  b->count = h->count;
  b->bins = h->samples / 2 + 1;
  const size_t n = (size_t)b->count * b->bins;
  auto bail = [&](hipError_t e, const char *where) {
    release(b);
    return gfs::hip_fail(e, where);
  };
  hipError_t e;
  if ((e = hipStreamCreate(&b->st)) != hipSuccess) return bail(e, "hipStreamCreate");
  if ((e = hipEventCreate(&b->e0)) != hipSuccess || (e = hipEventCreate(&b->e1)) != hipSuccess)
    return bail(e, "hipEventCreate");
  if ((e = hipMalloc(&b->data, sizeof(float2) * n)) != hipSuccess ||
      (e = hipMalloc(&b->logmag, sizeof(double) * n)) != hipSuccess ||
      (e = hipMalloc(&b->phase, sizeof(double) * n)) != hipSuccess ||
      (e = hipMalloc(&b->sensor, sizeof(double2) * b->bins)) != hipSuccess)
    return bail(e, "hipMalloc");
  if ((e = hipMemcpyAsync(b->data, payload, sizeof(float2) * n, hipMemcpyHostToDevice, b->st)) != hipSuccess)
    return bail(e, "hipMemcpyAsync");
  k_bank_prepare<<<(b->bins + 255) / 256, 256, 0, b->st>>>(b->count, b->bins, b->data, b->logmag, b->phase);
  if ((e = hipGetLastError()) != hipSuccess) return bail(e, "k_bank_prepare");
  if ((e = hipStreamSynchronize(b->st)) != hipSuccess) return bail(e, "hipStreamSynchronize");
  *out = b;
  return ok();
}

void gf_bank_dev_close(gf_bank_dev *b) {
  if (b) release(b);
}

int gf_bank_dev_response(gf_bank_dev *b, const double *r_m, const int32_t *bin, int64_t n, double *out) {
  if (!b) return fail(GF_ERR_ARGUMENT, "geofdtd: null bank handle");
  if (n < 0 || (n > 0 && (!r_m || !bin || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad query arrays");
  std::string err;
  for (int64_t i = 0; i < n; i++)
    if (int rc = check_query(b, r_m[i], bin[i], err)) return fail(rc, err);
  if (n == 0) return ok();
  if (int rc = reserve(b, (size_t)n, (size_t)n)) return rc;
  HIPCHK(hipMemcpyAsync(b->q_r, r_m, sizeof(double) * n, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->q_bin, bin, sizeof(int32_t) * n, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipEventRecord(b->e0, b->st));
  k_bank_points<<<(unsigned)((n + 255) / 256), 256, 0, b->st>>>(b->q_r, b->q_bin, n, b->r_min, b->r_max, b->count,
                                                                 b->bins, b->logmag, b->phase, b->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->e1, b->st));
  HIPCHK(hipMemcpyAsync(out, b->out, sizeof(double2) * n, hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;
  b->last_bytes = (double)n * (32.0 + 16.0 + 12.0);
  return ok();
}

int gf_bank_dev_spectra(gf_bank_dev *b, const double *r_m, int32_t n, double *out) {
  if (!b) return fail(GF_ERR_ARGUMENT, "geofdtd: null bank handle");
  if (n < 0 || (n > 0 && (!r_m || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad query arrays");
  std::string err;
  for (int32_t i = 0; i < n; i++)
    if (int rc = check_query(b, r_m[i], 0, err)) return fail(rc, err);
  if (n == 0) return ok();
  if (n > 65535) return fail(GF_ERR_ARGUMENT, "geofdtd: at most 65535 ranges per spectra call");
  const size_t nout = (size_t)n * b->bins;
  if (int rc = reserve(b, (size_t)n, nout)) return rc;
  // Sorted by range, stable, with each query's output row.
  std::vector<int32_t> order((size_t)n);
  for (int32_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) { return r_m[a] < r_m[c]; });
  std::vector<double> sorted((size_t)n);
  for (int32_t i = 0; i < n; i++) sorted[i] = r_m[order[i]];
  HIPCHK(hipMemcpyAsync(b->q_r, sorted.data(), sizeof(double) * n, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->q_bin, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipEventRecord(b->e0, b->st));
  // Workgroups along the bins: enough for about 8,192 workgroups in all (32
  // waves on each of the 256 CUs), one when the queries alone provide them.
  const int32_t per_q = (b->bins + 255) / 256;
  dim3 grid((unsigned)std::max(1, std::min(per_q, (8192 + n - 1) / n)), n);
  k_bank_spectra<<<grid, 256, 0, b->st>>>(b->q_r, b->q_bin, n, b->r_min, b->r_max, b->count, b->bins, b->logmag,
                                          b->phase, b->out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->e1, b->st));
  HIPCHK(hipMemcpyAsync(out, b->out, sizeof(double2) * nout, hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;
  b->last_bytes = (double)nout * (32.0 + 16.0);
  return ok();
}

int gf_bank_dev_impulse(gf_bank_dev *b, const double *r_m, int32_t n, int32_t length, double fs_hz,
                        const double *sensor, double *out) {
  if (!b) return fail(GF_ERR_ARGUMENT, "geofdtd: null bank handle");
  if (n < 0 || (n > 0 && (!r_m || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad query arrays");
  if (n > 65535) return fail(GF_ERR_ARGUMENT, "geofdtd: at most 65535 ranges per impulse call");
  if (int rc = gfbank::impulse_check(b, r_m, n, length, fs_hz)) return rc;
  if (n == 0) return ok();
  const size_t ncoef = (size_t)n * b->bins;
  if (int rc = reserve(b, (size_t)n, ncoef + ((size_t)n * length + 1) / 2 + 1)) return rc;
  double2 *coeff = b->out;
  double *trace = (double *)(b->out + ncoef);
  double2 *d_sensor = nullptr;
  if (sensor) {
    d_sensor = b->sensor;
    HIPCHK(hipMemcpyAsync(d_sensor, sensor, sizeof(double2) * b->bins, hipMemcpyHostToDevice, b->st));
  }
  HIPCHK(hipMemcpyAsync(b->q_r, r_m, sizeof(double) * n, hipMemcpyHostToDevice, b->st));
  HIPCHK(hipEventRecord(b->e0, b->st));
  HIPCHK(gfbank::impulse_launch(b, b->q_r, n, length, fs_hz, d_sensor, coeff, trace, b->st));
  HIPCHK(hipEventRecord(b->e1, b->st));
  HIPCHK(hipMemcpyAsync(out, trace, sizeof(double) * (size_t)n * length, hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->last_ms = ms;
  b->last_bytes = (double)ncoef * 48.0 + (double)n * length * 8.0;
  return ok();
}

int gf_bank_dev_stats(const gf_bank_dev *b, double *kernel_ms, double *bytes) {
  if (!b || !kernel_ms || !bytes) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *kernel_ms = b->last_ms;
  *bytes = b->last_bytes;
  return ok();
}

double gf_range_nyquist(double slowest, double max_freq) {
  if (max_freq <= 0) return INFINITY;
  return slowest / (2 * max_freq);
}

int gf_bank_check_range_sampling(const gf_bank_header *h, double max_freq) {
  if (!h) return fail(GF_ERR_ARGUMENT, "geofdtd: null header");
  if (h->count <= 1 || h->n_medium <= 0 || !h->medium) return ok();
  double slowest = h->medium[0].vs_mps;
  for (int i = 0; i < h->n_medium; i++) slowest = fmin(slowest, h->medium[i].vs_mps);
  const double limit = gf_range_nyquist(0.87 * slowest, max_freq) / 2;
  const double got = (h->r_max_m - h->r_min_m) / (double)(h->count - 1);
  if (got > limit) {
    char b1[64], b2[64];
    snprintf(b1, sizeof b1, "%.4f", got);
    if (std::isinf(limit))
      snprintf(b2, sizeof b2, "+Inf");
    else
      snprintf(b2, sizeof b2, "%.4f", limit);
    return fail(GF_ERR_INVALID, std::string("bank: range spacing ") + b1 + " m is too coarse to unwrap phase up to " +
                                    gfm::go_g(max_freq) + " Hz; need under " + b2 + " m");
  }
  return ok();
}

}  // extern "C"
