// engine.hip — sensing.Engine (internal/sensing/sensing.go of the reference)
// driven by a bank (sensing.FromBank), for many receivers at once, on gfx950.
//
// The reference gives every contact (a footfall) a voice: a streaming
// convolver of the contact's force with the impulse response at its range,
// kept until the response's coda has run out, and sums the voices a chunk at a
// time.  With a bank, the radial-force response is zero (propagation.go:
// 110-112), so each voice's second convolver contributes exactly nothing and
// only the vertical one is kept here.
//
// Device state:
//   impulse responses  one per range quantum (RangeQuantum = 0.02 m, shared by
//                      every receiver and contact that rounds to it), built in
//                      batches on the device from the bank's tables
//                      (FromBank's i w G(r, f), times the sensor, through
//                      CausalImpulseResponseAt) and kept for the engine's life.
//   voices             up to kMaxVoices slots: the contact's vertical force
//                      sampled from its admission chunk on, and per receiver
//                      the address of its impulse response.
// A voice admitted at sample a with force f[0..n) contributes
//     y(t) = sum_{m} f[m] h[t - a - m]
// at sample t, which is what the reference's convolver (started at the
// admission chunk, zero state) produces.  k_engine_mix computes every
// receiver's chunk in one launch: one 1024-thread workgroup per (receiver, up
// to 64 outputs), lanes split over outputs and tap slices, the slices added
// in a fixed order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "model.h"

namespace gfs {
int hip_fail(hipError_t e, const char *where);
int api_set_err(int code, const std::string &msg);
int check_device(std::string &why);
}  // namespace gfs

namespace gfbank {
int32_t bins(const gf_bank_dev *b);
double rate(const gf_bank_dev *b);
int impulse_check(const gf_bank_dev *b, const double *r_m, int32_t n, int32_t length, double fs_hz);
hipError_t impulse_launch(const gf_bank_dev *b, const double *d_r, int32_t n, int32_t length, double fs_hz,
                          const double2 *d_sensor, double2 *d_coeff, double *d_out, hipStream_t st);
hipError_t irdft_launch(double2 *d_coeff, int32_t bins, int32_t n, int32_t length, double *d_out, hipStream_t st);
}  // namespace gfbank

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

namespace {

constexpr int kMaxVoices = 64;
constexpr double kRangeQuantum = 0.02;  // sensing.go:49

struct MixArgs {
  int32_t nv;
  int32_t slot[kMaxVoices];
  int32_t n[kMaxVoices];
  int64_t first[kMaxVoices];
};

// One workgroup of 1024 threads per (receiver, np outputs): thread = (output
// t, slice q) with S = 1024 / np slices striding every active voice's taps in
// turn; np is chosen on the host so that the grid fills the chip (64
// receivers x 20 samples: 5 outputs a workgroup, 204 lanes an output).  The
// chain is latency-bound (two loads per multiply-add, the responses in L2 or
// the last-level cache), so lanes, not arithmetic, set its time.  The slices'
// sums are added in a fixed order: groups of 32, then the groups.
constexpr int kMixThreads = 1024;

template <bool RADIAL>
__global__ __launch_bounds__(kMixThreads) void k_engine_mix(MixArgs a, const double *__restrict__ force, int32_t cap,
                                                            const double *const *__restrict__ irp,
                                                            const double *const *__restrict__ irr,
                                                            const double *__restrict__ uv, int32_t nrecv,
                                                            int32_t irlen, int64_t emitted, int32_t chunk,
                                                            int32_t np, float *__restrict__ out) {
  __shared__ double red[kMixThreads];
  __shared__ double grp[kMixThreads / 32 + 64];
  const int r = blockIdx.x;
  const int t = threadIdx.x % np, q = threadIdx.x / np;
  const int S = kMixThreads / np;
  const int i = blockIdx.y * np + t;
  double acc = 0;
  if (i < chunk && q < S) {
    for (int v = 0; v < a.nv; v++) {
      const int s = a.slot[v];
      const int64_t T = emitted + i - a.first[v];
      const int64_t lo = T - irlen + 1 > 0 ? T - irlen + 1 : 0;
      const int64_t hi = T < a.n[v] - 1 ? T : a.n[v] - 1;
      // the response table holds device addresses: read them as global
      // memory, not flat
      const double *f = force + (int64_t)s * (RADIAL ? 3 : 1) * cap;
      const __attribute__((address_space(1))) double *h =
          (const __attribute__((address_space(1))) double *)irp[(int64_t)s * nrecv + r];
      if (RADIAL) {
        // the force along the contact-to-receiver line, f_x u_x + f_y u_y,
        // through the radial-force response (sensing.go:264-270)
        const __attribute__((address_space(1))) double *hr =
            (const __attribute__((address_space(1))) double *)irr[(int64_t)s * nrecv + r];
        const double ux = uv[2 * ((int64_t)s * nrecv + r)], uy = uv[2 * ((int64_t)s * nrecv + r) + 1];
        const double *fx = f + cap, *fy = f + 2 * cap;
#pragma unroll 4
        for (int64_t m = lo + q; m <= hi; m += S) {
          acc = fma(f[m], h[T - m], acc);
          acc = fma(fx[m] * ux + fy[m] * uy, hr[T - m], acc);
        }
      } else {
#pragma unroll 8
        for (int64_t m = lo + q; m <= hi; m += S) acc = fma(f[m], h[T - m], acc);
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  const int G = (S + 31) / 32;
  if (q < G && i < chunk) {
    double g = 0;
    for (int k = q * 32; k < q * 32 + 32 && k < S; k++) g += red[k * np + t];
    grp[q * np + t] = g;
  }
  __syncthreads();
  if (q == 0 && i < chunk) {
    double sum = 0;
    for (int k = 0; k < G; k++) sum += grp[k * np + t];
    out[(int64_t)r * chunk + i] = (float)sum;
  }
}

// sensing.Analytic (propagation.go:38-62) over green.HalfSpaceGF
// (green.go:120-266, 375-388): at bin k (f = k fs / nfft) and range rq[j],
// the vertical velocity response and the radial-force response, each times
// the sensor and the CausalImpulseResponseAt taper (dsp.go:245-258), with the
// Go operation order: Kjartansson's constant-Q wavenumber k(w) =
// (w / c0) (f / f0)^-g e^{-i pi g / 2}, phase velocity c = cR (f / f0)^g, the
// Rayleigh mode shapes at w / cR, the single-mode excitation
// uz^2 / (8 c u I1) sqrt(2 / (pi Re k r)) and e^{-i pi/4 - i k r}, times i w;
// the radial response is i (H(0) / V(0)) times the vertical one.
struct Halfspace {
  double vp, vs, rho, cr, ref, gam, c0, ere, eim;
};

__device__ __forceinline__ double2 cmulz(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ void k_analytic_spectra(Halfspace H, const double *__restrict__ rq, int32_t nq, int32_t bins,
                                   int32_t nfft, double fs, const double2 *__restrict__ sensor,
                                   double2 *__restrict__ cz, double2 *__restrict__ crad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= bins || j >= nq) return;
  const double f = (double)k * (fs / (double)nfft);
  double2 hz = make_double2(0, 0), hr = make_double2(0, 0);
  if (f > 0) {
    const double r = rq[j];
    const double omega = 2 * 3.141592653589793 * f;
    // ComplexWavenumber and PhaseVelocity (green.go:169-198)
    const double mag = omega / H.c0 * pow(f / H.ref, -H.gam);
    const double2 kc = make_double2(mag * H.ere - 0.0 * H.eim, mag * H.eim + 0.0 * H.ere);
    const double c = H.cr * pow(f / H.ref, H.gam);
    // NewEigenfunctions (green.go:58-80), EnergyIntegral (:112-117)
    const double kk = omega / H.cr;
    const double qq = kk * sqrt(1 - (H.cr * H.cr) / (H.vp * H.vp));
    const double ss = kk * sqrt(1 - (H.cr * H.cr) / (H.vs * H.vs));
    const double k2 = kk * kk, s2 = ss * ss;
    const double xq = kk, xs = -2 * kk * qq * ss / (k2 + s2);
    const double zq = -qq, zs = 2 * k2 * qq / (k2 + s2);
    auto sq = [&](double a, double b) { return a * a / (2 * qq) + 2 * a * b / (qq + ss) + b * b / (2 * ss); };
    const double i1 = 0.5 * H.rho * (sq(xq, xs) + sq(zq, zs));
    const double uz = zq + zs, ux = xq + xs;
    // VelocityResponse (green.go:218-266)
    const double kre = kc.x;
    const double amp = uz * uz / (8 * c * c * i1) * sqrt(2 / (3.141592653589793 * kre * r));
    const double px = kc.y * r, py = -3.141592653589793 / 4 - kc.x * r;  // -i pi/4 - i kc r
    const double ex = exp(px);
    double sn, cs;
    sincos(py, &sn, &cs);
    const double2 prop = make_double2(ex * cs, ex * sn);
    const double2 iwa = make_double2(0.0 * amp - omega * 0.0, 0.0 * 0.0 + omega * amp);
    hz = cmulz(iwa, prop);
    // RadialForceResponse (green.go:375-388)
    if (hz.x != 0 || hz.y != 0) {
      const double ratio = ux / uz;
      hr = cmulz(make_double2(0.0 * ratio - 1.0 * 0.0, 0.0 * 0.0 + 1.0 * ratio), hz);
    }
  }
  if (sensor) {
    hz = cmulz(hz, sensor[k]);
    hr = cmulz(hr, sensor[k]);
  }
  const double nyq = fs / 2, from = 0.8 * nyq;
  if (f > from) {
    const double w = 0.5 * (1 + cos(3.141592653589793 * (f - from) / (nyq - from)));
    hz = make_double2(hz.x * w - hz.y * 0.0, hz.x * 0.0 + hz.y * w);
    hr = make_double2(hr.x * w - hr.y * 0.0, hr.x * 0.0 + hr.y * w);
  }
  cz[(int64_t)j * bins + k] = hz;
  crad[(int64_t)j * bins + k] = hr;
}

int fail(int code, const std::string &m) { return gfs::api_set_err(code, m); }
int ok() { return gfs::api_set_err(GF_OK, ""); }

// math.Hypot (Go): scale by the larger magnitude.
double go_hypot(double p, double q) {
  p = fabs(p);
  q = fabs(q);
  if (std::isinf(p) || std::isinf(q)) return INFINITY;
  if (std::isnan(p) || std::isnan(q)) return NAN;
  if (p < q) std::swap(p, q);
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

// strconv.Quote for the printable ASCII a contact ID holds.
std::string go_quote(const std::string &s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}

}  // namespace

struct gf_engine {
  const gf_bank_dev *bank = nullptr;
  int32_t chunk = 0, irlen = 0, nrecv = 0, np = 0, cap = 0, bins = 0;
  double fs = 0;
  std::vector<double> rx, ry;
  int64_t emitted = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double2 *d_sensor = nullptr;
  std::map<int64_t, double *> cache;  // range quantum -> impulse response in HBM
  std::vector<double *> blocks;       // the allocations behind the cache
  double *d_r = nullptr;
  double2 *d_coeff = nullptr;
  size_t build_cap = 0;
  struct Voice {
    int32_t slot, n;
    int64_t first, silent;
  };
  std::vector<Voice> voices;  // in admission order
  std::vector<int32_t> free_slots;
  double *d_force = nullptr;        // [kMaxVoices][cap]
  const double **d_irp = nullptr;   // [kMaxVoices][nrecv]
  float *d_out = nullptr, *pin_out = nullptr;
  // pinned staging per voice slot, so an admission's copies run
  // asynchronously: a slot is reused only after a later Next has synchronised
  double *pin_force = nullptr;       // [kMaxVoices][cap]
  const double **pin_irp = nullptr;  // [kMaxVoices][nrecv]
  double last_ms = 0, total_ms = 0, build_ms = 0;
  int64_t builds = 0, macs = 0;
  // analytic propagation (no bank): both components, spectra built here
  bool radial = false;
  int32_t nfft = 0;
  Halfspace hs{};
  std::map<int64_t, double *> cache_r;  // range quantum -> radial-force response
  const double **d_irr = nullptr, **pin_irr = nullptr;  // [kMaxVoices][nrecv]
  double *d_uv = nullptr, *pin_uv = nullptr;            // [kMaxVoices][nrecv][2]
};

namespace {

void release(gf_engine *e) {
  if (e->st) hipStreamSynchronize(e->st);
  for (double *p : e->blocks) hipFree(p);
  hipFree(e->d_sensor);
  hipFree(e->d_r);
  hipFree(e->d_coeff);
  hipFree(e->d_force);
  hipFree(e->d_irp);
  hipFree(e->d_out);
  if (e->pin_out) hipHostFree(e->pin_out);
  if (e->pin_force) hipHostFree(e->pin_force);
  if (e->pin_irp) hipHostFree(e->pin_irp);
  hipFree(e->d_irr);
  hipFree(e->d_uv);
  if (e->pin_irr) hipHostFree(e->pin_irr);
  if (e->pin_uv) hipHostFree(e->pin_uv);
  if (e->e0) hipEventDestroy(e->e0);
  if (e->e1) hipEventDestroy(e->e1);
  if (e->st) hipStreamDestroy(e->st);
  delete e;
}

int64_t quantum(double r) { return (int64_t)round(r / kRangeQuantum); }

// responseFor (sensing.go:186-222) for every (contact, receiver) pair: the
// quanta not yet cached, in order of first use, checked as the Go calls
// would be and then built in one batch.
int build_responses(gf_engine *e, const double *xy, int32_t n) {
  std::vector<int64_t> keys;
  std::vector<double> rq;
  for (int32_t c = 0; c < n; c++)
    for (int32_t q = 0; q < e->nrecv; q++) {
      const int64_t k = quantum(go_hypot(e->rx[q] - xy[2 * c], e->ry[q] - xy[2 * c + 1]));
      if (e->cache.count(k) || std::find(keys.begin(), keys.end(), k) != keys.end()) continue;
      keys.push_back(k);
      rq.push_back(std::max((double)k * kRangeQuantum, kRangeQuantum));
    }
  if (keys.empty()) return GF_OK;
  const int32_t nk = (int32_t)keys.size();
  if (!e->radial)
    if (int rc = gfbank::impulse_check(e->bank, rq.data(), nk, e->irlen, e->fs)) return rc;
  const int ncomp = e->radial ? 2 : 1;
  if (e->build_cap < (size_t)nk) {
    HIPCHK(hipStreamSynchronize(e->st));
    hipFree(e->d_r);
    hipFree(e->d_coeff);
    e->d_r = nullptr;
    e->d_coeff = nullptr;
    HIPCHK(hipMalloc(&e->d_r, sizeof(double) * nk));
    HIPCHK(hipMalloc(&e->d_coeff, sizeof(double2) * (size_t)ncomp * nk * e->bins));
    e->build_cap = nk;
  }
  double *block = nullptr;
  HIPCHK(hipMalloc(&block, sizeof(double) * (size_t)ncomp * nk * e->irlen));
  e->blocks.push_back(block);
  HIPCHK(hipMemcpyAsync(e->d_r, rq.data(), sizeof(double) * nk, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipEventRecord(e->e0, e->st));
  if (e->radial) {
    // spectra of both components, then their inverse transforms: vertical
    // responses first in the block, radial ones after
    double2 *cz = e->d_coeff, *cr = e->d_coeff + (size_t)nk * e->bins;
    dim3 g((e->bins + 255) / 256, nk);
    k_analytic_spectra<<<g, 256, 0, e->st>>>(e->hs, e->d_r, nk, e->bins, e->nfft, e->fs, e->d_sensor, cz, cr);
    HIPCHK(hipGetLastError());
    HIPCHK(gfbank::irdft_launch(e->d_coeff, e->bins, 2 * nk, e->irlen, block, e->st));
  } else {
    HIPCHK(gfbank::impulse_launch(e->bank, e->d_r, nk, e->irlen, e->fs, e->d_sensor, e->d_coeff, block, e->st));
  }
  HIPCHK(hipEventRecord(e->e1, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e->e0, e->e1));
  e->build_ms += ms;
  e->builds += nk;
  for (int32_t j = 0; j < nk; j++) {
    e->cache[keys[j]] = block + (size_t)j * e->irlen;
    if (e->radial) e->cache_r[keys[j]] = block + (size_t)(nk + j) * e->irlen;
  }
  return GF_OK;
}

// soil.HalfSpace.Validate (soil.go:72-89), the Go texts.
int halfspace_validate(const gf_halfspace &h) {
  using gfm::go_g;
  if (h.vs_mps <= 0) return fail(GF_ERR_INVALID, "soil: Vs must be positive, got " + go_g(h.vs_mps) + " m/s");
  if (h.vp_mps <= 0) return fail(GF_ERR_INVALID, "soil: Vp must be positive, got " + go_g(h.vp_mps) + " m/s");
  if (h.density_kgm3 <= 0)
    return fail(GF_ERR_INVALID, "soil: density must be positive, got " + go_g(h.density_kgm3) + " kg/m^3");
  if (h.vp_mps <= sqrt(2.0) * h.vs_mps)
    return fail(GF_ERR_INVALID, "soil: Vp=" + go_g(h.vp_mps) + " m/s must exceed sqrt(2)*Vs=" +
                                    go_g(sqrt(2.0) * h.vs_mps) + " m/s for a non-negative Poisson ratio");
  if (h.qs <= 0) return fail(GF_ERR_INVALID, "soil: Qs must be positive, got " + go_g(h.qs));
  if (h.qp < 0) return fail(GF_ERR_INVALID, "soil: Qp must not be negative, got " + go_g(h.qp));
  return GF_OK;
}

// soil.HalfSpace.RayleighVelocity (soil.go:136-163): bisection on
// (2 - x)^2 - 4 sqrt(1 - x) sqrt(1 - x k^2), x = (c / Vs)^2, k = Vs / Vp.
int rayleigh_velocity(const gf_halfspace &h, double *out) {
  if (int rc = halfspace_validate(h)) return rc;
  const double k = h.vs_mps / h.vp_mps, k2 = k * k;
  auto f = [&](double x) { return (2 - x) * (2 - x) - 4 * sqrt(1 - x) * sqrt(1 - x * k2); };
  double lo = 1e-9, hi = 1.0;
  if (f(lo) >= 0 || f(hi) <= 0)
    return fail(GF_ERR_INVALID, "soil: no Rayleigh root bracketed for Vp=" + gfm::go_g(h.vp_mps) +
                                    " Vs=" + gfm::go_g(h.vs_mps));
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (f(mid) < 0)
      lo = mid;
    else
      hi = mid;
  }
  *out = sqrt(0.5 * (lo + hi)) * h.vs_mps;
  return GF_OK;
}

}  // namespace

extern "C" {

int gf_engine_create(const gf_bank_dev *bank, int32_t n_recv, const double *recv_xy, int32_t chunk, double fs_hz,
                     int32_t ir_len, const double *sensor, int32_t max_contact_samples, gf_engine **out) {
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = nullptr;
  if (chunk <= 0) return fail(GF_ERR_INVALID, "sensing: chunk size must be positive, got " + std::to_string(chunk));
  if (!bank) return fail(GF_ERR_INVALID, "sensing: bank is nil");
  if (n_recv <= 0 || !recv_xy) return fail(GF_ERR_ARGUMENT, "geofdtd: at least one receiver position is required");
  if (ir_len <= 0) return fail(GF_ERR_ARGUMENT, "geofdtd: impulse response length must be positive");
  if (max_contact_samples <= 0) return fail(GF_ERR_ARGUMENT, "geofdtd: max_contact_samples must be positive");
  std::string err;
  if (int rc = gfs::check_device(err)) return fail(rc, err);
  gf_engine *e = new gf_engine;
  e->bank = bank;
  e->chunk = chunk;
  e->fs = fs_hz;
  e->bins = gfbank::bins(bank);
  // A bank fixes the transform: keep at most three quarters of it
  // (sensing.go:141-148).
  const int32_t n = 2 * (e->bins - 1);
  e->irlen = ir_len > 3 * n / 4 ? 3 * n / 4 : ir_len;
  e->nrecv = n_recv;
  e->cap = max_contact_samples;
  // Outputs per workgroup: few enough that the grid covers the 256 CUs
  // (each output's taps then spread over up to 1024 lanes), at most 64.
  {
    const int64_t want = ((int64_t)chunk * n_recv + 255) / 256;
    e->np = (int32_t)std::max<int64_t>(1, std::min<int64_t>({want, (int64_t)chunk, 64}));
  }
  for (int32_t q = 0; q < n_recv; q++) {
    e->rx.push_back(recv_xy[2 * q]);
    e->ry.push_back(recv_xy[2 * q + 1]);
  }
  for (int32_t s = kMaxVoices - 1; s >= 0; s--) e->free_slots.push_back(s);
  auto bail = [&](hipError_t he, const char *where) {
    release(e);
    return gfs::hip_fail(he, where);
  };
  hipError_t he;
  if ((he = hipStreamCreate(&e->st)) != hipSuccess) return bail(he, "hipStreamCreate");
  if ((he = hipEventCreate(&e->e0)) != hipSuccess || (he = hipEventCreate(&e->e1)) != hipSuccess)
    return bail(he, "hipEventCreate");
  if ((he = hipMalloc(&e->d_force, sizeof(double) * (size_t)kMaxVoices * e->cap)) != hipSuccess ||
      (he = hipMalloc(&e->d_irp, sizeof(double *) * (size_t)kMaxVoices * n_recv)) != hipSuccess ||
      (he = hipMalloc(&e->d_out, sizeof(float) * (size_t)n_recv * chunk)) != hipSuccess)
    return bail(he, "hipMalloc");
  if ((he = hipHostMalloc((void **)&e->pin_out, sizeof(float) * (size_t)n_recv * chunk, hipHostMallocDefault)) !=
          hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_force, sizeof(double) * (size_t)kMaxVoices * e->cap,
                          hipHostMallocDefault)) != hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_irp, sizeof(double *) * (size_t)kMaxVoices * n_recv,
                          hipHostMallocDefault)) != hipSuccess)
    return bail(he, "hipHostMalloc");
  if (sensor) {
    if ((he = hipMalloc(&e->d_sensor, sizeof(double2) * e->bins)) != hipSuccess) return bail(he, "hipMalloc");
    if ((he = hipMemcpy(e->d_sensor, sensor, sizeof(double2) * e->bins, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(he, "hipMemcpy");
  }
  *out = e;
  return ok();
}

void gf_engine_destroy(gf_engine *e) {
  if (e) release(e);
}

int gf_rayleigh_velocity(const gf_halfspace *h, double *out_mps) {
  if (!h || !out_mps) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  if (int rc = rayleigh_velocity(*h, out_mps)) return rc;
  return ok();
}

int gf_engine_create_analytic(const gf_halfspace *h, double ref_freq_hz, int32_t n_recv, const double *recv_xy,
                              int32_t chunk, double fs_hz, int32_t ir_len, const double *sensor,
                              int32_t max_contact_samples, gf_engine **out) {
  if (!out || !h) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = nullptr;
  if (chunk <= 0) return fail(GF_ERR_INVALID, "sensing: chunk size must be positive, got " + std::to_string(chunk));
  if (n_recv <= 0 || !recv_xy) return fail(GF_ERR_ARGUMENT, "geofdtd: at least one receiver position is required");
  if (ir_len <= 0)
    return fail(GF_ERR_INVALID, "dsp: impulse response length must be positive, got " + std::to_string(ir_len));
  if (fs_hz <= 0) return fail(GF_ERR_INVALID, "dsp: sample rate must be positive, got " + gfm::go_g(fs_hz));
  if (max_contact_samples <= 0) return fail(GF_ERR_ARGUMENT, "geofdtd: max_contact_samples must be positive");
  double cr = 0;
  if (int rc = rayleigh_velocity(*h, &cr)) return rc;
  std::string err;
  if (int rc = gfs::check_device(err)) return fail(rc, err);
  gf_engine *e = new gf_engine;
  e->radial = true;
  e->chunk = chunk;
  e->fs = fs_hz;
  e->irlen = ir_len;
  int32_t n = 1;
  while (n < 4 * ir_len) n <<= 1;  // dsp.NextPow2(4 length)
  e->nfft = n;
  e->bins = n / 2 + 1;
  // HalfSpaceGF's constants (green.go:132-198)
  const double ref = ref_freq_hz > 0 ? ref_freq_hz : 30.0;
  const double gam = atan(1 / h->qs) / M_PI;
  e->hs.vp = h->vp_mps;
  e->hs.vs = h->vs_mps;
  e->hs.rho = h->density_kgm3;
  e->hs.cr = cr;
  e->hs.ref = ref;
  e->hs.gam = gam;
  e->hs.c0 = cr * cos(M_PI * gam / 2);
  e->hs.ere = cos(-M_PI * gam / 2);
  e->hs.eim = sin(-M_PI * gam / 2);
  e->nrecv = n_recv;
  e->cap = max_contact_samples;
  {
    const int64_t want = ((int64_t)chunk * n_recv + 255) / 256;
    e->np = (int32_t)std::max<int64_t>(1, std::min<int64_t>({want, (int64_t)chunk, 64}));
  }
  for (int32_t q = 0; q < n_recv; q++) {
    e->rx.push_back(recv_xy[2 * q]);
    e->ry.push_back(recv_xy[2 * q + 1]);
  }
  for (int32_t s = kMaxVoices - 1; s >= 0; s--) e->free_slots.push_back(s);
  auto bail = [&](hipError_t he, const char *where) {
    release(e);
    return gfs::hip_fail(he, where);
  };
  hipError_t he;
  const size_t vr = (size_t)kMaxVoices * n_recv;
  if ((he = hipStreamCreate(&e->st)) != hipSuccess) return bail(he, "hipStreamCreate");
  if ((he = hipEventCreate(&e->e0)) != hipSuccess || (he = hipEventCreate(&e->e1)) != hipSuccess)
    return bail(he, "hipEventCreate");
  if ((he = hipMalloc(&e->d_force, sizeof(double) * 3 * (size_t)kMaxVoices * e->cap)) != hipSuccess ||
      (he = hipMalloc(&e->d_irp, sizeof(double *) * vr)) != hipSuccess ||
      (he = hipMalloc(&e->d_irr, sizeof(double *) * vr)) != hipSuccess ||
      (he = hipMalloc(&e->d_uv, sizeof(double) * 2 * vr)) != hipSuccess ||
      (he = hipMalloc(&e->d_out, sizeof(float) * (size_t)n_recv * chunk)) != hipSuccess)
    return bail(he, "hipMalloc");
  if ((he = hipHostMalloc((void **)&e->pin_out, sizeof(float) * (size_t)n_recv * chunk, hipHostMallocDefault)) !=
          hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_force, sizeof(double) * 3 * (size_t)kMaxVoices * e->cap,
                          hipHostMallocDefault)) != hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_irp, sizeof(double *) * vr, hipHostMallocDefault)) != hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_irr, sizeof(double *) * vr, hipHostMallocDefault)) != hipSuccess ||
      (he = hipHostMalloc((void **)&e->pin_uv, sizeof(double) * 2 * vr, hipHostMallocDefault)) != hipSuccess)
    return bail(he, "hipHostMalloc");
  if (sensor) {
    if ((he = hipMalloc(&e->d_sensor, sizeof(double2) * e->bins)) != hipSuccess) return bail(he, "hipMalloc");
    if ((he = hipMemcpy(e->d_sensor, sensor, sizeof(double2) * e->bins, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(he, "hipMemcpy");
  }
  *out = e;
  return ok();
}

int gf_engine_transform(const gf_engine *e, int32_t *nfft) {
  if (!e || !nfft) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *nfft = e->radial ? e->nfft : 2 * (e->bins - 1);
  return ok();
}

int gf_engine_admit_xyz(gf_engine *e, const char *id, double x, double y, const double *fx, const double *fy,
                        const double *fz, int32_t n, double end_s) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (!e->radial) return fail(GF_ERR_ARGUMENT, "geofdtd: a bank engine takes the vertical force (gf_engine_admit)");
  if (n <= 0 || !fx || !fy || !fz) return fail(GF_ERR_ARGUMENT, "geofdtd: a voice needs its force samples");
  std::vector<double> r((size_t)e->nrecv);
  for (int32_t q = 0; q < e->nrecv; q++) {
    r[q] = go_hypot(e->rx[q] - x, e->ry[q] - y);
    if (r[q] <= 0)
      return fail(GF_ERR_INVALID, "sensing: contact " + go_quote(id ? id : "") + " at the receiver's own position");
  }
  if (n > e->cap)
    return fail(GF_ERR_ARGUMENT, "geofdtd: a contact of " + std::to_string(n) + " samples exceeds the engine's " +
                                     std::to_string(e->cap));
  if (e->free_slots.empty())
    return fail(GF_ERR_ARGUMENT, "geofdtd: more than " + std::to_string(kMaxVoices) + " voices at once");
  const double xy[2] = {x, y};
  if (int rc = build_responses(e, xy, 1)) return rc;
  const int32_t slot = e->free_slots.back();
  double *pf = e->pin_force + (size_t)slot * 3 * e->cap;
  memcpy(pf, fz, sizeof(double) * n);
  memcpy(pf + e->cap, fx, sizeof(double) * n);
  memcpy(pf + 2 * e->cap, fy, sizeof(double) * n);
  const double **pz = e->pin_irp + (size_t)slot * e->nrecv;
  const double **pr = e->pin_irr + (size_t)slot * e->nrecv;
  double *pu = e->pin_uv + (size_t)slot * e->nrecv * 2;
  for (int32_t q = 0; q < e->nrecv; q++) {
    const int64_t k = quantum(r[q]);
    pz[q] = e->cache[k];
    pr[q] = e->cache_r[k];
    // unit vector from the contact to the receiver (sensing.go:308-318)
    pu[2 * q] = (e->rx[q] - x) / r[q];
    pu[2 * q + 1] = (e->ry[q] - y) / r[q];
  }
  double *df = e->d_force + (size_t)slot * 3 * e->cap;
  HIPCHK(hipMemcpyAsync(df, pf, sizeof(double) * n, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemcpyAsync(df + e->cap, pf + e->cap, sizeof(double) * n, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemcpyAsync(df + 2 * e->cap, pf + 2 * e->cap, sizeof(double) * n, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemcpyAsync(e->d_irp + (size_t)slot * e->nrecv, pz, sizeof(double *) * e->nrecv, hipMemcpyHostToDevice,
                        e->st));
  HIPCHK(hipMemcpyAsync(e->d_irr + (size_t)slot * e->nrecv, pr, sizeof(double *) * e->nrecv, hipMemcpyHostToDevice,
                        e->st));
  HIPCHK(hipMemcpyAsync(e->d_uv + (size_t)slot * e->nrecv * 2, pu, sizeof(double) * 2 * e->nrecv,
                        hipMemcpyHostToDevice, e->st));
  e->free_slots.pop_back();
  e->voices.push_back(gf_engine::Voice{slot, n, e->emitted, (int64_t)ceil(end_s * e->fs) + e->irlen});
  return ok();
}

int gf_engine_warm(gf_engine *e, const double *contact_xy, int32_t n) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (n < 0 || (n > 0 && !contact_xy)) return fail(GF_ERR_ARGUMENT, "geofdtd: bad contact positions");
  if (int rc = build_responses(e, contact_xy, n)) return rc;
  return ok();
}

int gf_engine_admit(gf_engine *e, const char *id, double x, double y, const double *force_z, int32_t n,
                    double end_s) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (e->radial)
    return fail(GF_ERR_ARGUMENT, "geofdtd: an analytic engine takes the whole force (gf_engine_admit_xyz)");
  if (n <= 0 || !force_z) return fail(GF_ERR_ARGUMENT, "geofdtd: a voice needs its force samples");
  for (int32_t q = 0; q < e->nrecv; q++)
    if (go_hypot(e->rx[q] - x, e->ry[q] - y) <= 0)
      return fail(GF_ERR_INVALID, "sensing: contact " + go_quote(id ? id : "") + " at the receiver's own position");
  if (n > e->cap)
    return fail(GF_ERR_ARGUMENT, "geofdtd: a contact of " + std::to_string(n) + " samples exceeds the engine's " +
                                     std::to_string(e->cap));
  if (e->free_slots.empty())
    return fail(GF_ERR_ARGUMENT, "geofdtd: more than " + std::to_string(kMaxVoices) + " voices at once");
  const double xy[2] = {x, y};
  if (int rc = build_responses(e, xy, 1)) return rc;
  const int32_t slot = e->free_slots.back();
  double *pf = e->pin_force + (size_t)slot * e->cap;
  const double **pp = e->pin_irp + (size_t)slot * e->nrecv;
  memcpy(pf, force_z, sizeof(double) * n);
  for (int32_t q = 0; q < e->nrecv; q++) pp[q] = e->cache[quantum(go_hypot(e->rx[q] - x, e->ry[q] - y))];
  HIPCHK(hipMemcpyAsync(e->d_force + (size_t)slot * e->cap, pf, sizeof(double) * n, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipMemcpyAsync(e->d_irp + (size_t)slot * e->nrecv, pp, sizeof(double *) * e->nrecv, hipMemcpyHostToDevice,
                        e->st));
  e->free_slots.pop_back();
  // The voice runs until its force has ended and the response it excited has
  // run out (sensing.go:323-328).
  e->voices.push_back(gf_engine::Voice{slot, n, e->emitted, (int64_t)ceil(end_s * e->fs) + e->irlen});
  return ok();
}

int gf_engine_next(gf_engine *e, float *out, int32_t len) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (len != e->chunk)
    return fail(GF_ERR_INVALID,
                "sensing: output length " + std::to_string(len) + ", want " + std::to_string(e->chunk));
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  MixArgs a;
  a.nv = 0;
  for (const auto &v : e->voices) {
    if (e->emitted > v.silent) continue;
    a.slot[a.nv] = v.slot;
    a.n[a.nv] = v.n;
    a.first[a.nv] = v.first;
    a.nv++;
  }
  const size_t nout = (size_t)e->nrecv * e->chunk;
  if (a.nv == 0) {
    memset(out, 0, sizeof(float) * nout);
    e->last_ms = 0;
  } else {
    HIPCHK(hipEventRecord(e->e0, e->st));
    dim3 grid(e->nrecv, (e->chunk + e->np - 1) / e->np);
    if (e->radial)
      k_engine_mix<true><<<grid, kMixThreads, 0, e->st>>>(a, e->d_force, e->cap, e->d_irp, e->d_irr, e->d_uv,
                                                          e->nrecv, e->irlen, e->emitted, e->chunk, e->np, e->d_out);
    else
      k_engine_mix<false><<<grid, kMixThreads, 0, e->st>>>(a, e->d_force, e->cap, e->d_irp, nullptr, nullptr,
                                                           e->nrecv, e->irlen, e->emitted, e->chunk, e->np,
                                                           e->d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e->e1, e->st));
    HIPCHK(hipMemcpyAsync(e->pin_out, e->d_out, sizeof(float) * nout, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->e0, e->e1));
    e->last_ms = ms;
    e->total_ms += ms;
    memcpy(out, e->pin_out, sizeof(float) * nout);
    // multiply-adds of this mix: per voice and output, the taps where both
    // the force record and the response are defined
    int64_t taps = 0;
    for (int v = 0; v < a.nv; v++)
      for (int32_t i = 0; i < e->chunk; i++) {
        const int64_t T = e->emitted + i - a.first[v];
        const int64_t lo = std::max<int64_t>(0, T - e->irlen + 1), hi = std::min<int64_t>(a.n[v] - 1, T);
        if (hi >= lo) taps += hi - lo + 1;
      }
    e->macs += taps * e->nrecv * (e->radial ? 2 : 1);
  }
  // retire (sensing.go:332-341).  A slot goes back on the free list only once
  // the stream has finished the admission copies out of its pinned staging:
  // after a mix the synchronization above covers them; when nothing mixed
  // (every live voice starts later, or a voice retires unheard) wait here.
  bool retiring = false;
  for (const auto &v : e->voices) retiring |= e->emitted > v.silent;
  if (retiring && a.nv == 0) HIPCHK(hipStreamSynchronize(e->st));
  std::vector<gf_engine::Voice> kept;
  for (const auto &v : e->voices) {
    if (e->emitted <= v.silent)
      kept.push_back(v);
    else
      e->free_slots.push_back(v.slot);
  }
  e->voices.swap(kept);
  e->emitted += e->chunk;
  return ok();
}

int gf_engine_info(const gf_engine *e, int64_t *emitted, int32_t *active_voices, int32_t *responses,
                   int32_t *ir_len) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (emitted) *emitted = e->emitted;
  if (active_voices) *active_voices = (int32_t)e->voices.size();
  if (responses) *responses = (int32_t)e->cache.size();
  if (ir_len) *ir_len = e->irlen;
  return ok();
}

int gf_engine_stats(const gf_engine *e, double *mix_ms, double *mix_ms_total, int64_t *macs, double *build_ms,
                    int64_t *built) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (mix_ms) *mix_ms = e->last_ms;
  if (mix_ms_total) *mix_ms_total = e->total_ms;
  if (macs) *macs = e->macs;
  if (build_ms) *build_ms = e->build_ms;
  if (built) *built = e->builds;
  return ok();
}

}  // extern "C"
