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

__global__ __launch_bounds__(kMixThreads) void k_engine_mix(MixArgs a, const double *__restrict__ force, int32_t cap,
                                                            const double *const *__restrict__ irp, int32_t nrecv,
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
      const double *f = force + (int64_t)s * cap;
      // the response table holds device addresses: read them as global
      // memory, not flat
      const __attribute__((address_space(1))) double *h =
          (const __attribute__((address_space(1))) double *)irp[(int64_t)s * nrecv + r];
#pragma unroll 8
      for (int64_t m = lo + q; m <= hi; m += S) acc = fma(f[m], h[T - m], acc);
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
  if (int rc = gfbank::impulse_check(e->bank, rq.data(), nk, e->irlen, e->fs)) return rc;
  if (e->build_cap < (size_t)nk) {
    HIPCHK(hipStreamSynchronize(e->st));
    hipFree(e->d_r);
    hipFree(e->d_coeff);
    e->d_r = nullptr;
    e->d_coeff = nullptr;
    HIPCHK(hipMalloc(&e->d_r, sizeof(double) * nk));
    HIPCHK(hipMalloc(&e->d_coeff, sizeof(double2) * (size_t)nk * e->bins));
    e->build_cap = nk;
  }
  double *block = nullptr;
  HIPCHK(hipMalloc(&block, sizeof(double) * (size_t)nk * e->irlen));
  e->blocks.push_back(block);
  HIPCHK(hipMemcpyAsync(e->d_r, rq.data(), sizeof(double) * nk, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipEventRecord(e->e0, e->st));
  HIPCHK(gfbank::impulse_launch(e->bank, e->d_r, nk, e->irlen, e->fs, e->d_sensor, e->d_coeff, block, e->st));
  HIPCHK(hipEventRecord(e->e1, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e->e0, e->e1));
  e->build_ms += ms;
  e->builds += nk;
  for (int32_t j = 0; j < nk; j++) e->cache[keys[j]] = block + (size_t)j * e->irlen;
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

int gf_engine_warm(gf_engine *e, const double *contact_xy, int32_t n) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
  if (n < 0 || (n > 0 && !contact_xy)) return fail(GF_ERR_ARGUMENT, "geofdtd: bad contact positions");
  if (int rc = build_responses(e, contact_xy, n)) return rc;
  return ok();
}

int gf_engine_admit(gf_engine *e, const char *id, double x, double y, const double *force_z, int32_t n,
                    double end_s) {
  if (!e) return fail(GF_ERR_ARGUMENT, "geofdtd: null engine handle");
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
    k_engine_mix<<<grid, kMixThreads, 0, e->st>>>(a, e->d_force, e->cap, e->d_irp, e->nrecv, e->irlen, e->emitted,
                                                  e->chunk, e->np, e->d_out);
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
    e->macs += taps * e->nrecv;
  }
  // retire (sensing.go:332-341)
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
