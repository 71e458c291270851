// fk_api.cpp — C ABI of the frequency-wavenumber solver (include/geofdtd.h,
// the gf_fk_* entry points), mirroring the Go package internal/fk.
//
// The host does what the Go code does before any wavenumber is solved, in the
// same order and with the same texts: the layer-phase guard
// (checkLayerPhase, fk.go:253-268), the grid (GridFor, fk.go:456-489), the
// argument checks of VerticalDisplacement (fk.go:378-385) and of
// SurfaceResponse (fk.go:277-282, including layer.Stack.Validate).  Every
// solve and sum runs on the GPU (fk.hip).  A call over several frequencies
// reports what the Go calls, made one frequency after another, would have
// reported first.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "fk_kernels.h"
#include "model.h"
#include "sim.h"

namespace gfs {
int api_set_err(int code, const std::string &msg);
int bank_validate_header(const gf_bank_header &h, std::string &err);
}

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kMaxLayerPhase = 16.0;

thread_local gf_fk_stats_t g_stats;

int fail(int code, const std::string &msg) { return gfs::api_set_err(code, msg); }
int ok() { return gfs::api_set_err(GF_OK, ""); }

std::string g(double x) { return gfm::go_g(x); }

// Go's %.Nf: C's digits, Go's spelling of the infinities and NaN.
std::string go_f(double x, int prec) {
  if (isnan(x)) return "NaN";
  if (isinf(x)) return x > 0 ? "+Inf" : "-Inf";
  char b[64];
  snprintf(b, sizeof b, "%.*f", prec, x);
  return b;
}

// math.Max / math.Min.
double go_max(double x, double y) {
  if (isinf(x) && x > 0) return x;
  if (isinf(y) && y > 0) return y;
  if (isnan(x) || isnan(y)) return NAN;
  if (x == 0 && x == y) return signbit(x) ? y : x;
  return x > y ? x : y;
}
double go_min(double x, double y) {
  if (isinf(x) && x < 0) return x;
  if (isinf(y) && y < 0) return y;
  if (isnan(x) || isnan(y)) return NAN;
  if (x == 0 && x == y) return signbit(x) ? x : y;
  return x < y ? x : y;
}

// visco.SLS.Q (visco.go:84-90), with Go's Smith quotient.
double sls_q(const gf_sls &s, double f) {
  double w = 2 * kPi * f;
  double nr = 1, ni = w * s.tau_eps, mr = 1, mi = w * s.tau_sigma;
  double re, im;
  if (fabs(mr) >= fabs(mi)) {
    double ratio = mi / mr, denom = mr + ratio * mi;
    re = (nr + ni * ratio) / denom;
    im = (ni - nr * ratio) / denom;
  } else {
    double ratio = mr / mi, denom = mi + ratio * mr;
    re = (nr * ratio + ni) / denom;
    im = (ni * ratio - nr) / denom;
  }
  if (im == 0) return INFINITY;
  return re / im;
}

double ref_freq(const gf_medium &m) { return m.ref_freq_hz > 0 ? m.ref_freq_hz : 30.0; }

double q_of(const gf_medium &m, const gf_layer &l) {
  if (l.qs > 0) return l.qs;
  if (m.default_q > 0) return m.default_q;
  return 30.0;
}

// VelocityBounds' slowest (layer.go:86-93); the stack is not empty here.
double slowest(const gf_medium &m) {
  double s = m.stack[0].vs_mps;
  for (int i = 0; i < m.n_layers; i++) s = fmin(s, m.stack[i].vs_mps);
  return s;
}

// Null and size checks the Go types make impossible; an empty stack (a Go
// index panic) is reported as layer.Stack.Validate reports it.
int check_medium(const gf_medium *m, std::string &err) {
  if (!m) {
    err = "geofdtd: null medium";
    return GF_ERR_ARGUMENT;
  }
  if (m->n_layers < 0 || (m->n_layers > 0 && !m->stack)) {
    err = "geofdtd: bad layer array";
    return GF_ERR_ARGUMENT;
  }
  if (m->n_layers > GF_FK_MAX_LAYERS) {
    err = "geofdtd: the f-k solver takes at most " + std::to_string(GF_FK_MAX_LAYERS) + " layers";
    return GF_ERR_ARGUMENT;
  }
  if (m->n_layers == 0) {
    err = "layer: a stack needs at least a half-space";
    return GF_ERR_INVALID;
  }
  return GF_OK;
}

// checkLayerPhase (fk.go:253-268).
int check_layer_phase(const gf_medium &m, double f, std::string &err) {
  double sl = slowest(m);
  double k = 2 * kPi * f / (0.87 * sl);
  for (int i = 0; i < m.n_layers - 1; i++) {
    double th = m.stack[i].thickness_m;
    double kh = k * th;
    if (kh > kMaxLayerPhase) {
      err = "fk: layer " + std::to_string(i) + " is " + g(th) + " m thick, which is k*h = " + go_f(kh, 1) + " at " +
            g(f) + " Hz and past the " + go_f(kMaxLayerPhase, 0) +
            " this propagator stays accurate to; it would return a plausible wrong number rather than fail. "
            "Band-limit to " +
            go_f(kMaxLayerPhase * 0.87 * sl / (2 * kPi * th), 0) +
            " Hz for this stack, or reformulate the solve with compound matrices as internal/propmat does";
      return GF_ERR_INVALID;
    }
  }
  return GF_OK;
}

// SurfaceResponse's argument checks (fk.go:277-282).
int check_surface_args(const gf_medium &m, double k, double f, std::string &err) {
  if (k <= 0 || f <= 0) {
    err = "fk: wavenumber and frequency must be positive, got " + g(k) + " and " + g(f);
    return GF_ERR_INVALID;
  }
  return gfm::validate_stack(m.stack, m.n_layers, err);
}

int samples_of(const gf_integration &o) { return o.samples > 0 ? o.samples : 20000; }
double kmax_factor_of(const gf_integration &o) { return o.kmax_factor > 0 ? o.kmax_factor : 30.0; }

// int(math.Ceil(x)) as Go computes it on amd64: out-of-range and NaN give the
// minimum int64.  Counts this library cannot index are refused.
int ceil_count(double x, int64_t &n, std::string &err) {
  double c = ceil(x);
  if (!(c >= -9.2e18 && c <= 9.2e18)) {
    n = INT64_MIN;
    return GF_OK;
  }
  n = (int64_t)c;
  return GF_OK;
}

// GridFor (fk.go:456-489).
int grid_for(const gf_medium &m, const double *ranges, int32_t nr, double f, const gf_integration &opt, double &kmax,
             int64_t &n, std::string &err) {
  if (nr <= 0) {
    err = "fk: no ranges given";
    return GF_ERR_INVALID;
  }
  if (!ranges) {
    err = "geofdtd: null ranges";
    return GF_ERR_ARGUMENT;
  }
  double rmin = INFINITY, rmax = 0;
  for (int i = 0; i < nr; i++) {
    if (ranges[i] <= 0) {
      err = "fk: range must be positive, got " + g(ranges[i]) + " m";
      return GF_ERR_INVALID;
    }
    rmin = go_min(rmin, ranges[i]);
    rmax = go_max(rmax, ranges[i]);
  }
  double sl = slowest(m);
  double kbody = 2 * kPi * f / sl;
  kmax = go_max(kmax_factor_of(opt) * kbody, 40 / rmin);
  double dk = 2 * kPi / (8 * rmax);
  double q = m.has_relax ? sls_q(m.relax, f) : q_of(m, m.stack[m.n_layers - 1]);
  if (kbody > 0 && q > 0) dk = go_min(dk, kbody / (2 * q) / 4);
  int rc = ceil_count(kmax / dk, n, err);
  if (rc) return rc;
  if (opt.samples > 0) n = opt.samples;
  if (n < 64) n = 64;
  return GF_OK;
}

void to_device_medium(const gf_medium &m, gfw::Medium &M) {
  memset(&M, 0, sizeof M);
  M.nl = m.n_layers;
  M.has_relax = m.has_relax != 0;
  M.ref_freq = ref_freq(m);
  M.tau_sigma = m.relax.tau_sigma;
  M.tau_eps = m.relax.tau_eps;
  // referenceModulus (fk.go:216-223).
  double r = m.stack[0].density_kgm3 * m.stack[0].vs_mps * m.stack[0].vs_mps;
  for (int i = 0; i < m.n_layers; i++) {
    const gf_layer &l = m.stack[i];
    r = go_max(r, l.density_kgm3 * l.vs_mps * l.vs_mps);
    M.thick[i] = l.thickness_m;
    M.vp[i] = l.vp_mps;
    M.vs[i] = l.vs_mps;
    M.rho[i] = l.density_kgm3;
    M.q[i] = q_of(m, l);
  }
  M.mu_ref = r;
}

constexpr int64_t kMaxSamples = (int64_t)1 << 30;

// One frequency's host side of VerticalDisplacementMulti (fk.go:500-520):
// the guard, the grid and the static coefficient's argument checks.
int plan_multi(const gf_medium &m, const double *ranges, int32_t nr, double f, const gf_integration &opt,
               gfw::Job &job, std::string &err) {
  int rc = check_layer_phase(m, f, err);
  if (rc) return rc;
  double kmax;
  int64_t n;
  if ((rc = grid_for(m, ranges, nr, f, opt, kmax, n, err))) return rc;
  double sl = slowest(m);
  double kbody = 2 * kPi * f / sl;
  double rmin = INFINITY;
  for (int i = 0; i < nr; i++) rmin = go_min(rmin, ranges[i]);
  double kfar = go_max(200 * kbody, 400 / rmin);
  if ((rc = check_surface_args(m, kfar, f, err))) return rc;
  if (n > kMaxSamples) {
    err = "geofdtd: the f-k grid at " + g(f) + " Hz needs " + std::to_string(n) +
          " wavenumber samples, more than this library takes (2^30)";
    return GF_ERR_ARGUMENT;
  }
  job.f = f;
  job.dk = kmax / (double)n;
  job.kfar = kfar;
  job.n = (int32_t)n;
  job.first_block = 0;
  return GF_OK;
}

// The Go error for a job the GPU found singular, or GF_OK.
int job_error(const gfw::Job &J, const gfw::JobStatus &S, std::string &err) {
  auto singular = [&](int kind, double k) {
    err = std::string(kind == 1 ? "fk: layer propagator is singular at k=" : "fk: surface system is singular at k=") +
          g(k) + " f=" + g(J.f);
    return GF_ERR_INVALID;
  };
  if (S.far) return singular(S.far, J.kfar);
  if (S.sample != ~0ull) {
    int64_t i = (int64_t)(S.sample >> 1);
    return singular((int)(S.sample & 1) + 1, (double)i * J.dk);
  }
  return GF_OK;
}

void record_stats(const std::vector<gfw::Job> &jobs, int32_t nranges, const gfw::Timing &t) {
  gf_fk_stats_t s;
  memset(&s, 0, sizeof s);
  s.prologue_ms = t.prologue_ms;
  s.solve_ms = t.solve_ms;
  s.hankel_ms = t.hankel_ms;
  s.fold_ms = t.fold_ms;
  s.kernel_ms = t.prologue_ms + t.solve_ms + t.hankel_ms + t.fold_ms;
  for (const auto &j : jobs) s.samples += j.n;
  s.weights = s.samples * nranges;
  s.frequencies = (int32_t)jobs.size();
  s.launches = t.launches;
  g_stats = s;
}

int device_ready() {
  std::string why;
  int rc = gfs::check_device(why);
  return rc ? fail(rc, why) : GF_OK;
}

// Jobs for the frequencies before the first one the host rejects; that
// rejection (if any) is returned in host_rc/host_err for after the GPU's.
int plan_all(const gf_medium &m, const double *ranges, int32_t nr, const double *freqs, int32_t nf,
             const gf_integration &opt, std::vector<gfw::Job> &jobs, int &host_rc, std::string &host_err) {
  host_rc = GF_OK;
  for (int32_t b = 0; b < nf; b++) {
    gfw::Job j;
    int rc = plan_multi(m, ranges, nr, freqs[b], opt, j, host_err);
    if (rc) {
      host_rc = rc;
      break;
    }
    jobs.push_back(j);
  }
  return GF_OK;
}

int first_error(const std::vector<gfw::Job> &jobs, const std::vector<gfw::JobStatus> &st, int host_rc,
                const std::string &host_err) {
  std::string err;
  for (size_t j = 0; j < jobs.size(); j++)
    if (int rc = job_error(jobs[j], st[j], err)) return fail(rc, err);
  if (host_rc) return fail(host_rc, host_err);
  return GF_OK;
}

}  // namespace

extern "C" {

int gf_fk_check_layer_phase(const gf_medium *m, double f_hz) {
  std::string err;
  int rc = check_medium(m, err);
  if (!rc) rc = check_layer_phase(*m, f_hz, err);
  return rc ? fail(rc, err) : ok();
}

int gf_fk_grid_for(const gf_medium *m, const double *ranges, int32_t n_ranges, double f_hz, gf_integration opt,
                   gf_fk_grid *out) {
  std::string err;
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  int rc = check_medium(m, err);
  double kmax = 0;
  int64_t n = 0;
  if (!rc) rc = grid_for(*m, ranges, n_ranges, f_hz, opt, kmax, n, err);
  if (rc) return fail(rc, err);
  if (n > INT32_MAX) return fail(GF_ERR_ARGUMENT, "geofdtd: the f-k grid needs more samples than an int32 holds");
  out->kmax = kmax;
  out->samples = (int32_t)n;
  return ok();
}

int gf_fk_surface_response(const gf_medium *m, const double *k, const double *f_hz, int64_t n, double *out) {
  std::string err;
  int rc = check_medium(m, err);
  if (rc) return fail(rc, err);
  if (n < 0 || (n > 0 && (!k || !f_hz || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad point arrays");
  // Points before the first one whose arguments the Go call would reject.
  int64_t good = n;
  int host_rc = GF_OK;
  std::string host_err;
  for (int64_t i = 0; i < n; i++)
    if ((host_rc = check_surface_args(*m, k[i], f_hz[i], host_err))) {
      good = i;
      break;
    }
  if (good > 0) {
    if ((rc = device_ready())) return rc;
    gfw::Medium M;
    to_device_medium(*m, M);
    std::vector<int32_t> st((size_t)good);
    if ((rc = gfw::surface_points(M, k, f_hz, good, out, st.data()))) return rc;
    for (int64_t i = 0; i < good; i++)
      if (st[i])
        return fail(GF_ERR_INVALID,
                    std::string(st[i] == 1 ? "fk: layer propagator is singular at k=" : "fk: surface system is singular at k=") +
                        g(k[i]) + " f=" + g(f_hz[i]));
  }
  return host_rc ? fail(host_rc, host_err) : ok();
}

int gf_fk_static_coefficient(const gf_medium *m, double f_hz, double k_far, double out[2]) {
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  double u[2];
  int rc = gf_fk_surface_response(m, &k_far, &f_hz, 1, u);
  if (rc) return rc;
  // complex(kFar, 0) * u, as Go multiplies.
  out[0] = k_far * u[0] - 0.0 * u[1];
  out[1] = k_far * u[1] + 0.0 * u[0];
  return ok();
}

int gf_fk_vertical_displacement(const gf_medium *m, double r_m, double f_hz, gf_integration opt, double out[2]) {
  std::string err;
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  int rc = check_medium(m, err);
  if (rc) return fail(rc, err);
  if (r_m <= 0) return fail(GF_ERR_INVALID, "fk: range must be positive, got " + g(r_m) + " m");
  if (f_hz <= 0) return fail(GF_ERR_INVALID, "fk: frequency must be positive, got " + g(f_hz) + " Hz");
  if ((rc = check_layer_phase(*m, f_hz, err))) return fail(rc, err);
  double sl = slowest(*m);
  double kbody = 2 * kPi * f_hz / sl;
  double kmax = kmax_factor_of(opt) * kbody;
  kmax = go_max(kmax, 40 / r_m);
  double kfar = go_max(200 * kbody, 400 / r_m);
  if ((rc = check_surface_args(*m, kfar, f_hz, err))) return fail(rc, err);
  int n = samples_of(opt);
  if ((rc = device_ready())) return rc;
  gfw::Medium M;
  to_device_medium(*m, M);
  std::vector<gfw::Job> jobs{gfw::Job{f_hz, kmax / (double)n, kfar, n, 0}};
  std::vector<gfw::JobStatus> st(1);
  gfw::Timing t{};
  if ((rc = gfw::multi(M, jobs.data(), 1, &r_m, 1, out, st.data(), &t))) return rc;
  record_stats(jobs, 1, t);
  if ((rc = job_error(jobs[0], st[0], err))) return fail(rc, err);
  return ok();
}

int gf_fk_vertical_displacement_multi(const gf_medium *m, const double *ranges, int32_t n_ranges, const double *f_hz,
                                      int32_t n_freqs, gf_integration opt, double *out) {
  std::string err;
  int rc = check_medium(m, err);
  if (rc) return fail(rc, err);
  if (n_freqs < 0 || (n_freqs > 0 && (!f_hz || !out))) return fail(GF_ERR_ARGUMENT, "geofdtd: bad frequency arrays");
  std::vector<gfw::Job> jobs;
  int host_rc;
  std::string host_err;
  plan_all(*m, ranges, n_ranges, f_hz, n_freqs, opt, jobs, host_rc, host_err);
  if (!jobs.empty()) {
    if ((rc = device_ready())) return rc;
    gfw::Medium M;
    to_device_medium(*m, M);
    std::vector<gfw::JobStatus> st(jobs.size());
    gfw::Timing t{};
    std::vector<double> tmp((size_t)jobs.size() * n_ranges * 2);
    if ((rc = gfw::multi(M, jobs.data(), (int32_t)jobs.size(), ranges, n_ranges, tmp.data(), st.data(), &t)))
      return rc;
    record_stats(jobs, n_ranges, t);
    if ((rc = first_error(jobs, st, host_rc, host_err))) return rc;
    memcpy(out, tmp.data(), sizeof(double) * tmp.size());
  }
  return host_rc ? fail(host_rc, host_err) : ok();
}

int gf_fk_synthesise(const gf_medium *m, const double *ranges, int32_t n_ranges, double peak_n, double f0_hz,
                     double band_hz, double dt, double t0, int32_t n, int32_t samples, double *out) {
  std::string err;
  int rc = check_medium(m, err);
  if (rc) return fail(rc, err);
  if (n_ranges < 0 || n < 0 || (n_ranges > 0 && !ranges) || (n_ranges > 0 && n > 0 && !out))
    return fail(GF_ERR_ARGUMENT, "geofdtd: bad synthesis arrays");
  if (!(dt > 0)) return fail(GF_ERR_ARGUMENT, "geofdtd: synthesis needs a positive time step");
  // dsp.NextPow2 (dsp.go:20-31).
  int64_t nfft = 1;
  while (nfft < n) nfft <<= 1;
  if (nfft > INT32_MAX / 2) return fail(GF_ERR_ARGUMENT, "geofdtd: synthesis length too large");
  double fs = 1 / dt;
  double df = fs / (double)nfft;
  std::vector<double> freqs;
  for (int64_t b = 1; b <= nfft / 2; b++) {
    double f = (double)b * df;
    if (f > band_hz) break;
    freqs.push_back(f);
  }
  gf_integration opt{samples, 0};
  std::vector<gfw::Job> jobs;
  int host_rc;
  std::string host_err;
  plan_all(*m, ranges, n_ranges, freqs.data(), (int32_t)freqs.size(), opt, jobs, host_rc, host_err);
  if (host_rc && jobs.empty()) return fail(host_rc, host_err);
  if (n_ranges == 0 || n == 0) return ok();
  if ((rc = device_ready())) return rc;
  gfw::Medium M;
  to_device_medium(*m, M);
  std::vector<gfw::JobStatus> st(std::max<size_t>(jobs.size(), 1));
  gfw::Timing t{};
  if (host_rc) {
    // An earlier bin's GPU error would come first; otherwise the host's.
    if ((rc = gfw::multi(M, jobs.data(), (int32_t)jobs.size(), ranges, n_ranges, nullptr, st.data(), &t))) return rc;
    return first_error(jobs, st, host_rc, host_err);
  }
  if ((rc = gfw::synthesise(M, jobs.data(), (int32_t)jobs.size(), ranges, n_ranges, peak_n, f0_hz, t0, fs,
                            (int32_t)nfft, n, out, st.data(), &t)))
    return rc;
  record_stats(jobs, n_ranges, t);
  if ((rc = first_error(jobs, st, GF_OK, ""))) return rc;
  return ok();
}

int gf_fk_bank_fill(const gf_medium *m, const gf_bank_header *h, double max_freq_hz, float *payload) {
  std::string err;
  if (!h || !payload) return fail(GF_ERR_ARGUMENT, "geofdtd: null argument");
  int rc = check_medium(m, err);
  if (rc) return fail(rc, err);
  if ((rc = gfs::bank_validate_header(*h, err))) return fail(rc, err);
  if ((rc = gf_bank_check_range_sampling(h, max_freq_hz))) return rc;
  const int32_t count = h->count, bins = h->samples / 2 + 1;
  std::vector<double> ranges((size_t)count);
  for (int32_t i = 0; i < count; i++)  // RangeGrid.At (format.go:52-57)
    ranges[i] = count <= 1 ? h->r_min_m : h->r_min_m + (h->r_max_m - h->r_min_m) * (double)i / (double)(count - 1);
  // topBin (main.go:115-119) and FrequencyAt (format.go:137-139).
  int32_t top = (int32_t)(max_freq_hz / (h->sample_rate_hz / (double)h->samples));
  if (top >= bins) top = bins - 1;
  memset(payload, 0, sizeof(float) * 2 * (size_t)count * bins);
  if (top < 1) return ok();
  std::vector<double> freqs((size_t)top);
  for (int32_t k = 1; k <= top; k++) freqs[k - 1] = (double)k * h->sample_rate_hz / (double)h->samples;
  std::vector<double> vals((size_t)top * count * 2);
  gf_integration opt{0, 0};
  rc = gf_fk_vertical_displacement_multi(m, ranges.data(), count, freqs.data(), top, opt, vals.data());
  if (rc) {
    // Which bin: the first whose own call fails, as the error names it.
    for (int32_t k = 1; k <= top; k++) {
      std::vector<double> one((size_t)count * 2);
      if (gf_fk_vertical_displacement_multi(m, ranges.data(), count, &freqs[k - 1], 1, opt, one.data()) != GF_OK) {
        const std::string inner = gf_last_error();
        char head[96];
        snprintf(head, sizeof head, "bin %d (%.3f Hz): ", k, freqs[k - 1]);
        return fail(GF_ERR_INVALID, head + inner);
      }
    }
    return rc;
  }
  for (int32_t k = 1; k <= top; k++)
    for (int32_t i = 0; i < count; i++) {
      const double *v = &vals[((size_t)(k - 1) * count + i) * 2];
      float *o = payload + 2 * ((size_t)i * bins + k);
      o[0] = (float)v[0];  // bank.Set: float32(real(v)), float32(imag(v))
      o[1] = (float)v[1];
    }
  return ok();
}

int gf_fk_stats(gf_fk_stats_t *out) {
  if (!out) return fail(GF_ERR_ARGUMENT, "geofdtd: null output");
  *out = g_stats;
  return ok();
}

}  // extern "C"
