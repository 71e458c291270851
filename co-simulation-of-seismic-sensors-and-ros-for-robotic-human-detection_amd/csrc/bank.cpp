// bank.cpp — Green's-function banks from the batched FDTD.
//
// gf_bank_generate runs a batch of shots (one per soil model) with receivers
// on the bank's range grid, keeps the traces in HBM, and reduces each to the
// bank's frequency grid on the device:
//
//     V(f) = dt sum_n v[n] e^{-i w (T0 + n dt)}      (T0 = dt/2, run.go:99)
//     W(f) = dt sum_n F[n] e^{-i w n dt}             (the force the grid felt)
//     G(f) = V(f) / (i w W(f))                       [m/N]
//
// G is the vertical surface displacement per unit vertical point force,
// positive down: the same quantity and convention the reference's wavenumber
// banks hold (cmd/geobank/main.go:90-99), and the inverse of how its V5 test
// composes a velocity trace from one (cmd/geofdtd/main.go:171-202).  Bin 0
// and bins above max_freq are left zero, as geobank leaves them
// (main.go:116-121, 157-159).
//
// gf_bank_write serialises one bank in format v1 (internal/bank/format.go:
// 32-46, 215-240): magic, version, header length, the header as Go's
// encoding/json would write it (field order, omitempty, number and string
// formatting), zero padding to 4096, then little-endian float32 (re, im)
// pairs, range-major.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "batch.h"
#include "kernels.h"
#include "model.h"
#include "sim.h"

namespace gfs {
int api_set_err(int code, const std::string &msg);
}

namespace {

// ---- Go encoding/json formatting -------------------------------------------

// strconv.FormatFloat(f, 'f', -1, 64) or 'e' per encoding/json's floatEncoder:
// shortest round-trip digits; exponent form below 1e-6 or from 1e21, with
// "e-07" shortened to "e-7".
std::string json_number(double f) {
  if (f == 0) return signbit(f) ? "-0" : "0";
  char buf[64];
  int p = 1;
  for (; p < 17; p++) {
    snprintf(buf, sizeof buf, "%.*e", p - 1, f);
    if (strtod(buf, nullptr) == f) break;
  }
  snprintf(buf, sizeof buf, "%.*e", p - 1, f);
  const double a = fabs(f);
  if (a < 1e-6 || a >= 1e21) {
    std::string s = buf;
    const size_t e = s.find('e');
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);  // ex like "-07" or "+21"
    char sign = ex[0];
    std::string digits = ex.substr(1);
    while (digits.size() > 1 && digits[0] == '0') digits.erase(0, 1);
    // Go writes at least two exponent digits and then trims one leading zero
    // of a negative exponent: e-07 -> e-7, e+21 stays.
    if (sign == '-') return mant + "e-" + digits;
    if (digits.size() < 2) digits = "0" + digits;
    return mant + "e+" + digits;
  }
  // Fixed notation with exactly the shortest digits.
  const int ex = atoi(strchr(buf, 'e') + 1);
  snprintf(buf, sizeof buf, "%.*f", std::max(p - 1 - ex, 0), f);
  return buf;
}

// encoding/json string escaping with its default HTML-safe set.
std::string json_string(const char *s) {
  std::string o = "\"";
  for (const unsigned char *p = (const unsigned char *)(s ? s : ""); *p; p++) {
    const unsigned char c = *p;
    char buf[8];
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case '<': case '>': case '&':
        snprintf(buf, sizeof buf, "\\u%04x", c);
        o += buf;
        break;
      default:
        if (c < 0x20) {
          snprintf(buf, sizeof buf, "\\u%04x", c);
          o += buf;
        } else {
          o += (char)c;
        }
    }
  }
  return o + "\"";
}

// layer.Layer's JSON (layer.go:24-36): qs is omitempty; qp is not a field of
// the reference's layer and is not written.
std::string json_layer(const gf_layer &l) {
  std::string s = "{\"thickness_m\":" + json_number(l.thickness_m) + ",\"vp_mps\":" + json_number(l.vp_mps) +
                  ",\"vs_mps\":" + json_number(l.vs_mps) + ",\"density_kgm3\":" + json_number(l.density_kgm3);
  if (l.qs != 0) s += ",\"qs\":" + json_number(l.qs);
  return s + "}";
}

std::string header_json(const gf_bank_header &h) {
  std::string s = "{\"format_version\":1,\"provenance\":{\"solver\":" + json_string(h.solver);
  if (h.config_hash && *h.config_hash) s += ",\"config_hash\":" + json_string(h.config_hash);
  if (h.created && *h.created) s += ",\"created\":" + json_string(h.created);
  if (h.notes && *h.notes) s += ",\"notes\":" + json_string(h.notes);
  s += "},\"medium\":";
  if (h.n_medium <= 0 || !h.medium) {
    s += "null";
  } else {
    s += "[";
    for (int i = 0; i < h.n_medium; i++) s += (i ? "," : "") + json_layer(h.medium[i]);
    s += "]";
  }
  s += ",\"sample_rate_hz\":" + json_number(h.sample_rate_hz) + ",\"samples\":" + std::to_string(h.samples);
  s += ",\"ranges\":{\"min_m\":" + json_number(h.r_min_m) + ",\"max_m\":" + json_number(h.r_max_m) +
       ",\"count\":" + std::to_string(h.count) + "}";
  s += ",\"component\":" + json_string(h.component) + ",\"units\":" + json_string(h.units) + "}";
  return s;
}

// bank.Header.Validate (format.go:142-165).
int validate_header(const gf_bank_header &h, std::string &err) {
  char buf[256];
  if (h.samples <= 0 || h.samples % 2 != 0) {
    snprintf(buf, sizeof buf, "bank: samples must be positive and even, got %d", h.samples);
  } else if (h.sample_rate_hz <= 0) {
    snprintf(buf, sizeof buf, "bank: sample rate must be positive, got %s", gfm::go_g(h.sample_rate_hz).c_str());
  } else if (h.count <= 0) {
    snprintf(buf, sizeof buf, "bank: range grid must have at least one entry, got %d", h.count);
  } else if (h.r_min_m <= 0) {
    snprintf(buf, sizeof buf, "bank: minimum range must be positive, got %s m", gfm::go_g(h.r_min_m).c_str());
  } else if (h.count > 1 && h.r_max_m <= h.r_min_m) {
    snprintf(buf, sizeof buf, "bank: maximum range %s m must exceed minimum %s m", gfm::go_g(h.r_max_m).c_str(),
             gfm::go_g(h.r_min_m).c_str());
  } else if (!h.component || !*h.component) {
    snprintf(buf, sizeof buf, "bank: component must be named");
  } else {
    if (h.n_medium > 0 && h.medium) {
      gf_model m;
      memset(&m, 0, sizeof m);
      m.stack = h.medium;
      m.n_layers = h.n_medium;
      m.spacing_m = 1;
      m.max_range_m = 1;
      m.depth_m = 1;
      gfm::Resolved R;
      std::string e;
      if (gfm::resolve(m, R, e) == GF_ERR_INVALID && e.rfind("layer", 0) == 0) {
        err = "bank: " + e;
        return GF_ERR_INVALID;
      }
    }
    return GF_OK;
  }
  err = buf;
  return GF_ERR_INVALID;
}

double range_at(const gf_bank_spec &s, int i) {
  if (s.count <= 1) return s.r_min_m;
  return s.r_min_m + (s.r_max_m - s.r_min_m) * (double)i / (double)(s.count - 1);
}

}  // namespace

namespace gfs {
// bank.Header.Validate, for the device bank lookup.
int bank_validate_header(const gf_bank_header &h, std::string &err) { return validate_header(h, err); }
}  // namespace gfs

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

namespace {

// G(f_b) = V / (i w W) for rows of device traces; out[row][b] = (re, im) in
// FP64.  Rows where W vanishes are left zero.  d_freq holds nb frequencies.
int reduce_to_g(const float *d_traces, int64_t ld, int rows, int n, double dt, double t0,
                const std::vector<double> &force, const double *d_freq, const std::vector<double> &freqs, int nb,
                hipStream_t st, std::vector<double> &out) {
  double *d_v = nullptr, *d_w = nullptr, *d_force = nullptr;
  auto cleanup = [&]() {
    if (d_v) (void)hipFree(d_v);
    if (d_w) (void)hipFree(d_w);
    if (d_force) (void)hipFree(d_force);
  };
  out.assign(2 * (size_t)rows * nb, 0.0);
  if (nb <= 0 || rows <= 0 || force.empty()) return GF_OK;
  std::vector<double> v(2 * (size_t)rows * nb), w(2 * (size_t)nb);
  hipError_t e = hipMalloc(&d_v, sizeof(double) * v.size());
  if (e == hipSuccess) e = hipMalloc(&d_w, sizeof(double) * w.size());
  if (e == hipSuccess) e = hipMalloc(&d_force, sizeof(double) * force.size());
  if (e == hipSuccess) e = hipMemcpyAsync(d_force, force.data(), sizeof(double) * force.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = gfk::launch_dft_f32(d_traces, ld, rows, n, dt, t0, d_freq, nb, d_v, st);
  if (e == hipSuccess) e = gfk::launch_dft_f64(d_force, (int64_t)force.size(), 1, (int32_t)force.size(), dt, 0.0, d_freq, nb, d_w, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(v.data(), d_v, sizeof(double) * v.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(w.data(), d_w, sizeof(double) * w.size(), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return gfs::hip_fail(e, "spectrum reduction");
  for (int b = 0; b < nb; b++) {
    const double om = 2.0 * 3.141592653589793 * freqs[b];
    const double dr = -om * w[2 * b + 1], di = om * w[2 * b];  // i w W
    const double den = dr * dr + di * di;
    if (!(den > 0)) continue;
    for (int r = 0; r < rows; r++) {
      const double vr = v[2 * ((size_t)r * nb + b)], vi = v[2 * ((size_t)r * nb + b) + 1];
      out[2 * ((size_t)r * nb + b)] = (vr * dr + vi * di) / den;
      out[2 * ((size_t)r * nb + b) + 1] = (vi * dr - vr * di) / den;
    }
  }
  return GF_OK;
}

}  // namespace

extern "C" int gf_spectrum(const float *traces, int32_t rows, int32_t n, double dt, double t0, const double *force,
                           int64_t n_force, const double *freqs, int32_t nbins, double *out) {
  if (!traces || !force || !freqs || !out || rows <= 0 || n <= 0 || n_force <= 0 || nbins <= 0)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad spectrum arguments");
  std::string why;
  int rc = gfs::check_device(why);
  if (rc) return gfs::api_set_err(rc, why);
  hipStream_t st = nullptr;
  float *d_tr = nullptr;
  double *d_freq = nullptr;
  std::vector<double> g, fv(force, force + n_force), fr(freqs, freqs + nbins);
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipError_t e = hipMalloc(&d_tr, sizeof(float) * (size_t)rows * n);
  if (e == hipSuccess) e = hipMalloc(&d_freq, sizeof(double) * nbins);
  if (e == hipSuccess) e = hipMemcpy(d_tr, traces, sizeof(float) * (size_t)rows * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_freq, freqs, sizeof(double) * nbins, hipMemcpyHostToDevice);
  if (e == hipSuccess) rc = reduce_to_g(d_tr, n, rows, n, dt, t0, fv, d_freq, fr, nbins, st, g);
  if (d_tr) (void)hipFree(d_tr);
  if (d_freq) (void)hipFree(d_freq);
  (void)hipStreamDestroy(st);
  if (e != hipSuccess) return gfs::hip_fail(e, "spectrum set-up");
  if (rc) return rc;
  memcpy(out, g.data(), sizeof(double) * g.size());
  return gfs::api_set_err(GF_OK, "");
}

extern "C" int gf_bank_write(const char *path, const gf_bank_header *h, const float *data) {
  if (!path || !h || !data) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  std::string err;
  int rc = validate_header(*h, err);
  if (rc) return gfs::api_set_err(rc, err);
  const std::string hdr = header_json(*h);
  std::vector<unsigned char> buf;
  const unsigned char magic[8] = {'G', 'E', 'O', 'B', 'A', 'N', 'K', 0};
  buf.insert(buf.end(), magic, magic + 8);
  auto u32 = [&](uint32_t v) {
    for (int k = 0; k < 4; k++) buf.push_back((unsigned char)(v >> (8 * k)));
  };
  u32(1);
  u32((uint32_t)hdr.size());
  buf.insert(buf.end(), hdr.begin(), hdr.end());
  while (buf.size() % 4096) buf.push_back(0);
  const size_t nf = (size_t)2 * h->count * (h->samples / 2 + 1);
  for (size_t i = 0; i < nf; i++) {  // little-endian float32, as binary.Write(LittleEndian)
    uint32_t bits;
    memcpy(&bits, &data[i], sizeof bits);
    u32(bits);
  }
  FILE *fp = fopen(path, "wb");
  if (!fp) return gfs::api_set_err(GF_ERR_ARGUMENT, std::string("geofdtd: cannot create ") + path);
  const size_t w = fwrite(buf.data(), 1, buf.size(), fp);
  const int cl = fclose(fp);
  if (w != buf.size() || cl != 0) return gfs::api_set_err(GF_ERR_ARGUMENT, std::string("geofdtd: short write to ") + path);
  return gfs::api_set_err(GF_OK, "");
}

extern "C" int gf_bank_header_json(const gf_bank_header *h, char *out, int32_t cap) {
  if (!h || !out || cap <= 0) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  const std::string s = header_json(*h);
  if ((int64_t)s.size() + 1 > cap) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: header buffer too small");
  memcpy(out, s.c_str(), s.size() + 1);
  return gfs::api_set_err(GF_OK, "");
}

namespace {

int check_spec(const gf_bank_spec &spec) {
  gf_bank_header probe_h;
  memset(&probe_h, 0, sizeof probe_h);
  probe_h.sample_rate_hz = spec.sample_rate_hz;
  probe_h.samples = spec.samples;
  probe_h.r_min_m = spec.r_min_m;
  probe_h.r_max_m = spec.r_max_m;
  probe_h.count = spec.count;
  probe_h.component = "x";
  std::string err;
  int rc = validate_header(probe_h, err);
  return rc ? gfs::api_set_err(rc, err) : GF_OK;
}

std::vector<std::vector<double>> spec_ranges(const gf_bank_spec &spec, int32_t n_shots) {
  std::vector<std::vector<double>> ranges(n_shots);
  for (int i = 0; i < n_shots; i++)
    for (int r = 0; r < spec.count; r++) ranges[i].push_back(range_at(spec, r));
  return ranges;
}

// The bank spectra of a finished batch whose receivers sit on the spec's
// range grid: each shot's traces, still in HBM, reduced to G(f) on the device.
int bank_reduce(gfb::Batch &B, const gf_bank_spec &spec, float *spectra, double *snapped, gf_result *results) {
  const int n_shots = (int)B.sims.size();
  const int bins = spec.samples / 2 + 1;
  int top = (int)(spec.max_freq_hz / (spec.sample_rate_hz / spec.samples));
  if (top >= bins) top = bins - 1;
  const int nb = std::max(top, 0);  // bins 1..top
  std::vector<double> freqs(std::max(nb, 1));
  for (int k = 1; k <= top; k++) freqs[k - 1] = (double)k * spec.sample_rate_hz / spec.samples;
  memset(spectra, 0, sizeof(float) * 2 * (size_t)n_shots * spec.count * bins);
  int rc = GF_OK;
  double *d_freq = nullptr;
  HIPCHK(hipMalloc(&d_freq, sizeof(double) * freqs.size()));
  if (hipMemcpy(d_freq, freqs.data(), sizeof(double) * freqs.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d_freq);
    return gfs::api_set_err(GF_ERR_DEVICE, "geofdtd: copying the bank's frequencies failed");
  }
  std::vector<double> g;
  for (int i = 0; i < n_shots; i++) {
    if (!B.sims[i]) continue;
    double peak = 0;
    B.probes(i, &peak);
    if (results) results[i].peak = peak;
    if (!B.sims[i]) continue;
    gfs::Sim &s = *B.sims[i];
    for (int r = 0; r < spec.count && snapped; r++) snapped[(size_t)i * spec.count + r] = B.snapped[i][r];
    if (nb <= 0) continue;
    const double dt = s.R.dt;
    if ((rc = reduce_to_g(s.d_traces, s.g.trace_cap, spec.count, B.steps[i], dt, 0.5 * dt, s.force, d_freq, freqs,
                          nb, B.stream, g))) {
      B.fail_shot(i, rc, gf_last_error());
      continue;
    }
    float *out = spectra + (size_t)2 * i * spec.count * bins;
    for (int r = 0; r < spec.count; r++)
      for (int k = 1; k <= top; k++) {
        out[2 * ((size_t)r * bins + k)] = (float)g[2 * ((size_t)r * nb + k - 1)];
        out[2 * ((size_t)r * bins + k) + 1] = (float)g[2 * ((size_t)r * nb + k - 1) + 1];
      }
  }
  (void)hipFree(d_freq);
  for (int i = 0; i < n_shots; i++) {
    if (results) {
      results[i].status = B.st[i];
      snprintf(results[i].error, sizeof results[i].error, "%s", B.msg[i].c_str());
      if (B.sims[i]) {
        results[i].dt = B.sims[i]->R.dt;
        results[i].t0 = 0.5 * B.sims[i]->R.dt;
        results[i].nr = B.sims[i]->R.nr;
        results[i].nz = B.sims[i]->R.nz;
      }
    }
  }
  if (B.first_err) return gfs::api_set_err(B.first_err, B.first_msg);
  return gfs::api_set_err(GF_OK, "");
}

}  // namespace

extern "C" int gf_bank_generate(const gf_shot *shots, int32_t n_shots, const gf_bank_spec *spec, float *spectra,
                                double *snapped, gf_result *results) {
  if (!shots || !spec || !spectra || n_shots <= 0)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad bank arguments");
  int rc = check_spec(*spec);
  if (rc) return rc;
  gfb::Batch B;
  B.setup(shots, n_shots, spec_ranges(*spec, n_shots));
  rc = B.run();
  if (rc) return rc;
  return bank_reduce(B, *spec, spectra, snapped, results);
}

// ---- the batch as a handle (gf_batch_*) -------------------------------------

struct gf_batch {
  gfb::Batch B;
  bool bank = false;
  gf_bank_spec spec;
};

extern "C" int gf_batch_create(const gf_shot *shots, int32_t n_shots, const gf_bank_spec *spec, gf_batch **out) {
  if (!shots || !out || n_shots <= 0) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad batch arguments");
  *out = nullptr;
  std::vector<std::vector<double>> ranges(n_shots);
  if (spec) {
    int rc = check_spec(*spec);
    if (rc) return rc;
    ranges = spec_ranges(*spec, n_shots);
  } else {
    for (int i = 0; i < n_shots; i++) {
      if (shots[i].n_ranges > 0 && !shots[i].ranges)
        return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null ranges or traces in shot");
      if (shots[i].n_ranges > 0) ranges[i].assign(shots[i].ranges, shots[i].ranges + shots[i].n_ranges);
    }
  }
  std::unique_ptr<gf_batch> b(new gf_batch());
  b->bank = spec != nullptr;
  if (spec) b->spec = *spec;
  b->B.setup(shots, n_shots, ranges);
  int rc = b->B.plan();
  if (rc) return rc;
  *out = b.release();
  return gfs::api_set_err(GF_OK, "");
}

extern "C" int gf_batch_advance(gf_batch *b, int32_t num, int32_t den) {
  if (!b) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null batch");
  int rc = b->B.advance(num, den);
  return rc ? rc : gfs::api_set_err(GF_OK, "");
}

extern "C" int gf_batch_reset(gf_batch *b) {
  if (!b) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null batch");
  int rc = b->B.reset();
  return rc ? rc : gfs::api_set_err(GF_OK, "");
}

extern "C" int gf_batch_results(gf_batch *b, gf_result *results, int32_t *statuses) {
  if (!b || !results) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  if (!b->B.finished()) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: the batch has not run to its end");
  for (size_t i = 0; i < b->B.sims.size(); i++)
    if (!b->B.cols[i].empty() && !results[i].traces)
      return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null ranges or traces in shot");
  gfb::describe_results(b->B, results);
  return gfb::finish_results(b->B, results, statuses);
}

extern "C" int gf_batch_spectra(gf_batch *b, float *spectra, double *snapped, gf_result *results) {
  if (!b || !spectra) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  if (!b->bank) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: the batch was created without a bank spec");
  if (!b->B.finished()) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: the batch has not run to its end");
  return bank_reduce(b->B, b->spec, spectra, snapped, results);
}

extern "C" void gf_batch_destroy(gf_batch *b) { delete b; }
