// model.cpp — see model.h.  Host-only FP64; compiled with -ffp-contract=off.
#include "model.h"

#include "gofmt.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

namespace gfm {

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kSqrt2 = 1.4142135623730951;

struct Cx {
  double re, im;
};

// runtime/complex.go's complex128div: Smith's algorithm.
Cx cdiv(Cx n, Cx m) {
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
  return {e, f};
}

// math.Hypot: max * sqrt(1 + (min/max)^2).
double hypot_go(double p, double q) {
  p = fabs(p);
  q = fabs(q);
  if (isinf(p) || isinf(q)) return INFINITY;
  if (isnan(p) || isnan(q)) return NAN;
  if (p < q) std::swap(p, q);
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

// math/cmplx.Sqrt, scaled to avoid overflow exactly as Go scales it.
Cx csqrt_go(Cx x) {
  if (x.im == 0) {
    if (x.re == 0) return {0, x.im};
    if (x.re < 0) return {0, copysign(sqrt(-x.re), x.im)};
    return {sqrt(x.re), x.im};
  }
  if (x.re == 0) {
    if (x.im < 0) {
      double t = sqrt(-0.5 * x.im);
      return {t, -t};
    }
    double t = sqrt(0.5 * x.im);
    return {t, t};
  }
  double a = x.re, b = x.im, scale;
  if (fabs(a) > 4 || fabs(b) > 4) {
    a *= 0.25;
    b *= 0.25;
    scale = 2;
  } else {
    a *= 1.8014398509481984e16;  // 2^54
    b *= 1.8014398509481984e16;
    scale = 7.450580596923828125e-9;  // 2^-27
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
  return b < 0 ? Cx{t, -r} : Cx{t, r};
}

// math.Max.
double max_go(double x, double y) {
  if (isinf(x) && x > 0) return x;
  if (isinf(y) && y > 0) return y;
  if (isnan(x) || isnan(y)) return NAN;
  if (x == 0 && x == y) return signbit(x) ? y : x;
  return x > y ? x : y;
}

// ---- visco.SLS (visco.go:42-124) ----

double sls_delta(const gf_sls &s) { return s.tau_eps / s.tau_sigma - 1; }

Cx sls_modulus(const gf_sls &s, double f) {
  const double w = 2 * kPi * f;
  // (1 + i w TauEps) / (1 + i w TauSigma), Go's complex multiply expanded.
  Cx num = {1 + (0 * s.tau_eps - w * 0.0), 0 + (0 * 0.0 + w * s.tau_eps)};
  Cx den = {1 + (0 * s.tau_sigma - w * 0.0), 0 + (0 * 0.0 + w * s.tau_sigma)};
  return cdiv(num, den);
}

double relaxed_velocity(const gf_sls &s, double v, double fref) {
  return v * cdiv({1, 0}, csqrt_go(sls_modulus(s, fref))).re;
}

double unrelaxed_modulus(const gf_sls &s, double v, double fref, double rho) {
  double c0 = relaxed_velocity(s, v, fref);
  double relaxed = rho * c0 * c0;
  return relaxed * (1 + sls_delta(s));
}

// ---- GSLS extension: generalized Maxwell body, Emmerich-Korn form ----
// M(w) = M_U [1 - sum_l Y_l w_l / (w_l + i w)]; Y_l fitted so that
// Q^-1(w_k) = sum_l Y_l (w_l w_k + w_l^2 Q^-1) / (w_l^2 + w_k^2) holds in least
// squares at 2N-1 log-spaced frequencies (Kristek & Moczo 2003).

struct Band {
  int n = 0;
  double fmin = 5, fmax = 150;
  double w[gfk::kMaxMech] = {0};
};

Band make_band(const gf_model &m) {
  Band b;
  b.n = m.gsls_mechanisms;
  b.fmin = m.gsls_fmin_hz > 0 ? m.gsls_fmin_hz : 5.0;
  b.fmax = m.gsls_fmax_hz > 0 ? m.gsls_fmax_hz : 150.0;
  double wmin = 2 * kPi * b.fmin, wmax = 2 * kPi * b.fmax;
  for (int l = 0; l < b.n; l++) {
    double x = b.n == 1 ? 0.5 : (double)l / (double)(b.n - 1);
    b.w[l] = wmin * pow(wmax / wmin, x);
  }
  return b;
}

void fit_weights(const Band &b, double q, double *y) {
  const int n = b.n;
  for (int l = 0; l < n; l++) y[l] = 0;
  if (!(q > 0)) return;  // Q == 0: lossless layer
  const double wmin = 2 * kPi * b.fmin, wmax = 2 * kPi * b.fmax, qi = 1.0 / q;
  const int kn = 2 * n - 1;
  double a[gfk::kMaxMech][gfk::kMaxMech + 1] = {{0}};
  for (int k = 0; k < kn; k++) {
    double x = kn == 1 ? 0.5 : (double)k / (double)(kn - 1);
    double wk = wmin * pow(wmax / wmin, x);
    double row[gfk::kMaxMech];
    for (int l = 0; l < n; l++) row[l] = (b.w[l] * wk + b.w[l] * b.w[l] * qi) / (b.w[l] * b.w[l] + wk * wk);
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < n; j++) a[i][j] += row[i] * row[j];
      a[i][n] += row[i] * qi;
    }
  }
  for (int c = 0; c < n; c++) {  // Gaussian elimination, partial pivoting
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
    if (p != c)
      for (int j = 0; j <= n; j++) std::swap(a[c][j], a[p][j]);
    for (int r = c + 1; r < n; r++) {
      double f = a[r][c] / a[c][c];
      for (int j = c; j <= n; j++) a[r][j] -= f * a[c][j];
    }
  }
  for (int c = n - 1; c >= 0; c--) {
    double s = a[c][n];
    for (int j = c + 1; j < n; j++) s -= a[c][j] * y[j];
    y[c] = s / a[c][c];
  }
}

Cx band_r(const Band &b, const double *y, double f) {
  const double w = 2 * kPi * f;
  Cx r = {1, 0};
  for (int l = 0; l < b.n; l++) {
    Cx t = cdiv({y[l] * b.w[l], 0}, {b.w[l], w});
    r.re -= t.re;
    r.im -= t.im;
  }
  return r;
}

// M_U such that the phase velocity at fref is v.
double band_unrelaxed(const Band &b, const double *y, double v, double fref, double rho) {
  double s = cdiv({1, 0}, csqrt_go(band_r(b, y, fref))).re;
  double c = v * s;
  return rho * c * c;
}

double layer_qp(const gf_layer &l) { return l.qp > 0 ? l.qp : l.qs; }

double ref_freq(const gf_model &m) { return m.ref_freq_hz > 0 ? m.ref_freq_hz : 30.0; }

int validate_ext(const gf_model &m, std::string &err) {
  char buf[256];
  if (m.gsls_mechanisms < 0 || m.gsls_mechanisms > 4) {
    snprintf(buf, sizeof buf, "fdtd: gsls mechanisms must be in [0, 4], got %d", m.gsls_mechanisms);
    err = buf;
    return GF_ERR_INVALID;
  }
  if (m.gsls_mechanisms > 0) {
    Band b = make_band(m);
    if (!(b.fmax > b.fmin)) {
      err = "fdtd: gsls band must have fmax > fmin";
      return GF_ERR_INVALID;
    }
    for (int i = 0; i < m.n_layers; i++)
      if (m.stack[i].qs < 0 || m.stack[i].qp < 0) {
        snprintf(buf, sizeof buf, "fdtd: layer %d: Q must not be negative", i);
        err = buf;
        return GF_ERR_INVALID;
      }
  }
  return GF_OK;
}

// The slowest relaxed shear velocity (SpacingFor, dominantFreq).
double slowest_relaxed(const gf_model &m) {
  double slowest = m.stack[0].vs_mps;
  for (int i = 0; i < m.n_layers; i++) slowest = std::min(slowest, m.stack[i].vs_mps);
  if (m.gsls_mechanisms > 0) {
    Band b = make_band(m);
    double best = INFINITY;
    for (int i = 0; i < m.n_layers; i++) {
      if (m.stack[i].vs_mps != slowest) continue;
      double y[gfk::kMaxMech];
      fit_weights(b, m.stack[i].qs, y);
      double mu = band_unrelaxed(b, y, slowest, ref_freq(m), 1.0), sy = 0;
      for (int l = 0; l < b.n; l++) sy += y[l];
      best = std::min(best, sqrt(mu * (1 - sy)));
    }
    return best;
  }
  if (m.relax.tau_eps > m.relax.tau_sigma) return relaxed_velocity(m.relax, slowest, ref_freq(m));
  return slowest;
}

double max_velocity(const gf_model &m) {
  double vp = 0.0;
  const double fref = ref_freq(m);
  if (m.gsls_mechanisms > 0) {
    Band b = make_band(m);
    for (int i = 0; i < m.n_layers; i++) {
      double y[gfk::kMaxMech];
      fit_weights(b, layer_qp(m.stack[i]), y);
      vp = max_go(vp, sqrt(band_unrelaxed(b, y, m.stack[i].vp_mps, fref, 1.0)));
    }
    return vp;
  }
  const bool lossy = m.relax.tau_eps > m.relax.tau_sigma;
  for (int i = 0; i < m.n_layers; i++) {
    double v = m.stack[i].vp_mps;
    if (lossy) {
      double c0 = relaxed_velocity(m.relax, v, fref);
      v = c0 * sqrt(1 + sls_delta(m.relax));
    }
    vp = max_go(vp, v);
  }
  return vp;
}

// newProfile's `first` (pml.go:40-59): the lowest index with any damping.
int profile_first(int n, double h, double offset, double inner, double thick, double d0, double amax) {
  if (thick <= 0) return n;
  for (int i = 0; i < n; i++) {
    double s = ((double)i + offset) * h - inner;
    if (s <= 0) continue;
    double x = std::min(s / thick, 1.0);
    double d = d0 * (x * x);
    double alpha = amax * (1 - x);
    if (d + alpha <= 0) continue;
    return i;
  }
  return n;
}

}  // namespace

// layer.Stack.Validate (internal/layer/layer.go:53-73), with its messages.
int validate_stack(const gf_layer *st, int n, std::string &err) {
  char buf[512];
  if (n <= 0 || st == nullptr) {
    err = "layer: a stack needs at least a half-space";
    return GF_ERR_INVALID;
  }
  for (int i = 0; i < n; i++) {
    const gf_layer &l = st[i];
    if (l.vs_mps <= 0) {
      snprintf(buf, sizeof buf, "layer %d: Vs must be positive, got %s m/s", i, go_g(l.vs_mps).c_str());
    } else if (l.vp_mps <= 0) {
      snprintf(buf, sizeof buf, "layer %d: Vp must be positive, got %s m/s", i, go_g(l.vp_mps).c_str());
    } else if (l.density_kgm3 <= 0) {
      snprintf(buf, sizeof buf, "layer %d: density must be positive, got %s kg/m^3", i,
               go_g(l.density_kgm3).c_str());
    } else if (l.vp_mps <= kSqrt2 * l.vs_mps) {
      snprintf(buf, sizeof buf, "layer %d: Vp=%s must exceed sqrt(2)*Vs=%s for a non-negative Poisson ratio", i,
               go_g(l.vp_mps).c_str(), go_g(kSqrt2 * l.vs_mps).c_str());
    } else if (i < n - 1 && l.thickness_m <= 0) {
      snprintf(buf, sizeof buf, "layer %d: thickness must be positive, got %s m", i, go_g(l.thickness_m).c_str());
    } else {
      continue;
    }
    err = buf;
    return GF_ERR_INVALID;
  }
  return GF_OK;
}

std::string go_g(double x) { return gofmt::g(x); }

int visco_fit(double q, double f0, gf_sls &out, std::string &err) {
  if (q <= 0) {
    err = "visco: Q must be positive, got " + go_g(q);
    return GF_ERR_INVALID;
  }
  if (f0 <= 0) {
    err = "visco: reference frequency must be positive, got " + go_g(f0) + " Hz";
    return GF_ERR_INVALID;
  }
  double w0 = 2 * kPi * f0;
  double root = sqrt(1 + 1 / (q * q));
  out.tau_sigma = (root - 1 / q) / w0;
  out.tau_eps = (root + 1 / q) / w0;
  return GF_OK;
}

int spacing_for(const gf_model &m, double max_freq, double cpw, double &out, std::string &err) {
  if (max_freq <= 0 || cpw <= 0) {
    err = "fdtd: bandwidth and cells per wavelength must be positive";
    return GF_ERR_INVALID;
  }
  int rc = validate_stack(m.stack, m.n_layers, err);
  if (rc) return rc;
  if ((rc = validate_ext(m, err))) return rc;
  out = slowest_relaxed(m) / (max_freq * cpw);
  return GF_OK;
}

void ricker(double peak, double f0, double dt, int32_t n, double *out) {
  const double t0 = 1.5 / f0;
  const double a = 9.869604401089358 * f0 * f0;  // Go folds math.Pi*math.Pi to this constant
  for (int32_t i = 0; i < n; i++) {
    double tau = (double)i * dt - t0;
    out[i] = peak * (1 - 2 * a * tau * tau) * exp(-a * tau * tau);
  }
}

void ricker_spectrum(double peak, double f0, double f, double &re, double &im) {
  double fr = f / f0;
  double mag = 2 / sqrt(kPi) * fr * fr * exp(-fr * fr) / f0;
  double phase = -2 * kPi * f * 1.5 / f0;
  double pm = peak * mag, c = cos(phase), s = sin(phase);
  re = pm * c - 0 * s;
  im = pm * s + 0 * c;
}

int resolve(const gf_model &m, Resolved &R, std::string &err) {
  char buf[256];
  int rc = validate_stack(m.stack, m.n_layers, err);
  if (rc) return rc;
  if (m.spacing_m <= 0) {
    err = "fdtd: spacing must be positive, got " + go_g(m.spacing_m) + " m";
    return GF_ERR_INVALID;
  }
  if (m.max_range_m <= 0 || m.depth_m <= 0) {
    err = "fdtd: domain must be positive, got " + go_g(m.max_range_m) + " m by " + go_g(m.depth_m) + " m";
    return GF_ERR_INVALID;
  }
  if (m.relax.tau_eps < m.relax.tau_sigma) {
    err = "fdtd: TauEps must be at least TauSigma for a dissipative solid";
    return GF_ERR_INVALID;
  }
  if ((rc = validate_ext(m, err))) return rc;

  R.stack.assign(m.stack, m.stack + m.n_layers);
  const double h = m.spacing_m;
  R.h = h;
  R.pml = m.pml_cells > 0 ? m.pml_cells : 20;
  const double nr_d = round(m.max_range_m / h), nz_d = round(m.depth_m / h);
  if (nr_d + 1 + R.pml > 1 << 20 || nz_d + 1 + R.pml > 1 << 20) {
    snprintf(buf, sizeof buf, "fdtd: grid of %.0f x %.0f cells is too large", nr_d + 1 + R.pml, nz_d + 1 + R.pml);
    err = buf;
    return GF_ERR_INVALID;
  }
  R.nr = (int)nr_d + 1 + R.pml;
  R.nz = (int)nz_d + 1 + R.pml;
  R.max_range = m.max_range_m;
  R.depth = m.depth_m;
  R.vmax = max_velocity(m);
  const double courant = m.courant > 0 ? m.courant : 0.5;
  R.dt = courant * (h / (R.vmax * kSqrt2));
  R.src_area = kPi * h * h / 4;
  R.refl = m.pml_reflection > 0 ? m.pml_reflection : 1e-4;
  const double fref = ref_freq(m);

  R.layers.resize(m.n_layers);
  if (m.gsls_mechanisms > 0) {
    Band b = make_band(m);
    R.scalar = false;
    R.nmech = b.n;
    bool any = false;
    for (int q = 0; q < b.n; q++) R.decay[q] = (1 - R.dt * b.w[q] / 2) / (1 + R.dt * b.w[q] / 2);
    for (int i = 0; i < m.n_layers; i++) {
      const gf_layer &l = m.stack[i];
      gfk::LayerParams &L = R.layers[i];
      double yp[gfk::kMaxMech], ys[gfk::kMaxMech];
      fit_weights(b, layer_qp(l), yp);
      fit_weights(b, l.qs, ys);
      L.thickness = l.thickness_m;
      L.rho = l.density_kgm3;
      L.mp = band_unrelaxed(b, yp, l.vp_mps, fref, l.density_kgm3);
      L.ms = band_unrelaxed(b, ys, l.vs_mps, fref, l.density_kgm3);
      for (int q = 0; q < gfk::kMaxMech; q++) L.drive_p[q] = L.drive_s[q] = 0;
      for (int q = 0; q < b.n; q++) {
        double den = 1 + R.dt * b.w[q] / 2;
        L.drive_p[q] = R.dt * b.w[q] * yp[q] / den;
        L.drive_s[q] = R.dt * b.w[q] * ys[q] / den;
        if (L.drive_p[q] != 0 || L.drive_s[q] != 0) any = true;
      }
    }
    R.lossy = any;
    if (!any) R.nmech = 0;
  } else {
    R.scalar = true;
    R.lossy = m.relax.tau_eps > m.relax.tau_sigma;
    R.nmech = R.lossy ? 1 : 0;
    if (R.lossy) {
      const double tau = m.relax.tau_sigma;
      const double kappa = sls_delta(m.relax) / (1 + sls_delta(m.relax));
      const double den = 1 + R.dt / (2 * tau);
      R.mem_decay = (1 - R.dt / (2 * tau)) / den;
      R.mem_drive = (R.dt * kappa / tau) / den;
    }
    for (int i = 0; i < m.n_layers; i++) {
      const gf_layer &l = m.stack[i];
      gfk::LayerParams &L = R.layers[i];
      L.thickness = l.thickness_m;
      L.rho = l.density_kgm3;
      if (R.lossy) {
        L.mp = unrelaxed_modulus(m.relax, l.vp_mps, fref, l.density_kgm3);
        L.ms = unrelaxed_modulus(m.relax, l.vs_mps, fref, l.density_kgm3);
      } else {
        L.mp = l.density_kgm3 * l.vp_mps * l.vp_mps;
        L.ms = l.density_kgm3 * l.vs_mps * l.vs_mps;
      }
      for (int q = 0; q < gfk::kMaxMech; q++) L.drive_p[q] = L.drive_s[q] = 0;
    }
  }

  // The surface row's layer (layerAt(0)).
  {
    double top = 0;
    int li = m.n_layers - 1;
    for (int i = 0; i < m.n_layers - 1; i++) {
      top += m.stack[i].thickness_m;
      if (0.0 < top) {
        li = i;
        break;
      }
    }
    R.mu0 = R.layers[li].ms;
    R.lam0 = R.layers[li].mp - 2 * R.layers[li].ms;
  }

  R.dominant = m.dominant_freq_hz > 0 ? m.dominant_freq_hz : slowest_relaxed(m) / (20 * h);
  R.pml_thickness = (double)R.pml * h;
  const double order = 2;
  R.d0 = -(order + 1) * R.vmax * log(R.refl) / (2 * R.pml_thickness);
  R.amax = kPi * R.dominant;
  R.first_rint = profile_first(R.nr, h, 0, m.max_range_m, R.pml_thickness, R.d0, R.amax);
  R.first_rhalf = profile_first(R.nr, h, 0.5, m.max_range_m, R.pml_thickness, R.d0, R.amax);
  R.first_zint = profile_first(R.nz, h, 0, m.depth_m, R.pml_thickness, R.d0, R.amax);
  R.first_zhalf = profile_first(R.nz, h, 0.5, m.depth_m, R.pml_thickness, R.d0, R.amax);
  return GF_OK;
}

}  // namespace gfm
