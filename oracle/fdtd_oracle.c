/*
 * fdtd_oracle.c — FP64 CPU restatement of the reference's axisymmetric
 * viscoelastic FDTD, used ONLY as the parity checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and there only to check or time the
 * GPU path against it.  The product (libgeofdtd) never links or calls it.
 *
 * What it restates, and from where (paths relative to the reference repo):
 *   visco.SLS, Fit, Delta, Modulus, Q, RelaxedVelocity, Moduli
 *                                       internal/visco/visco.go:42-124
 *   layer.Stack.Validate, VelocityBounds, HalfSpace
 *                                       internal/layer/layer.go:40-100
 *   Model defaults, SpacingFor, materials, harmonic, layerAt, maxVelocity,
 *   timeStep                            internal/fdtd/model.go:78-246
 *   newProfile, profile.update          internal/fdtd/pml.go:29-70
 *   New, dominantFreq, Dt, Cells, Time, Drive, forceAt, Column, Step,
 *   velocities, stresses, relax, SurfaceVelocity, RadialVelocity, MaxSpeed
 *                                       internal/fdtd/fdtd.go:45-360
 *   Ricker, RickerSpectrum, Run         internal/fdtd/run.go:19-127
 *
 * Every expression keeps the Go operand order, and the file is compiled with
 * -ffp-contract=off because Go on amd64 (GOAMD64=v1) does not fuse multiply
 * and add.  Go's complex division (Smith's algorithm, runtime/complex.go),
 * cmplx.Sqrt and math.Hypot are restated rather than taken from libm, since
 * their rounding differs from C99's.  math.Exp/Log/Pow/Erf come from libm;
 * they can differ from Go's by an ulp, which moves results at the 1e-16 level.
 *
 * The extension (gf_model.gsls_mechanisms > 0: per-layer Q, separate P and S
 * relaxation, N mechanisms) has no reference counterpart.  Its attenuation
 * law is the generalized Maxwell body in the Emmerich-Korn form
 * (Emmerich & Korn 1987, Geophysics 52:1252; fitted as in Kristek & Moczo
 * 2003, BSSA 93:2273):
 *     M(w) = M_U [1 - sum_l Y_l w_l / (w_l + i w)],
 *     Q^-1(w_k) = sum_l Y_l (w_l w_k + w_l^2 Q^-1) / (w_l^2 + w_k^2)
 * solved for Y_l in least squares over 2N-1 log-spaced frequencies, with w_l
 * log-spaced over the band.  In the time domain each mechanism is the
 * reference's Crank-Nicolson memory variable with w_l = 1/TauSigma and
 * Y_l = Delta/(1+Delta) (fdtd.go:68-75), so one mechanism with Y_P == Y_S is
 * the reference scheme term for term.
 *
 * Parity status of each function is recorded in the tests that pin it:
 * tests/test_oracle_visco.py (closed forms), tests/test_oracle_known_answers.py
 * (the reference's logged Boussinesq, Rayleigh, C-PML and attenuation numbers).
 */
#include <errno.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "geofdtd.h"
#include "goarith.h"

#define ORC_PI 3.141592653589793
#define ORC_SQRT2 1.4142135623730951
#define ORC_MAX_MECH 8

/* ---------------------------------------------------------------- errors */

__thread char orc_err[512];

int orc_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(orc_err, sizeof orc_err, fmt, ap);
  va_end(ap);
  return code;
}

const char *orc_last_error(void) { return orc_err; }

/* ------------------------------------------------------------------ visco */

int orc_visco_fit(double q, double f0, gf_sls *out) {
  char b[64];
  if (q <= 0) return fail(GF_ERR_INVALID, "visco: Q must be positive, got %s", gofmt_g(q, b, sizeof b));
  if (f0 <= 0)
    return fail(GF_ERR_INVALID, "visco: reference frequency must be positive, got %s Hz",
                gofmt_g(f0, b, sizeof b));
  double w0 = 2 * ORC_PI * f0;
  double root = sqrt(1 + 1 / (q * q));
  out->tau_sigma = (root - 1 / q) / w0;
  out->tau_eps = (root + 1 / q) / w0;
  orc_err[0] = 0;
  return GF_OK;
}

double orc_sls_delta(gf_sls s) { return s.tau_eps / s.tau_sigma - 1; }

static cx sls_modulus(gf_sls s, double f) {
  cx w = {0, 2 * ORC_PI * f};
  /* 1 + w*complex(TauEps, 0), with Go's complex multiply expanded. */
  cx num = {1 + (w.re * s.tau_eps - w.im * 0.0), 0 + (w.re * 0.0 + w.im * s.tau_eps)};
  cx den = {1 + (w.re * s.tau_sigma - w.im * 0.0), 0 + (w.re * 0.0 + w.im * s.tau_sigma)};
  return go_cdiv(num, den);
}

void orc_sls_modulus(gf_sls s, double f, double *re, double *im) {
  cx m = sls_modulus(s, f);
  *re = m.re;
  *im = m.im;
}

double orc_sls_q(gf_sls s, double f) {
  cx m = sls_modulus(s, f);
  if (m.im == 0) return INFINITY;
  return m.re / m.im;
}

double orc_sls_relaxed_velocity(gf_sls s, double v, double fref) {
  cx one = {1, 0};
  return v * go_cdiv(one, go_csqrt(sls_modulus(s, fref))).re;
}

/* SLS.Velocity: complex(c0, 0) * cmplx.Sqrt(Modulus(f)). */
void orc_sls_velocity(gf_sls s, double v, double fref, double f, double *re, double *im) {
  double c0 = orc_sls_relaxed_velocity(s, v, fref);
  cx q = go_csqrt(sls_modulus(s, f));
  *re = c0 * q.re - 0 * q.im;
  *im = c0 * q.im + 0 * q.re;
}

void orc_sls_moduli(gf_sls s, double v, double fref, double rho, double *relaxed, double *unrelaxed) {
  double c0 = orc_sls_relaxed_velocity(s, v, fref);
  double r = rho * c0 * c0;
  *relaxed = r;
  *unrelaxed = r * (1 + orc_sls_delta(s));
}

/* ------------------------------------------------------------------ layer */

static int stack_validate(const gf_layer *st, int n) {
  char a[64], b[64];
  if (n <= 0 || st == NULL) return fail(GF_ERR_INVALID, "layer: a stack needs at least a half-space");
  for (int i = 0; i < n; i++) {
    const gf_layer *l = &st[i];
    if (l->vs_mps <= 0)
      return fail(GF_ERR_INVALID, "layer %d: Vs must be positive, got %s m/s", i, gofmt_g(l->vs_mps, a, 64));
    if (l->vp_mps <= 0)
      return fail(GF_ERR_INVALID, "layer %d: Vp must be positive, got %s m/s", i, gofmt_g(l->vp_mps, a, 64));
    if (l->density_kgm3 <= 0)
      return fail(GF_ERR_INVALID, "layer %d: density must be positive, got %s kg/m^3", i,
                  gofmt_g(l->density_kgm3, a, 64));
    if (l->vp_mps <= ORC_SQRT2 * l->vs_mps)
      return fail(GF_ERR_INVALID, "layer %d: Vp=%s must exceed sqrt(2)*Vs=%s for a non-negative Poisson ratio", i,
                  gofmt_g(l->vp_mps, a, 64), gofmt_g(ORC_SQRT2 * l->vs_mps, b, 64));
    if (i < n - 1 && l->thickness_m <= 0)
      return fail(GF_ERR_INVALID, "layer %d: thickness must be positive, got %s m", i,
                  gofmt_g(l->thickness_m, a, 64));
  }
  return GF_OK;
}

static void velocity_bounds(const gf_layer *st, int n, double *slowest, double *fastest) {
  double s = st[0].vs_mps, f = st[0].vs_mps;
  for (int i = 0; i < n; i++) {
    s = fmin(s, st[i].vs_mps);
    f = fmax(f, st[i].vs_mps);
  }
  *slowest = s;
  *fastest = f;
}

/* ------------------------------------------------------------------ model */

static double ref_freq(const gf_model *m) { return m->ref_freq_hz > 0 ? m->ref_freq_hz : 30.0; }
static int pml_cells(const gf_model *m) { return m->pml_cells > 0 ? m->pml_cells : 20; }
static double pml_refl(const gf_model *m) { return m->pml_reflection > 0 ? m->pml_reflection : 1e-4; }
static double courant(const gf_model *m) { return m->courant > 0 ? m->courant : 0.5; }
static int ref_lossy(const gf_model *m) { return m->relax.tau_eps > m->relax.tau_sigma; }

/* ---- GSLS extension: per-layer, per-wave-type relaxation spectra ---- */

typedef struct {
  int n;                    /* mechanisms */
  double w[ORC_MAX_MECH];   /* relaxation angular frequencies, shared */
} gsls_band;

static void gsls_band_make(const gf_model *m, gsls_band *b) {
  double fmin = m->gsls_fmin_hz > 0 ? m->gsls_fmin_hz : 5.0;
  double fmax = m->gsls_fmax_hz > 0 ? m->gsls_fmax_hz : 150.0;
  b->n = m->gsls_mechanisms;
  double wmin = 2 * ORC_PI * fmin, wmax = 2 * ORC_PI * fmax;
  for (int l = 0; l < b->n; l++) {
    double x = b->n == 1 ? 0.5 : (double)l / (double)(b->n - 1);
    b->w[l] = wmin * pow(wmax / wmin, x);
  }
}

/* Least-squares weights Y_l for constant Q over the band (normal equations,
 * Gaussian elimination with partial pivoting; N <= 8). */
static void gsls_fit(const gf_model *m, const gsls_band *b, double q, double *y) {
  int n = b->n;
  for (int l = 0; l < n; l++) y[l] = 0;
  if (!(q > 0)) return;
  double fmin = m->gsls_fmin_hz > 0 ? m->gsls_fmin_hz : 5.0;
  double fmax = m->gsls_fmax_hz > 0 ? m->gsls_fmax_hz : 150.0;
  double wmin = 2 * ORC_PI * fmin, wmax = 2 * ORC_PI * fmax;
  int k_n = 2 * n - 1;
  double qi = 1.0 / q;
  double ata[ORC_MAX_MECH][ORC_MAX_MECH + 1];
  memset(ata, 0, sizeof ata);
  for (int k = 0; k < k_n; k++) {
    double x = k_n == 1 ? 0.5 : (double)k / (double)(k_n - 1);
    double wk = wmin * pow(wmax / wmin, x);
    double row[ORC_MAX_MECH];
    for (int l = 0; l < n; l++) {
      double wl = b->w[l];
      row[l] = (wl * wk + wl * wl * qi) / (wl * wl + wk * wk);
    }
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < n; j++) ata[i][j] += row[i] * row[j];
      ata[i][n] += row[i] * qi;
    }
  }
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int r = c + 1; r < n; r++)
      if (fabs(ata[r][c]) > fabs(ata[p][c])) p = r;
    if (p != c)
      for (int j = 0; j <= n; j++) {
        double t = ata[c][j];
        ata[c][j] = ata[p][j];
        ata[p][j] = t;
      }
    for (int r = c + 1; r < n; r++) {
      double f = ata[r][c] / ata[c][c];
      for (int j = c; j <= n; j++) ata[r][j] -= f * ata[c][j];
    }
  }
  for (int c = n - 1; c >= 0; c--) {
    double s = ata[c][n];
    for (int j = c + 1; j < n; j++) s -= ata[c][j] * y[j];
    y[c] = s / ata[c][c];
  }
}

/* R(w) = 1 - sum_l Y_l w_l / (w_l + i w): the modulus normalised to M_U. */
static cx gsls_r(const gsls_band *b, const double *y, double f) {
  double w = 2 * ORC_PI * f;
  cx r = {1, 0};
  for (int l = 0; l < b->n; l++) {
    cx num = {y[l] * b->w[l], 0};
    cx den = {b->w[l], w};
    cx t = go_cdiv(num, den);
    r.re -= t.re;
    r.im -= t.im;
  }
  return r;
}

/* Unrelaxed modulus that makes the phase velocity v at fref. */
static double gsls_unrelaxed(const gsls_band *b, const double *y, double v, double fref, double rho) {
  cx one = {1, 0};
  double s = go_cdiv(one, go_csqrt(gsls_r(b, y, fref))).re;
  double c = v * s;
  return rho * c * c;
}

/* Q(f) of a fitted spectrum, for the tests. */
double orc_gsls_q(const gf_model *m, double q, double f) {
  gsls_band b;
  double y[ORC_MAX_MECH];
  gsls_band_make(m, &b);
  gsls_fit(m, &b, q, y);
  cx r = gsls_r(&b, y, f);
  return r.re / r.im;
}

int orc_gsls_weights(const gf_model *m, double q, double *w_out, double *y_out) {
  gsls_band b;
  gsls_band_make(m, &b);
  gsls_fit(m, &b, q, y_out);
  for (int l = 0; l < b.n; l++) w_out[l] = b.w[l];
  return b.n;
}

static double layer_qp(const gf_layer *l) { return l->qp > 0 ? l->qp : l->qs; }

static int model_validate_ext(const gf_model *m) {
  if (m->gsls_mechanisms < 0 || m->gsls_mechanisms > ORC_MAX_MECH)
    return fail(GF_ERR_INVALID, "fdtd: gsls mechanisms must be in [0, %d], got %d", ORC_MAX_MECH,
                m->gsls_mechanisms);
  if (m->gsls_mechanisms > 0) {
    double fmin = m->gsls_fmin_hz > 0 ? m->gsls_fmin_hz : 5.0;
    double fmax = m->gsls_fmax_hz > 0 ? m->gsls_fmax_hz : 150.0;
    if (!(fmax > fmin)) return fail(GF_ERR_INVALID, "fdtd: gsls band must have fmax > fmin");
    for (int i = 0; i < m->n_layers; i++)
      if (m->stack[i].qs < 0 || m->stack[i].qp < 0)
        return fail(GF_ERR_INVALID, "fdtd: layer %d: Q must not be negative", i);
  }
  return GF_OK;
}

/* Slowest relaxed shear velocity: what SpacingFor and dominantFreq use. */
static double slowest_relaxed(const gf_model *m) {
  double slowest, fastest;
  velocity_bounds(m->stack, m->n_layers, &slowest, &fastest);
  if (m->gsls_mechanisms > 0) {
    /* The layer with the slowest nominal Vs, at its own shear relaxation:
     * sqrt(M_R/rho) with M_R = M_U (1 - sum Y). */
    gsls_band b;
    gsls_band_make(m, &b);
    double best = INFINITY;
    for (int i = 0; i < m->n_layers; i++) {
      if (m->stack[i].vs_mps != slowest) continue;
      double y[ORC_MAX_MECH];
      gsls_fit(m, &b, m->stack[i].qs, y);
      double mu = gsls_unrelaxed(&b, y, slowest, ref_freq(m), 1.0), sy = 0;
      for (int l = 0; l < b.n; l++) sy += y[l];
      best = fmin(best, sqrt(mu * (1 - sy)));
    }
    return best;
  }
  if (ref_lossy(m)) return orc_sls_relaxed_velocity(m->relax, slowest, ref_freq(m));
  return slowest;
}

int orc_spacing_for(const gf_model *m, double max_freq, double cpw, double *out) {
  if (max_freq <= 0 || cpw <= 0)
    return fail(GF_ERR_INVALID, "fdtd: bandwidth and cells per wavelength must be positive");
  int rc = stack_validate(m->stack, m->n_layers);
  if (rc) return rc;
  double v = slowest_relaxed(m);
  *out = v / (max_freq * cpw);
  orc_err[0] = 0;
  return GF_OK;
}

/* Unrelaxed P velocity of the stiffest layer (model.go:226-242). */
static double max_velocity(const gf_model *m) {
  double vp = 0.0, fref = ref_freq(m);
  if (m->gsls_mechanisms > 0) {
    gsls_band b;
    gsls_band_make(m, &b);
    for (int i = 0; i < m->n_layers; i++) {
      double y[ORC_MAX_MECH];
      gsls_fit(m, &b, layer_qp(&m->stack[i]), y);
      double v = sqrt(gsls_unrelaxed(&b, y, m->stack[i].vp_mps, fref, 1.0));
      vp = go_max(vp, v);
    }
    return vp;
  }
  int lossy = ref_lossy(m);
  for (int i = 0; i < m->n_layers; i++) {
    double v = m->stack[i].vp_mps;
    if (lossy) {
      double c0 = orc_sls_relaxed_velocity(m->relax, v, fref);
      v = c0 * sqrt(1 + orc_sls_delta(m->relax));
    }
    vp = go_max(vp, v);
  }
  return vp;
}

static double time_step(const gf_model *m, double h) { return h / (max_velocity(m) * ORC_SQRT2); }

static const gf_layer *layer_at(const gf_model *m, double z) {
  double top = 0;
  for (int i = 0; i < m->n_layers - 1; i++) {
    top += m->stack[i].thickness_m;
    if (z < top) return &m->stack[i];
  }
  return &m->stack[m->n_layers - 1];
}

static double harmonic(double a, double b) {
  if (a <= 0 || b <= 0) return 0;
  return 2 * a * b / (a + b);
}

/* ---------------------------------------------------------------- profile */

typedef struct {
  double *a, *b;
  int first;
} profile;

static void new_profile(profile *p, int n, double h, double offset, double inner, double thickness, double vmax,
                        double reflection, double dominant, double dt) {
  p->a = (double *)calloc((size_t)n, sizeof(double));
  p->b = (double *)malloc((size_t)n * sizeof(double));
  p->first = n;
  for (int i = 0; i < n; i++) p->b[i] = 1;
  if (thickness <= 0) return;
  const double order = 2;
  double d0 = -(order + 1) * vmax * log(reflection) / (2 * thickness);
  double amax = ORC_PI * dominant;
  for (int i = 0; i < n; i++) {
    double s = ((double)i + offset) * h - inner;
    if (s <= 0) continue;
    double x = fmin(s / thickness, 1);
    double d = d0 * pow(x, order);
    double alpha = amax * (1 - x);
    if (d + alpha <= 0) continue;
    double b = exp(-(d + alpha) * dt);
    p->a[i] = d * (b - 1) / (d + alpha);
    p->b[i] = b;
    if (i < p->first) p->first = i;
  }
}

static inline double pupdate(const profile *p, double *psi, int i, double deriv) {
  *psi = p->b[i] * *psi + p->a[i] * deriv;
  return *psi;
}

/* -------------------------------------------------------------------- Sim */

typedef struct orc_sim {
  gf_model m;
  gf_layer *stack; /* owned copy */
  int nr, nz;
  double dr, dz, dt;
  double *rho, *lamU, *muU, *rhoZ, *muZ;
  double *vr, *vz, *trr, *tqq, *tzz, *trz;
  int lossy;   /* any memory variables at all */
  int nmech;   /* mechanisms (1 in the reference path) */
  int scalar;  /* reference form: one drive per mechanism for P and S alike */
  double memDecay, memDrive; /* reference path */
  double decay[ORC_MAX_MECH];
  double *driveP, *driveS; /* [nmech][nz], extension path */
  double *rrr, *rqq, *rzz, *rrz; /* [nmech][n] */
  profile pRint, pRhalf, pZint, pZhalf;
  double *psiTrrR, *psiTrzZ, *psiTrzR, *psiTzzZ, *psiVrR, *psiVzZ, *psiVzR, *psiVrZ;
  double *force;
  int64_t nforce;
  double srcArea;
  int step;
} orc_sim;

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  double *arrs[] = {s->rho, s->lamU, s->muU, s->rhoZ, s->muZ, s->vr, s->vz, s->trr, s->tqq, s->tzz, s->trz,
                    s->driveP, s->driveS, s->rrr, s->rqq, s->rzz, s->rrz, s->pRint.a, s->pRint.b, s->pRhalf.a,
                    s->pRhalf.b, s->pZint.a, s->pZint.b, s->pZhalf.a, s->pZhalf.b, s->psiTrrR, s->psiTrzZ,
                    s->psiTrzR, s->psiTzzZ, s->psiVrR, s->psiVzZ, s->psiVzR, s->psiVrZ, s->force};
  for (size_t i = 0; i < sizeof arrs / sizeof arrs[0]; i++) free(arrs[i]);
  free(s->stack);
  free(s);
}

static void materials(orc_sim *s) {
  const gf_model *m = &s->m;
  int nz = s->nz;
  double dz = s->dz, fref = ref_freq(m);
  s->rho = calloc(nz, sizeof(double));
  s->lamU = calloc(nz, sizeof(double));
  s->muU = calloc(nz, sizeof(double));
  s->rhoZ = calloc(nz, sizeof(double));
  s->muZ = calloc(nz, sizeof(double));
  gsls_band band;
  if (m->gsls_mechanisms > 0) gsls_band_make(m, &band);
  for (int j = 0; j < nz; j++) {
    const gf_layer *l = layer_at(m, (double)j * dz);
    double rho = l->density_kgm3, vp = l->vp_mps, vs = l->vs_mps, mp, ms;
    if (m->gsls_mechanisms > 0) {
      double yp[ORC_MAX_MECH], ys[ORC_MAX_MECH];
      gsls_fit(m, &band, layer_qp(l), yp);
      gsls_fit(m, &band, l->qs, ys);
      mp = gsls_unrelaxed(&band, yp, vp, fref, rho);
      ms = gsls_unrelaxed(&band, ys, vs, fref, rho);
      for (int q = 0; q < band.n; q++) {
        double w = band.w[q], den = 1 + s->dt * w / 2;
        s->driveP[(size_t)q * nz + j] = s->dt * w * yp[q] / den;
        s->driveS[(size_t)q * nz + j] = s->dt * w * ys[q] / den;
      }
    } else if (ref_lossy(m)) {
      double r;
      orc_sls_moduli(m->relax, vp, fref, rho, &r, &mp);
      orc_sls_moduli(m->relax, vs, fref, rho, &r, &ms);
    } else {
      mp = rho * vp * vp;
      ms = rho * vs * vs;
    }
    s->rho[j] = rho;
    s->muU[j] = ms;
    s->lamU[j] = mp - 2 * ms;
  }
  for (int j = 0; j < nz; j++) {
    int k = j + 1 < nz - 1 ? j + 1 : nz - 1;
    s->rhoZ[j] = 0.5 * (s->rho[j] + s->rho[k]);
    s->muZ[j] = harmonic(s->muU[j], s->muU[k]);
  }
}

int orc_sim_create(const gf_model *min, orc_sim **out) {
  char a[64], b[64];
  *out = NULL;
  int rc = stack_validate(min->stack, min->n_layers);
  if (rc) return rc;
  if (min->spacing_m <= 0)
    return fail(GF_ERR_INVALID, "fdtd: spacing must be positive, got %s m", gofmt_g(min->spacing_m, a, 64));
  if (min->max_range_m <= 0 || min->depth_m <= 0)
    return fail(GF_ERR_INVALID, "fdtd: domain must be positive, got %s m by %s m", gofmt_g(min->max_range_m, a, 64),
                gofmt_g(min->depth_m, b, 64));
  if (min->relax.tau_eps < min->relax.tau_sigma)
    return fail(GF_ERR_INVALID, "fdtd: TauEps must be at least TauSigma for a dissipative solid");
  if ((rc = model_validate_ext(min))) return rc;

  orc_sim *s = calloc(1, sizeof *s);
  s->m = *min;
  s->stack = malloc(sizeof(gf_layer) * (size_t)min->n_layers);
  memcpy(s->stack, min->stack, sizeof(gf_layer) * (size_t)min->n_layers);
  s->m.stack = s->stack;
  const gf_model *m = &s->m;
  double h = m->spacing_m;
  int pml = pml_cells(m);
  s->nr = (int)round(m->max_range_m / h) + 1 + pml;
  s->nz = (int)round(m->depth_m / h) + 1 + pml;
  s->dr = s->dz = h;
  s->dt = courant(m) * time_step(m, h);
  s->nmech = m->gsls_mechanisms > 0 ? m->gsls_mechanisms : 1;
  if (m->gsls_mechanisms > 0) {
    s->driveP = calloc((size_t)s->nmech * s->nz, sizeof(double));
    s->driveS = calloc((size_t)s->nmech * s->nz, sizeof(double));
    gsls_band band;
    gsls_band_make(m, &band);
    for (int q = 0; q < band.n; q++) {
      double w = band.w[q];
      s->decay[q] = (1 - s->dt * w / 2) / (1 + s->dt * w / 2);
    }
  }
  materials(s);
  s->srcArea = ORC_PI * h * h / 4;

  if (m->gsls_mechanisms > 0) {
    s->scalar = 0;
    s->lossy = 0;
    for (size_t i = 0; i < (size_t)s->nmech * s->nz; i++)
      if (s->driveP[i] != 0 || s->driveS[i] != 0) s->lossy = 1;
  } else {
    s->scalar = 1;
    s->lossy = m->relax.tau_eps > m->relax.tau_sigma;
    if (s->lossy) {
      double tau = m->relax.tau_sigma;
      double kappa = orc_sls_delta(m->relax) / (1 + orc_sls_delta(m->relax));
      double den = 1 + s->dt / (2 * tau);
      s->memDecay = (1 - s->dt / (2 * tau)) / den;
      s->memDrive = (s->dt * kappa / tau) / den;
    }
  }

  size_t n = (size_t)s->nr * s->nz;
  double **fields[] = {&s->vr, &s->vz, &s->trr, &s->tqq, &s->tzz, &s->trz, &s->psiTrrR, &s->psiTrzZ,
                       &s->psiTrzR, &s->psiTzzZ, &s->psiVrR, &s->psiVzZ, &s->psiVzR, &s->psiVrZ};
  for (size_t i = 0; i < sizeof fields / sizeof fields[0]; i++) *fields[i] = calloc(n, sizeof(double));
  if (s->lossy) {
    size_t nm = n * (size_t)s->nmech;
    s->rrr = calloc(nm, sizeof(double));
    s->rqq = calloc(nm, sizeof(double));
    s->rzz = calloc(nm, sizeof(double));
    s->rrz = calloc(nm, sizeof(double));
  }

  double thickness = (double)pml * h;
  double vmax = max_velocity(m);
  double refl = pml_refl(m);
  double dom = m->dominant_freq_hz > 0 ? m->dominant_freq_hz : slowest_relaxed(m) / (20 * s->dr);
  new_profile(&s->pRint, s->nr, h, 0, m->max_range_m, thickness, vmax, refl, dom, s->dt);
  new_profile(&s->pRhalf, s->nr, h, 0.5, m->max_range_m, thickness, vmax, refl, dom, s->dt);
  new_profile(&s->pZint, s->nz, h, 0, m->depth_m, thickness, vmax, refl, dom, s->dt);
  new_profile(&s->pZhalf, s->nz, h, 0.5, m->depth_m, thickness, vmax, refl, dom, s->dt);
  *out = s;
  orc_err[0] = 0;
  return GF_OK;
}

void orc_sim_info(const orc_sim *s, double *dt, int32_t *nr, int32_t *nz) {
  *dt = s->dt;
  *nr = s->nr;
  *nz = s->nz;
}

int orc_sim_lossy(const orc_sim *s) { return s->lossy; }
int orc_sim_has_memory(const orc_sim *s) { return s->rrr != NULL; }
double orc_sim_time(const orc_sim *s) { return ((double)s->step - 0.5) * s->dt; }

/* Coefficient tables, for comparing the device-built ones. which: 0 rho,
 * 1 lamU, 2 muU, 3 rhoZ, 4 muZ (length nz); 5..8 profile a of rint, rhalf,
 * zint, zhalf; 9..12 profile b of the same.  Returns the length. */
int orc_sim_table(const orc_sim *s, int which, double *out) {
  const double *src = NULL;
  int len = 0;
  const profile *ps[4] = {&s->pRint, &s->pRhalf, &s->pZint, &s->pZhalf};
  if (which >= 0 && which <= 4) {
    const double *t[5] = {s->rho, s->lamU, s->muU, s->rhoZ, s->muZ};
    src = t[which];
    len = s->nz;
  } else if (which >= 5 && which <= 12) {
    const profile *p = ps[(which - 5) % 4];
    src = which < 9 ? p->a : p->b;
    len = (which - 5) % 4 < 2 ? s->nr : s->nz;
  }
  if (src && out) memcpy(out, src, sizeof(double) * (size_t)len);
  return len;
}

int orc_sim_profile_first(const orc_sim *s, int which) {
  const profile *ps[4] = {&s->pRint, &s->pRhalf, &s->pZint, &s->pZhalf};
  return ps[which & 3]->first;
}

void orc_sim_memcoef(const orc_sim *s, double *decay, double *drive) {
  *decay = s->memDecay;
  *drive = s->memDrive;
}

void orc_sim_drive(orc_sim *s, const double *force, int64_t n) {
  free(s->force);
  s->force = NULL;
  s->nforce = n > 0 ? n : 0;
  if (n > 0) {
    s->force = malloc(sizeof(double) * (size_t)n);
    memcpy(s->force, force, sizeof(double) * (size_t)n);
  }
}

static inline double force_at(const orc_sim *s, int n) {
  if (n < 0 || n >= s->nforce) return 0;
  return s->force[n];
}

void orc_sim_column(const orc_sim *s, double r, int32_t *col, double *snapped) {
  int i = (int)round(r / s->dr);
  if (i > s->nr - 2) i = s->nr - 2;
  if (i < 0) i = 0;
  *col = i;
  *snapped = (double)i * s->dr;
}

/* Each row of a pass of velocities or stresses writes only its own cells,
 * so the passes may be split over threads by row (the parallel
 * build, liboracle_fdtd_omp.so, used to write long golden runs); every value
 * is computed by the same expression either way, so both builds give the
 * same bits. */
#ifdef _OPENMP
#define ORC_ROWS _Pragma("omp parallel for schedule(static)")
#else
#define ORC_ROWS
#endif

static void velocities(orc_sim *s) {
  int nr = s->nr, nz = s->nz;
  double dr = s->dr, dz = s->dz, dt = s->dt;
  ORC_ROWS
  for (int j = 0; j < nz - 1; j++) {
    int zp = j >= s->pZint.first;
    size_t row = (size_t)j * nr;
    double rho = s->rho[j];
    for (int i = 0; i < nr - 1; i++) {
      size_t k = row + i;
      double dtrr = (s->trr[k + 1] - s->trr[k]) / dr;
      if (i >= s->pRhalf.first) dtrr += pupdate(&s->pRhalf, &s->psiTrrR[k], i, dtrr);
      double dtrz;
      if (j == 0)
        dtrz = 2 * s->trz[k] / dz;
      else
        dtrz = (s->trz[k] - s->trz[k - nr]) / dz;
      if (zp) dtrz += pupdate(&s->pZint, &s->psiTrzZ[k], j, dtrz);
      double hoop = 0.5 * ((s->trr[k] - s->tqq[k]) + (s->trr[k + 1] - s->tqq[k + 1])) / (((double)i + 0.5) * dr);
      s->vr[k] += dt / rho * (dtrr + dtrz + hoop);
    }
  }
  ORC_ROWS
  for (int j = 0; j < nz - 1; j++) {
    int zh = j >= s->pZhalf.first;
    size_t row = (size_t)j * nr;
    double rho = s->rhoZ[j];
    for (int i = 0; i < nr - 1; i++) {
      size_t k = row + i;
      double dtzz = (s->tzz[k + nr] - s->tzz[k]) / dz;
      if (zh) dtzz += pupdate(&s->pZhalf, &s->psiTzzZ[k], j, dtzz);
      double radial;
      if (i == 0) {
        radial = 4 * s->trz[k] / dr;
      } else {
        double d = (s->trz[k] - s->trz[k - 1]) / dr;
        if (i >= s->pRint.first) d += pupdate(&s->pRint, &s->psiTrzR[k], i, d);
        radial = d + 0.5 * (s->trz[k] + s->trz[k - 1]) / ((double)i * dr);
      }
      s->vz[k] += dt / rho * (dtzz + radial);
    }
  }
}

/* Sim.relax (fdtd.go:315-320). */
static inline double relax1(double decay, double drive, double *r, double e) {
  double next = decay * *r - drive * e;
  double avg = 0.5 * (*r + next);
  *r = next;
  return avg;
}

/* Extension: one mechanism driven by an already-formed drive term. */
static inline double relaxg(double decay, double *r, double drive_term) {
  double next = decay * *r - drive_term;
  double avg = 0.5 * (*r + next);
  *r = next;
  return avg;
}

static void stresses(orc_sim *s, double dTdt) {
  int nr = s->nr, nz = s->nz, nm = s->nmech;
  double dr = s->dr, dz = s->dz, dt = s->dt;
  size_t n = (size_t)nr * nz;
  ORC_ROWS
  for (int j = 0; j < nz; j++) {
    int zp = j >= s->pZint.first;
    size_t row = (size_t)j * nr;
    double lam = s->lamU[j], mu = s->muU[j];
    double lam2mu = lam + 2 * mu;
    for (int i = 0; i < nr; i++) {
      size_t k = row + i;
      double dvr, vrr;
      if (i == 0) {
        dvr = 2 * s->vr[k] / dr;
        vrr = dvr;
      } else {
        dvr = (s->vr[k] - s->vr[k - 1]) / dr;
        vrr = 0.5 * (s->vr[k] + s->vr[k - 1]) / ((double)i * dr);
      }
      if (i >= s->pRint.first) dvr += pupdate(&s->pRint, &s->psiVrR[k], i, dvr);
      double dvz;
      if (j == 0) {
        double applied = 0.0;
        if (i == 0) applied = dTdt;
        dvz = (applied - lam * (dvr + vrr)) / lam2mu;
      } else {
        dvz = (s->vz[k] - s->vz[k - nr]) / dz;
        if (zp) dvz += pupdate(&s->pZint, &s->psiVzZ[k], j, dvz);
      }
      double err = lam2mu * dvr + lam * (vrr + dvz);
      double eqq = lam * (dvr + dvz) + lam2mu * vrr;
      if (s->lossy && s->scalar) {
        s->trr[k] += dt * (err + relax1(s->memDecay, s->memDrive, &s->rrr[k], err));
        s->tqq[k] += dt * (eqq + relax1(s->memDecay, s->memDrive, &s->rqq[k], eqq));
      } else if (s->lossy) {
        double pb = lam2mu * (dvr + vrr + dvz);
        double sr = 0, sq = 0;
        for (int q = 0; q < nm; q++) {
          double dP = s->driveP[(size_t)q * nz + j], dS = s->driveS[(size_t)q * nz + j];
          sr += relaxg(s->decay[q], &s->rrr[q * n + k], dP * pb - dS * (2 * mu * (vrr + dvz)));
          sq += relaxg(s->decay[q], &s->rqq[q * n + k], dP * pb - dS * (2 * mu * (dvr + dvz)));
        }
        s->trr[k] += dt * (err + sr);
        s->tqq[k] += dt * (eqq + sq);
      } else {
        s->trr[k] += dt * err;
        s->tqq[k] += dt * eqq;
      }
      if (j == 0) continue;
      double ezz = lam * (dvr + vrr) + lam2mu * dvz;
      if (s->lossy && s->scalar) {
        s->tzz[k] += dt * (ezz + relax1(s->memDecay, s->memDrive, &s->rzz[k], ezz));
      } else if (s->lossy) {
        double pb = lam2mu * (dvr + vrr + dvz);
        double sz = 0;
        for (int q = 0; q < nm; q++) {
          double dP = s->driveP[(size_t)q * nz + j], dS = s->driveS[(size_t)q * nz + j];
          sz += relaxg(s->decay[q], &s->rzz[q * n + k], dP * pb - dS * (2 * mu * (dvr + vrr)));
        }
        s->tzz[k] += dt * (ezz + sz);
      } else {
        s->tzz[k] += dt * ezz;
      }
    }
  }
  ORC_ROWS
  for (int j = 0; j < nz - 1; j++) {
    int zh = j >= s->pZhalf.first;
    size_t row = (size_t)j * nr;
    double mu = s->muZ[j];
    for (int i = 0; i < nr - 1; i++) {
      size_t k = row + i;
      double dvzdr = (s->vz[k + 1] - s->vz[k]) / dr;
      if (i >= s->pRhalf.first) dvzdr += pupdate(&s->pRhalf, &s->psiVzR[k], i, dvzdr);
      double dvrdz = (s->vr[k + nr] - s->vr[k]) / dz;
      if (zh) dvrdz += pupdate(&s->pZhalf, &s->psiVrZ[k], j, dvrdz);
      double e = mu * (dvzdr + dvrdz);
      if (s->lossy && s->scalar) {
        s->trz[k] += dt * (e + relax1(s->memDecay, s->memDrive, &s->rrz[k], e));
      } else if (s->lossy) {
        double sz = 0;
        /* muZ is harmonic across an interface; the shear drive uses the
         * mechanism weights of the row above, like muU[j]. */
        for (int q = 0; q < nm; q++) sz += relaxg(s->decay[q], &s->rrz[q * n + k], s->driveS[(size_t)q * nz + j] * e);
        s->trz[k] += dt * (e + sz);
      } else {
        s->trz[k] += dt * e;
      }
    }
  }
}

void orc_sim_step(orc_sim *s) {
  s->tzz[0] = -force_at(s, s->step) / s->srcArea;
  velocities(s);
  stresses(s, (force_at(s, s->step + 1) - force_at(s, s->step)) / s->dt);
  s->step++;
}

/* Checkpoint of the evolving state (the six fields, the eight C-PML psi
 * arrays, the memory variables and the step count) to `path` (save != 0) or
 * back from it, so a run of hours can be continued after an interruption.
 * The file is raw doubles in a fixed order, for the same model only. */
int orc_sim_checkpoint(orc_sim *s, const char *path, int save) {
  FILE *f = fopen(path, save ? "wb" : "rb");
  if (!f) return fail(GF_ERR_ARGUMENT, "oracle: cannot open checkpoint %s", path);
  size_t n = (size_t)s->nr * s->nz;
  size_t nm = s->rrr ? (size_t)s->nmech * n : 0;
  double *arrs[] = {s->vr, s->vz, s->trr, s->tqq, s->tzz, s->trz, s->psiTrrR, s->psiTrzZ, s->psiTrzR, s->psiTzzZ,
                    s->psiVrR, s->psiVzZ, s->psiVzR, s->psiVrZ, s->rrr, s->rqq, s->rzz, s->rrz};
  int64_t head[3] = {s->step, (int64_t)n, (int64_t)nm};
  int ok = 1;
  if (save) {
    ok = fwrite(head, sizeof head, 1, f) == 1;
  } else {
    int64_t h[3];
    ok = fread(h, sizeof h, 1, f) == 1 && h[1] == head[1] && h[2] == head[2];
    if (ok) s->step = (int)h[0];
  }
  for (size_t a = 0; ok && a < sizeof arrs / sizeof arrs[0]; a++) {
    size_t len = a < 14 ? n : nm;
    if (!len || !arrs[a]) continue;
    ok = save ? fwrite(arrs[a], sizeof(double), len, f) == len : fread(arrs[a], sizeof(double), len, f) == len;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? GF_OK : fail(GF_ERR_ARGUMENT, "oracle: checkpoint %s is incomplete or of another model", path);
}

double orc_sim_surface_velocity(const orc_sim *s, int i) {
  double dvr, vrr;
  if (i == 0) {
    dvr = 2 * s->vr[0] / s->dr;
    vrr = dvr;
  } else {
    dvr = (s->vr[i] - s->vr[i - 1]) / s->dr;
    vrr = 0.5 * (s->vr[i] + s->vr[i - 1]) / ((double)i * s->dr);
  }
  double lam = s->lamU[0], mu = s->muU[0];
  double dvz = -lam / (lam + 2 * mu) * (dvr + vrr);
  return s->vz[i] - 0.5 * s->dz * dvz;
}

double orc_sim_radial_velocity(const orc_sim *s, int i) { return s->vr[i]; }

double orc_sim_max_speed(const orc_sim *s) {
  double m = 0.0;
  size_t n = (size_t)s->nr * s->nz;
  for (size_t k = 0; k < n; k++) m = go_max(m, fabs(s->vz[k]));
  for (size_t k = 0; k < n; k++) m = go_max(m, fabs(s->vr[k]));
  return m;
}

/* Read a whole field: 0 vr, 1 vz, 2 trr, 3 tqq, 4 tzz, 5 trz. */
void orc_sim_field(const orc_sim *s, int which, double *out) {
  const double *f[6] = {s->vr, s->vz, s->trr, s->tqq, s->tzz, s->trz};
  memcpy(out, f[which], sizeof(double) * (size_t)s->nr * s->nz);
}

/* -------------------------------------------------------------- run.go */

int orc_ricker(double peak, double f0, double dt, int32_t n, double *out) {
  if (f0 <= 0 || dt <= 0 || n <= 0) return fail(GF_ERR_INVALID, "fdtd: Ricker needs a positive frequency, step and length");
  double t0 = 1.5 / f0;
  double a = 9.869604401089358 * f0 * f0; /* math.Pi*math.Pi folds to this */
  for (int i = 0; i < n; i++) {
    double tau = (double)i * dt - t0;
    out[i] = peak * (1 - 2 * a * tau * tau) * exp(-a * tau * tau);
  }
  orc_err[0] = 0;
  return GF_OK;
}

void orc_ricker_spectrum(double peak, double f0, double f, double *re, double *im) {
  double fr = f / f0;
  double mag = 2 / sqrt(ORC_PI) * fr * fr * exp(-fr * fr) / f0;
  double phase = -2 * ORC_PI * f * 1.5 / f0;
  double pm = peak * mag, c = cos(phase), s = sin(phase);
  *re = pm * c - 0 * s;
  *im = pm * s + 0 * c;
}

/* fdtd.Run.  out_traces is [nranges][steps]. */
int orc_run(const gf_model *m, const double *force, int64_t nforce, const double *ranges, int32_t nranges,
            int32_t steps, double *out_traces, double *out_snapped, double *out_peak, double *out_dt, int32_t *out_nr,
            int32_t *out_nz) {
  char a[64], b[64];
  orc_sim *s;
  int rc = orc_sim_create(m, &s);
  if (rc) return rc;
  if (steps <= 0) {
    orc_sim_destroy(s);
    return fail(GF_ERR_INVALID, "fdtd: a shot needs a positive number of steps");
  }
  orc_sim_drive(s, force, nforce);
  int *cols = malloc(sizeof(int) * (size_t)(nranges > 0 ? nranges : 1));
  for (int i = 0; i < nranges; i++) {
    int32_t c;
    double snapped;
    orc_sim_column(s, ranges[i], &c, &snapped);
    if (snapped >= m->max_range_m) {
      free(cols);
      orc_sim_destroy(s);
      return fail(GF_ERR_INVALID, "fdtd: receiver at %s m is outside the interior (%s m)", gofmt_g(ranges[i], a, 64),
                  gofmt_g(m->max_range_m, b, 64));
    }
    cols[i] = c;
    if (out_snapped) out_snapped[i] = snapped;
  }
  double peak = 0;
  for (int n = 0; n < steps; n++) {
    orc_sim_step(s);
    for (int i = 0; i < nranges; i++) out_traces[(size_t)i * steps + n] = orc_sim_surface_velocity(s, cols[i]);
    if (n % 256 == 0) {
      double p = orc_sim_max_speed(s);
      peak = go_max(peak, p);
      if (isnan(p) || isinf(p)) {
        free(cols);
        orc_sim_destroy(s);
        if (out_peak) *out_peak = peak;
        return fail(GF_ERR_DIVERGED, "fdtd: diverged at step %d of %d", n, steps);
      }
    }
  }
  if (out_peak) *out_peak = peak;
  if (out_dt) *out_dt = s->dt;
  if (out_nr) *out_nr = s->nr;
  if (out_nz) *out_nz = s->nz;
  free(cols);
  orc_sim_destroy(s);
  orc_err[0] = 0;
  return GF_OK;
}

/* ------------------------------------------------------ CPU baseline pool */

typedef struct {
  const gf_model *m;
  const double *force;
  int64_t nforce;
  const double *ranges;
  int32_t nranges, steps;
  double *traces;
  int status;
} orc_job;

typedef struct {
  orc_job *jobs;
  int njobs;
  int next;
  pthread_mutex_t mu;
} orc_pool;

static void *pool_worker(void *arg) {
  orc_pool *p = arg;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    int i = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (i >= p->njobs) return NULL;
    orc_job *j = &p->jobs[i];
    j->status = orc_run(j->m, j->force, j->nforce, j->ranges, j->nranges, j->steps, j->traces, NULL, NULL, NULL,
                        NULL, NULL);
  }
}

/* Independent shots, one per thread, over `threads` workers: the CPU path
 * shaped like cmd/geobank's GOMAXPROCS pool (cmd/geobank/main.go:128-167).
 * traces[i] points at shot i's [nranges][steps] output. */
int orc_run_many(const gf_model *models, int32_t nshots, const double *const *forces, const int64_t *nforces,
                 const double *ranges, int32_t nranges, const int32_t *steps, double *const *traces, int32_t threads,
                 int32_t *statuses) {
  orc_pool p;
  p.jobs = calloc((size_t)nshots, sizeof(orc_job));
  p.njobs = nshots;
  p.next = 0;
  pthread_mutex_init(&p.mu, NULL);
  for (int i = 0; i < nshots; i++) {
    p.jobs[i].m = &models[i];
    p.jobs[i].force = forces[i];
    p.jobs[i].nforce = nforces[i];
    p.jobs[i].ranges = ranges;
    p.jobs[i].nranges = nranges;
    p.jobs[i].steps = steps[i];
    p.jobs[i].traces = traces[i];
  }
  if (threads < 1) threads = 1;
  pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, pool_worker, &p);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  int rc = GF_OK;
  for (int i = 0; i < nshots; i++) {
    if (statuses) statuses[i] = p.jobs[i].status;
    if (p.jobs[i].status && !rc) rc = p.jobs[i].status;
  }
  free(th);
  free(p.jobs);
  pthread_mutex_destroy(&p.mu);
  return rc;
}
