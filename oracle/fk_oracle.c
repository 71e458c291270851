/*
 * fk_oracle.c -- FP64 CPU restatement of the reference's frequency-wavenumber
 * solver, used ONLY as the parity checker of the GPU f-k path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this code, and there only to check or time the GPU
 * path against it.  The product (libgeofdtd) never links or calls it.
 *
 * What it restates, and from where (paths relative to the reference repo):
 *   Medium defaults, refFreq, qOf, poleQ, complexVelocity, propsOf
 *                                       internal/fk/fk.go:35-125
 *   props.eigen, props.propagator, invert
 *                                       internal/fk/fk.go:127-214
 *   referenceModulus, MaxLayerPhase, checkLayerPhase
 *                                       internal/fk/fk.go:216-268
 *   SurfaceResponse                     internal/fk/fk.go:270-307
 *   Integration defaults, StaticCoefficient, VerticalDisplacement
 *                                       internal/fk/fk.go:309-431
 *   Grid, GridFor, VerticalDisplacementMulti
 *                                       internal/fk/fk.go:433-550
 *   cmat.mul, column, solve             internal/fk/cmat.go:14-96
 *   synthesise (the spectra before the inverse transform)
 *                                       cmd/geofdtd/main.go:168-200
 *
 * Operand order follows the Go expressions, and the file is compiled with
 * -ffp-contract=off (Go on amd64 does not fuse multiply-add).  Complex
 * products are written out as Go computes them; division, Sqrt and Hypot are
 * Go's (goarith.h).  math.Exp/Sincos/Pow/Atan/J0 come from libm (glibc's j0,
 * like Go's math.J0, follows FreeBSD's e_j0.c), so results can differ from
 * Go's in the last bits.
 *
 * Pinned by tests/test_fk_oracle.py against the closed forms and properties
 * the reference's own f-k tests assert (fk_test.go, multi_test.go,
 * stability_test.go, causality_test.go).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "geofdtd.h"
#include "goarith.h"

#define FK_PI 3.141592653589793
#define FK_MAX_LAYER_PHASE 16.0

double orc_sls_q(gf_sls s, double f);
double orc_sls_relaxed_velocity(gf_sls s, double v, double fref);
void orc_sls_velocity(gf_sls s, double v, double fref, double f, double *re, double *im);
void orc_ricker_spectrum(double peak, double f0, double f, double *re, double *im);

/* ------------------------------------------------------ complex helpers */

static cx C(double re, double im) {
  cx r = {re, im};
  return r;
}
static cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static cx cneg(cx a) { return C(-a.re, -a.im); }
/* Go's complex128 product: (ac - bd) + (ad + bc)i. */
static cx cmul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
/* math/cmplx.Exp for finite arguments. */
static cx cexp_go(cx x) {
  double r = exp(x.re);
  double s = sin(x.im), c = cos(x.im);
  return C(r * c, r * s);
}
static double go_min(double x, double y) {
  if (isinf(x) && x < 0) return x;
  if (isinf(y) && y < 0) return y;
  if (isnan(x) || isnan(y)) return NAN;
  if (x == 0 && x == y) return signbit(x) ? x : y;
  return x < y ? x : y;
}

typedef struct {
  cx m[4][4];
} cmat;
typedef struct {
  cx v[4];
} cvec;

/* cmat.mul (cmat.go:16-28). */
static cmat mat_mul(const cmat *a, const cmat *b) {
  cmat out;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      cx s = C(0, 0);
      for (int k = 0; k < 4; k++) s = cadd(s, cmul(a->m[i][k], b->m[k][j]));
      out.m[i][j] = s;
    }
  return out;
}

/* cmat.solve (cmat.go:54-96): Gaussian elimination, partial pivoting. */
static int mat_solve(const cmat *a, cvec b, cvec *out) {
  cmat m = *a;
  cvec x = b;
  for (int col = 0; col < 4; col++) {
    double best = 0;
    int best_row = col;
    for (int r = col; r < 4; r++) {
      double v = go_hypot(m.m[r][col].re, m.m[r][col].im);
      if (v > best) {
        best = v;
        best_row = r;
      }
    }
    if (best == 0) return 0;
    for (int c = 0; c < 4; c++) {
      cx t = m.m[col][c];
      m.m[col][c] = m.m[best_row][c];
      m.m[best_row][c] = t;
    }
    cx t = x.v[col];
    x.v[col] = x.v[best_row];
    x.v[best_row] = t;
    cx p = m.m[col][col];
    for (int r = col + 1; r < 4; r++) {
      cx f = go_cdiv(m.m[r][col], p);
      if (f.re == 0 && f.im == 0) continue;
      for (int c = col; c < 4; c++) m.m[r][c] = csub(m.m[r][c], cmul(f, m.m[col][c]));
      x.v[r] = csub(x.v[r], cmul(f, x.v[col]));
    }
  }
  for (int i = 3; i >= 0; i--) {
    cx s = x.v[i];
    for (int j = i + 1; j < 4; j++) s = csub(s, cmul(m.m[i][j], out->v[j]));
    out->v[i] = go_cdiv(s, m.m[i][i]);
  }
  return 1;
}

/* invert (fk.go:199-214). */
static int mat_invert(const cmat *a, cmat *out) {
  for (int j = 0; j < 4; j++) {
    cvec e = {{C(0, 0), C(0, 0), C(0, 0), C(0, 0)}};
    e.v[j] = C(1, 0);
    cvec col;
    if (!mat_solve(a, e, &col)) return 0;
    for (int i = 0; i < 4; i++) out->m[i][j] = col.v[i];
  }
  return 1;
}

/* ------------------------------------------------------------- the medium */

static double fk_ref_freq(const gf_medium *m) { return m->ref_freq_hz > 0 ? m->ref_freq_hz : 30.0; }

static double fk_q_of(const gf_medium *m, const gf_layer *l) {
  if (l->qs > 0) return l->qs;
  if (m->default_q > 0) return m->default_q;
  return 30.0;
}

static double fk_pole_q(const gf_medium *m, double f) {
  if (m->has_relax) return orc_sls_q(m->relax, f);
  return fk_q_of(m, &m->stack[m->n_layers - 1]);
}

/* complexVelocity (fk.go:92-104). */
static cx fk_complex_velocity(const gf_medium *m, double v, double q, double f) {
  if (m->has_relax) {
    if (f <= 0) return C(orc_sls_relaxed_velocity(m->relax, v, fk_ref_freq(m)), 0);
    cx r;
    orc_sls_velocity(m->relax, v, fk_ref_freq(m), f, &r.re, &r.im);
    return r;
  }
  if (q <= 0 || f <= 0) return C(v, 0);
  double gam = atan(1 / q) / FK_PI;
  double c0 = v * cos(FK_PI * gam / 2);
  /* cmplx.Pow of a positive real to a real power is math.Pow, phase 0. */
  cx p = C(pow(f / fk_ref_freq(m), gam), 0);
  cx e = cexp_go(C(0, FK_PI * gam / 2));
  return cmul(cmul(C(c0, 0), p), e);
}

void orc_fk_complex_velocity(const gf_medium *m, double v, double q, double f, double *re, double *im) {
  cx c = fk_complex_velocity(m, v, q, f);
  *re = c.re;
  *im = c.im;
}

typedef struct {
  cx mu, lambda;
  double rho;
  cx alpha, beta;
} props;

/* propsOf (fk.go:114-125). */
static props fk_props(const gf_medium *m, const gf_layer *l, double f) {
  double q = fk_q_of(m, l);
  cx a = fk_complex_velocity(m, l->vp_mps, 2.25 * q, f);
  cx b = fk_complex_velocity(m, l->vs_mps, q, f);
  double rho = l->density_kgm3;
  props p;
  p.mu = cmul(cmul(C(rho, 0), b), b);
  p.lambda = csub(cmul(cmul(C(rho, 0), a), a), cmul(C(2, 0), p.mu));
  p.rho = rho;
  p.alpha = a;
  p.beta = b;
  return p;
}

/* props.eigen (fk.go:135-167): eigenvalues and eigenvector columns in the
 * order (P down, SV down, P up, SV up). */
static void fk_eigen(const props *p, double k, double omega, double mu_ref, cx lam[4], cmat *vec) {
  cx kc = C(k, 0), w2 = C(omega * omega, 0);
  cx kk = cmul(kc, kc);
  cx ra = go_csqrt(csub(C(1, 0), go_cdiv(w2, cmul(cmul(kk, p->alpha), p->alpha))));
  cx rb = go_csqrt(csub(C(1, 0), go_cdiv(w2, cmul(cmul(kk, p->beta), p->beta))));
  if (ra.re < 0) ra = cneg(ra);
  if (rb.re < 0) rb = cneg(rb);
  cx rc2 = go_cdiv(cmul(C(p->rho, 0), w2), kk);
  cx mr = C(mu_ref, 0), mu = p->mu;
  lam[0] = cneg(ra);
  lam[1] = cneg(rb);
  lam[2] = ra;
  lam[3] = rb;
  cx two_mu = cmul(C(2, 0), mu);
  cx m2mu_mr = go_cdiv(cmul(C(-2, 0), mu), mr); /* -2*mu/mr */
  cx p2mu_mr = go_cdiv(two_mu, mr);             /* 2*mu/mr */
  cx pd[4] = {C(1, 0), cneg(ra), cmul(m2mu_mr, ra), go_cdiv(csub(two_mu, rc2), mr)};
  cx pu[4] = {C(1, 0), ra, cmul(p2mu_mr, ra), go_cdiv(csub(two_mu, rc2), mr)};
  cx sd[4] = {rb, C(-1, 0), go_cdiv(csub(rc2, two_mu), mr), cmul(p2mu_mr, rb)};
  cx su[4] = {cneg(rb), C(-1, 0), go_cdiv(csub(rc2, two_mu), mr), cmul(m2mu_mr, rb)};
  for (int i = 0; i < 4; i++) {
    vec->m[i][0] = pd[i];
    vec->m[i][1] = sd[i];
    vec->m[i][2] = pu[i];
    vec->m[i][3] = su[i];
  }
}

/* props.propagator (fk.go:176-196). */
static int fk_propagator(const props *p, double k, double omega, double mu_ref, double thickness, cmat *out) {
  cx lam[4];
  cmat e;
  fk_eigen(p, k, omega, mu_ref, lam, &e);
  cx zeta = C(k * thickness, 0);
  double max_re = 0;
  for (int i = 0; i < 4; i++) max_re = go_max(max_re, lam[i].re);
  cmat d;
  memset(&d, 0, sizeof d);
  for (int i = 0; i < 4; i++) d.m[i][i] = cexp_go(cmul(csub(lam[i], C(max_re, 0)), zeta));
  cmat inv;
  if (!mat_invert(&e, &inv)) return 0;
  cmat ed = mat_mul(&e, &d);
  *out = mat_mul(&ed, &inv);
  return 1;
}

/* referenceModulus (fk.go:216-223). */
static double fk_reference_modulus(const gf_medium *m) {
  const gf_layer *s = m->stack;
  double r = s[0].density_kgm3 * s[0].vs_mps * s[0].vs_mps;
  for (int i = 0; i < m->n_layers; i++) r = go_max(r, s[i].density_kgm3 * s[i].vs_mps * s[i].vs_mps);
  return r;
}

/* layer.Stack.Validate (layer.go:53-73) and VelocityBounds (layer.go:86-93). */
static int fk_validate(const gf_medium *m) {
  char b1[40], b2[40];
  if (m->n_layers <= 0) return fail(GF_ERR_INVALID, "layer: a stack needs at least a half-space");
  for (int i = 0; i < m->n_layers; i++) {
    const gf_layer *l = &m->stack[i];
    if (l->vs_mps <= 0) return fail(GF_ERR_INVALID, "layer %d: Vs must be positive, got %s m/s", i, gofmt_g(l->vs_mps, b1, 40));
    if (l->vp_mps <= 0) return fail(GF_ERR_INVALID, "layer %d: Vp must be positive, got %s m/s", i, gofmt_g(l->vp_mps, b1, 40));
    if (l->density_kgm3 <= 0)
      return fail(GF_ERR_INVALID, "layer %d: density must be positive, got %s kg/m^3", i, gofmt_g(l->density_kgm3, b1, 40));
    if (l->vp_mps <= 1.4142135623730951 * l->vs_mps)
      return fail(GF_ERR_INVALID, "layer %d: Vp=%s must exceed sqrt(2)*Vs=%s for a non-negative Poisson ratio", i,
                  gofmt_g(l->vp_mps, b1, 40), gofmt_g(1.4142135623730951 * l->vs_mps, b2, 40));
    if (i < m->n_layers - 1 && l->thickness_m <= 0)
      return fail(GF_ERR_INVALID, "layer %d: thickness must be positive, got %s m", i, gofmt_g(l->thickness_m, b1, 40));
  }
  return GF_OK;
}

static double fk_slowest(const gf_medium *m) {
  double s = m->stack[0].vs_mps;
  for (int i = 0; i < m->n_layers; i++) s = fmin(s, m->stack[i].vs_mps);
  return s;
}

/* Go's %.Nf: C's, with Go's spelling of the infinities and NaN. */
static const char *gofmt_f(double x, int prec, char *buf, size_t n) {
  if (isnan(x))
    snprintf(buf, n, "NaN");
  else if (isinf(x))
    snprintf(buf, n, x > 0 ? "+Inf" : "-Inf");
  else
    snprintf(buf, n, "%.*f", prec, x);
  return buf;
}

/* checkLayerPhase (fk.go:253-268). */
int orc_fk_check_layer_phase(const gf_medium *m, double f) {
  if (m->n_layers <= 0) return fail(GF_ERR_INVALID, "layer: a stack needs at least a half-space");
  double slowest = fk_slowest(m);
  double k = 2 * FK_PI * f / (0.87 * slowest);
  for (int i = 0; i < m->n_layers - 1; i++) {
    const gf_layer *l = &m->stack[i];
    double kh = k * l->thickness_m;
    if (kh > FK_MAX_LAYER_PHASE) {
      char a[40], b[40], c[40], d[40], e[40];
      return fail(GF_ERR_INVALID,
                  "fk: layer %d is %s m thick, which is k*h = %s at %s Hz and past the %s this propagator stays "
                  "accurate to; it would return a plausible wrong number rather than fail. Band-limit to %s Hz for "
                  "this stack, or reformulate the solve with compound matrices as internal/propmat does",
                  i, gofmt_g(l->thickness_m, a, 40), gofmt_f(kh, 1, b, 40), gofmt_g(f, c, 40),
                  gofmt_f(FK_MAX_LAYER_PHASE, 0, d, 40),
                  gofmt_f(FK_MAX_LAYER_PHASE * 0.87 * slowest / (2 * FK_PI * l->thickness_m), 0, e, 40));
    }
  }
  return GF_OK;
}

/* SurfaceResponse (fk.go:276-307). */
int orc_fk_surface_response(const gf_medium *m, double k, double f, double *re, double *im) {
  char a[40], b[40];
  if (k <= 0 || f <= 0)
    return fail(GF_ERR_INVALID, "fk: wavenumber and frequency must be positive, got %s and %s", gofmt_g(k, a, 40),
                gofmt_g(f, b, 40));
  int rc = fk_validate(m);
  if (rc) return rc;
  double omega = 2 * FK_PI * f;
  double mu_ref = fk_reference_modulus(m);
  cmat total;
  memset(&total, 0, sizeof total);
  for (int i = 0; i < 4; i++) total.m[i][i] = C(1, 0);
  for (int l = 0; l < m->n_layers - 1; l++) {
    props p = fk_props(m, &m->stack[l], f);
    cmat pm;
    if (!fk_propagator(&p, k, omega, mu_ref, m->stack[l].thickness_m, &pm))
      return fail(GF_ERR_INVALID, "fk: layer propagator is singular at k=%s f=%s", gofmt_g(k, a, 40), gofmt_g(f, b, 40));
    total = mat_mul(&pm, &total);
  }
  props hp = fk_props(m, &m->stack[m->n_layers - 1], f);
  cx lam[4];
  cmat e;
  fk_eigen(&hp, k, omega, mu_ref, lam, &e);
  cx b4 = go_cdiv(C(0, -1), C(2 * FK_PI * mu_ref * k, 0));
  cmat A;
  cvec rhs;
  for (int i = 0; i < 4; i++) {
    A.m[i][0] = total.m[i][0];
    A.m[i][1] = total.m[i][1];
    A.m[i][2] = cneg(e.m[i][0]);
    A.m[i][3] = cneg(e.m[i][1]);
    rhs.v[i] = cmul(cneg(b4), total.m[i][3]);
  }
  cvec x;
  if (!mat_solve(&A, rhs, &x))
    return fail(GF_ERR_INVALID, "fk: surface system is singular at k=%s f=%s", gofmt_g(k, a, 40), gofmt_g(f, b, 40));
  cx u = cmul(C(0, -1), x.v[1]);
  *re = u.re;
  *im = u.im;
  return GF_OK;
}

static int surface(const gf_medium *m, double k, double f, cx *u) { return orc_fk_surface_response(m, k, f, &u->re, &u->im); }

/* StaticCoefficient (fk.go:347-353). */
int orc_fk_static_coefficient(const gf_medium *m, double f, double k_far, double *re, double *im) {
  cx u;
  int rc = surface(m, k_far, f, &u);
  if (rc) return rc;
  cx c = cmul(C(k_far, 0), u);
  *re = c.re;
  *im = c.im;
  return GF_OK;
}

static int samples_of(gf_integration o) { return o.samples > 0 ? o.samples : 20000; }
static double kmax_factor_of(gf_integration o) { return o.kmax_factor > 0 ? o.kmax_factor : 30.0; }

/* VerticalDisplacement (fk.go:378-431). */
int orc_fk_vertical_displacement(const gf_medium *m, double r, double f, gf_integration opt, double *re, double *im) {
  char a[40];
  if (r <= 0) return fail(GF_ERR_INVALID, "fk: range must be positive, got %s m", gofmt_g(r, a, 40));
  if (f <= 0) return fail(GF_ERR_INVALID, "fk: frequency must be positive, got %s Hz", gofmt_g(f, a, 40));
  int rc = orc_fk_check_layer_phase(m, f);
  if (rc) return rc;
  double slowest = fk_slowest(m);
  double k_body = 2 * FK_PI * f / slowest;
  double k_max = kmax_factor_of(opt) * k_body;
  k_max = go_max(k_max, 40 / r);
  cx c;
  rc = orc_fk_static_coefficient(m, f, go_max(200 * k_body, 400 / r), &c.re, &c.im);
  if (rc) return rc;
  int n = samples_of(opt);
  double dk = k_max / (double)n;
  cx sum = cmul(C(0.5 * dk, 0), cneg(c));
  for (int i = 1; i <= n; i++) {
    double k = (double)i * dk;
    cx u;
    rc = surface(m, k, f, &u);
    if (rc) return rc;
    double w = i == n ? 0.5 : 1.0;
    cx rem = csub(cmul(C(k, 0), u), c);
    sum = cadd(sum, cmul(C(w * j0(k * r) * dk, 0), rem));
  }
  cx out = cadd(go_cdiv(c, C(r, 0)), sum);
  *re = out.re;
  *im = out.im;
  return GF_OK;
}

/* GridFor (fk.go:456-489). */
int orc_fk_grid_for(const gf_medium *m, const double *ranges, int n_ranges, double f, gf_integration opt,
                    double *kmax_out, long long *samples_out) {
  char a[40];
  if (n_ranges <= 0) return fail(GF_ERR_INVALID, "fk: no ranges given");
  double r_min = INFINITY, r_max = 0;
  for (int i = 0; i < n_ranges; i++) {
    if (ranges[i] <= 0) return fail(GF_ERR_INVALID, "fk: range must be positive, got %s m", gofmt_g(ranges[i], a, 40));
    r_min = go_min(r_min, ranges[i]);
    r_max = go_max(r_max, ranges[i]);
  }
  if (m->n_layers <= 0) return fail(GF_ERR_INVALID, "layer: a stack needs at least a half-space");
  double slowest = fk_slowest(m);
  double k_body = 2 * FK_PI * f / slowest;
  double k_max = go_max(kmax_factor_of(opt) * k_body, 40 / r_min);
  double dk = 2 * FK_PI / (8 * r_max);
  double q = fk_pole_q(m, f);
  if (k_body > 0 && q > 0) dk = go_min(dk, k_body / (2 * q) / 4);
  /* int(math.Ceil(x)) on amd64: NaN and out-of-range give the minimum int64. */
  double nf = ceil(k_max / dk);
  long long n = (nf >= -9.2e18 && nf <= 9.2e18) ? (long long)nf : (long long)(-9223372036854775807LL - 1);
  if (opt.samples > 0) n = opt.samples;
  if (n < 64) n = 64;
  *kmax_out = k_max;
  *samples_out = n;
  return GF_OK;
}

/* VerticalDisplacementMulti (fk.go:500-550): out [n_ranges][2]. */
int orc_fk_vertical_displacement_multi(const gf_medium *m, const double *ranges, int n_ranges, double f,
                                       gf_integration opt, double *out) {
  int rc = orc_fk_check_layer_phase(m, f);
  if (rc) return rc;
  double kmax;
  long long n;
  rc = orc_fk_grid_for(m, ranges, n_ranges, f, opt, &kmax, &n);
  if (rc) return rc;
  double slowest = fk_slowest(m);
  double k_body = 2 * FK_PI * f / slowest;
  double r_min = INFINITY;
  for (int i = 0; i < n_ranges; i++) r_min = go_min(r_min, ranges[i]);
  cx c;
  rc = orc_fk_static_coefficient(m, f, go_max(200 * k_body, 400 / r_min), &c.re, &c.im);
  if (rc) return rc;
  double dk = kmax / (double)n;
  cx *acc = (cx *)malloc(sizeof(cx) * (size_t)n_ranges);
  if (!acc) return fail(GF_ERR_ARGUMENT, "oracle: out of memory");
  for (int j = 0; j < n_ranges; j++) acc[j] = csub(go_cdiv(c, C(ranges[j], 0)), cmul(C(0.5 * dk, 0), c));
  for (long long i = 1; i <= n; i++) {
    double k = (double)i * dk;
    cx u;
    rc = surface(m, k, f, &u);
    if (rc) {
      free(acc);
      return rc;
    }
    double w = i == n ? 0.5 : 1.0;
    cx rem = csub(cmul(C(k, 0), u), c);
    for (int j = 0; j < n_ranges; j++) acc[j] = cadd(acc[j], cmul(C(w * j0(k * ranges[j]) * dk, 0), rem));
  }
  for (int j = 0; j < n_ranges; j++) {
    out[2 * j] = acc[j].re;
    out[2 * j + 1] = acc[j].im;
  }
  free(acc);
  return GF_OK;
}

/* The spectra synthesise builds before its inverse transform
 * (cmd/geofdtd/main.go:168-195): coeff [n_ranges][nfft/2+1][2], where nfft
 * is dsp.NextPow2(n); returns nfft through *nfft_out. */
int orc_fk_synthesise_coeff(const gf_medium *m, const double *ranges, int n_ranges, double peak, double f0,
                            double band, double dt, double t0, int n, int samples, double *coeff, int *nfft_out) {
  int nfft = 1;
  while (nfft < n) nfft <<= 1;
  *nfft_out = nfft;
  if (!coeff) return GF_OK;
  int nb = nfft / 2 + 1;
  memset(coeff, 0, sizeof(double) * 2 * (size_t)nb * (size_t)n_ranges);
  double fs = 1 / dt;
  double df = fs / (double)nfft;
  double *u = (double *)malloc(sizeof(double) * 2 * (size_t)n_ranges);
  if (!u) return fail(GF_ERR_ARGUMENT, "oracle: out of memory");
  gf_integration opt = {samples, 0};
  for (int b = 1; b <= nfft / 2; b++) {
    double f = (double)b * df;
    if (f > band) break;
    int rc = orc_fk_vertical_displacement_multi(m, ranges, n_ranges, f, opt, u);
    if (rc) {
      free(u);
      return rc;
    }
    cx w;
    orc_ricker_spectrum(peak, f0, f, &w.re, &w.im);
    cx jw = C(0, 2 * FK_PI * f);
    cx shift = cexp_go(C(0, -2 * FK_PI * f * t0));
    for (int i = 0; i < n_ranges; i++) {
      cx v = cmul(cmul(cmul(cmul(C(u[2 * i], u[2 * i + 1]), w), jw), shift), C(fs, 0));
      coeff[2 * ((size_t)i * nb + b)] = v.re;
      coeff[2 * ((size_t)i * nb + b) + 1] = v.im;
    }
  }
  free(u);
  return GF_OK;
}

/* The quadrature of VerticalDisplacementMulti on a given grid (k_max, n) and
 * static-coefficient wavenumber k_far, summed in the order the GPU kernels
 * sum (libgeofdtd's fk.hip): samples in workgroups of 256; within one, for
 * each group of G = min(256, ranges left) ranges, S = 256 / G interleaved
 * slices (sample ii in slice ii mod S) summed in order, then the slices in
 * order; the workgroup partials are folded by 64 lanes (lane l takes every
 * 64th partial from l, then lane sums pairwise at distances 32 .. 1) and
 * added to C/r - (dk/2) C.  Every
 * term is the Go code's w * J0(k r) * dk * (k u - C); only the order of the
 * additions differs from the reference's sequential sum, so that comparing
 * the GPU with this isolates per-term rounding.  out [n_ranges][2]. */
int orc_fk_quadrature_workgroup_order(const gf_medium *m, const double *ranges, int n_ranges, double f, double k_max,
                                      long long n, double k_far, double *out) {
  enum { B = 256 };
  cx c;
  int rc = orc_fk_static_coefficient(m, f, k_far, &c.re, &c.im);
  if (rc) return rc;
  double dk = k_max / (double)n;
  cx *acc = (cx *)malloc(sizeof(cx) * (size_t)n_ranges);
  cx *rem = (cx *)malloc(sizeof(cx) * B);
  double *kk = (double *)malloc(sizeof(double) * B), *ww = (double *)malloc(sizeof(double) * B);
  cx *slice = (cx *)malloc(sizeof(cx) * B);
  if (!acc || !rem || !kk || !ww || !slice) {
    free(acc), free(rem), free(kk), free(ww), free(slice);
    return fail(GF_ERR_ARGUMENT, "oracle: out of memory");
  }
  long long nb = (n + B - 1) / B;
  cx *part = (cx *)malloc(sizeof(cx) * (size_t)nb * (size_t)n_ranges);
  if (!part) {
    free(acc), free(rem), free(kk), free(ww), free(slice);
    return fail(GF_ERR_ARGUMENT, "oracle: out of memory");
  }
  for (long long first = 1; first <= n; first += B) {
    int cnt = (int)((n - first + 1) < B ? (n - first + 1) : B);
    for (int t = 0; t < cnt; t++) {
      long long i = first + t;
      double k = (double)i * dk;
      cx u;
      rc = surface(m, k, f, &u);
      if (rc) {
        free(acc), free(rem), free(kk), free(ww), free(slice);
        return rc;
      }
      kk[t] = k;
      ww[t] = i == n ? 0.5 : 1.0;
      rem[t] = csub(cmul(C(k, 0), u), c);
    }
    for (int g0 = 0; g0 < n_ranges; g0 += B) {
      int G = (n_ranges - g0) < B ? (n_ranges - g0) : B;
      int S = B / G;
      for (int jr = 0; jr < G; jr++) {
        double r = ranges[g0 + jr];
        for (int s = 0; s < S; s++) {
          cx a = C(0, 0);
          for (int ii = s; ii < cnt; ii += S) a = cadd(a, cmul(C(ww[ii] * j0(kk[ii] * r) * dk, 0), rem[ii]));
          slice[s] = a;
        }
        cx tot = slice[0];
        for (int s = 1; s < S; s++) tot = cadd(tot, slice[s]);
        part[(size_t)((first - 1) / B) * n_ranges + g0 + jr] = tot;
      }
    }
  }
  /* The fold: 64 lanes each add every 64th partial in order, then the lane
   * sums are added pairwise at distances 32, 16, ..., 1. */
  for (int j = 0; j < n_ranges; j++) {
    cx lane[64];
    for (int l = 0; l < 64; l++) {
      lane[l] = C(0, 0);
      for (long long b = l; b < nb; b += 64) lane[l] = cadd(lane[l], part[(size_t)b * n_ranges + j]);
    }
    for (int d = 32; d >= 1; d >>= 1)
      for (int l = 0; l < d; l++) lane[l] = cadd(lane[l], lane[l + d]);
    acc[j] = cadd(csub(go_cdiv(c, C(ranges[j], 0)), cmul(C(0.5 * dk, 0), c)), lane[0]);
    out[2 * j] = acc[j].re;
    out[2 * j + 1] = acc[j].im;
  }
  free(acc), free(rem), free(kk), free(ww), free(slice), free(part);
  return GF_OK;
}
