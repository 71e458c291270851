/*
 * goarith.h -- Go's arithmetic and formatting, restated in C for the FP64
 * oracle (test infrastructure only; nothing in the product includes it).
 *
 * Go's complex division (Smith's algorithm, runtime/complex.go), cmplx.Sqrt,
 * math.Hypot and math.Max are restated rather than taken from libm, since
 * their rounding and special cases differ from C99's; gofmt_g is Go's %g.
 */
#ifndef ORACLE_GOARITH_H
#define ORACLE_GOARITH_H

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The calling thread's last error, and the setter that fills it. */
extern __thread char orc_err[512];
int orc_fail(int code, const char *fmt, ...);
#define fail orc_fail

#define ORC_UNUSED __attribute__((unused))

/* Go's %g (strconv.FormatFloat(x, 'g', -1, 64)): the shortest digit string
 * that reads back exactly, in exponent form only when the decimal exponent is
 * below -4 or at least 6 -- unlike C's %g, whose switch point is the digit
 * count. */
ORC_UNUSED static const char *gofmt_g(double x, char *buf, size_t n) {
  int p = 1;
  for (; p < 17; p++) {
    snprintf(buf, n, "%.*e", p - 1, x);
    if (strtod(buf, NULL) == x) break;
  }
  snprintf(buf, n, "%.*e", p - 1, x);
  if (x != 0 && isfinite(x)) {
    int e = atoi(strchr(buf, 'e') + 1);
    if (e >= -4 && e < 6) snprintf(buf, n, "%.*f", p - 1 - e > 0 ? p - 1 - e : 0, x);
  } else if (x == 0) {
    snprintf(buf, n, "%s", signbit(x) ? "-0" : "0");
  }
  return buf;
}

/* ------------------------------------------------- Go arithmetic helpers */

typedef struct {
  double re, im;
} cx;

/* runtime/complex.go complex128div (Smith 1962). */
ORC_UNUSED static cx go_cdiv(cx n, cx m) {
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
  cx r = {e, f};
  return r;
}

/* math.Hypot (math/hypot.go; the amd64 assembly computes the same). */
ORC_UNUSED static double go_hypot(double p, double q) {
  p = fabs(p);
  q = fabs(q);
  if (isinf(p) || isinf(q)) return INFINITY;
  if (isnan(p) || isnan(q)) return NAN;
  if (p < q) {
    double t = p;
    p = q;
    q = t;
  }
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

/* math/cmplx.Sqrt. */
ORC_UNUSED static cx go_csqrt(cx x) {
  cx r;
  if (x.im == 0) {
    if (x.re == 0) {
      r.re = 0;
      r.im = x.im;
      return r;
    }
    if (x.re < 0) {
      r.re = 0;
      r.im = copysign(sqrt(-x.re), x.im);
      return r;
    }
    r.re = sqrt(x.re);
    r.im = x.im;
    return r;
  }
  if (x.re == 0) {
    if (x.im < 0) {
      double t = sqrt(-0.5 * x.im);
      r.re = t;
      r.im = -t;
      return r;
    }
    double t = sqrt(0.5 * x.im);
    r.re = t;
    r.im = t;
    return r;
  }
  double a = x.re, b = x.im, scale;
  if (fabs(a) > 4 || fabs(b) > 4) {
    a *= 0.25;
    b *= 0.25;
    scale = 2;
  } else {
    a *= 1.8014398509481984e16;
    b *= 1.8014398509481984e16;
    scale = 7.450580596923828125e-9;
  }
  double rr = go_hypot(a, b), t;
  if (a > 0) {
    t = sqrt(0.5 * rr + 0.5 * a);
    rr = scale * fabs((0.5 * b) / t);
    t *= scale;
  } else {
    rr = sqrt(0.5 * rr - 0.5 * a);
    t = scale * fabs((0.5 * b) / rr);
    rr *= scale;
  }
  if (b < 0) {
    r.re = t;
    r.im = -rr;
  } else {
    r.re = t;
    r.im = rr;
  }
  return r;
}

/* math.Max: NaN wins, then +Inf. */
ORC_UNUSED static double go_max(double x, double y) {
  if (isinf(x) && x > 0) return x;
  if (isinf(y) && y > 0) return y;
  if (isnan(x) || isnan(y)) return NAN;
  if (x == 0 && x == y) return signbit(x) ? y : x;
  return x > y ? x : y;
}

#endif
