/*
 * bank_oracle.c -- FP64 CPU restatement of the reference's bank interpolation,
 * used ONLY as the parity checker of the GPU bank lookup.
 *
 * TEST INFRASTRUCTURE.  Only tests/ and bench.py's cpu_baseline leg load
 * this code.  The product (libgeofdtd) never links or calls it.
 *
 * What it restates (paths relative to the reference repo):
 *   Bank.prepare      internal/bank/interp.go:64-101 (log-magnitude and
 *                     range-unwrapped phase tables, negInf for zero values)
 *   Bank.Response     internal/bank/interp.go:29-54 (bin check, bracket,
 *                     linear interpolation of both, cmplx.Rect)
 *   Bank.bracket      internal/bank/format.go:295-309
 *   RangeGrid.Spacing internal/bank/format.go:60-66
 *
 * cmplx.Abs is Go's Hypot (goarith.h); math.Log/Atan2/Exp/Sincos come from
 * libm; math.Round and C's round both round half away from zero.
 *
 * Pinned by tests/test_bank_lookup_oracle.py against the reference's own
 * bank_test.go interpolation tests.
 */
#include <math.h>
#include <stdint.h>

#include "geofdtd.h"
#include "goarith.h"

#define BANK_NEG_INF (-1e308)
#define BANK_PI 3.141592653589793

/* data [count][bins][2] float32 -> logmag, phase [count][bins]. */
void orc_bank_prepare(int count, int bins, const float *data, double *logmag, double *phase) {
  for (int k = 0; k < bins; k++) {
    double prev = 0, unwrapped = 0;
    for (int i = 0; i < count; i++) {
      long j = 2L * ((long)i * bins + k);
      double re = (double)data[j], im = (double)data[j + 1];
      long idx = (long)i * bins + k;
      double mag = go_hypot(re, im);
      if (mag == 0) {
        logmag[idx] = BANK_NEG_INF;
        phase[idx] = unwrapped;
        continue;
      }
      logmag[idx] = log(mag);
      double w = atan2(im, re);
      if (i == 0) {
        unwrapped = w;
      } else {
        double d = w - prev;
        d -= 2 * BANK_PI * round(d / (2 * BANK_PI));
        unwrapped += d;
      }
      prev = w;
      phase[idx] = unwrapped;
    }
  }
}

/* Bank.Response at (r, bin) from prepared tables: GF_OK or GF_ERR_INVALID
 * with the reference's message. */
int orc_bank_response(double r_min, double r_max, int count, int bins, const double *logmag, const double *phase,
                      double r, int bin, double *re, double *im) {
  char a[40], b[40], c[40];
  if (bin < 0 || bin >= bins) return fail(GF_ERR_INVALID, "bank: bin %d out of range [0, %d)", bin, bins);
  if (r < r_min || r > r_max)
    return fail(GF_ERR_INVALID, "bank: range %s m is outside the bank's %s to %s m", gofmt_g(r, a, 40),
                gofmt_g(r_min, b, 40), gofmt_g(r_max, c, 40));
  int lo = 0;
  double frac = 0;
  if (count > 1) {
    double spacing = (r_max - r_min) / (double)(count - 1);
    double x = (r - r_min) / spacing;
    lo = (int)floor(x);
    if (lo >= count - 1) {
      lo = count - 2;
      frac = 1;
    } else {
      frac = x - (double)lo;
    }
  }
  long i0 = (long)lo * bins + bin;
  long i1 = (long)(lo + 1) * bins + bin;
  if (count == 1) i1 = i0;
  if (logmag[i0] == BANK_NEG_INF || logmag[i1] == BANK_NEG_INF) {
    *re = 0;
    *im = 0;
    return GF_OK;
  }
  double lm = logmag[i0] + frac * (logmag[i1] - logmag[i0]);
  double ph = phase[i0] + frac * (phase[i1] - phase[i0]);
  double mag = exp(lm);
  *re = mag * cos(ph);
  *im = mag * sin(ph);
  return GF_OK;
}

/* Bank.Response at every bin for each of nq ranges (all inside the bank):
 * out [nq][bins][2].  The CPU baseline of the batched lookup: one call per
 * thread over its share of the ranges. */
int orc_bank_spectra(double r_min, double r_max, int count, int bins, const double *logmag, const double *phase,
                     const double *r, int nq, double *out) {
  for (int q = 0; q < nq; q++)
    for (int k = 0; k < bins; k++) {
      long o = 2L * ((long)q * bins + k);
      int rc = orc_bank_response(r_min, r_max, count, bins, logmag, phase, r[q], k, &out[o], &out[o + 1]);
      if (rc) return rc;
    }
  return GF_OK;
}
