// Host check of the device e^x and sin/cos (csrc/fastmath.h) against the C
// library: prints the worst error of each in units in the last place of the
// library's result (exp: relative; sin and cos: absolute, in ulps of 1, since
// near their zeros a relative measure says nothing about a Taylor series),
// then the special values.
#include <cmath>
#include <cstdio>
#include <random>

#include "fastmath.h"

static double ulp(double x) { return std::nextafter(std::fabs(x), INFINITY) - std::fabs(x); }

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 2000000;
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> ue(-700, 700), up(-1e4, 1e4), us(-8, 8);
  double we = 0, ws = 0, wc = 0;
  for (long i = 0; i < n; i++) {
    const double x = ue(g);
    const double r = std::exp(x);
    we = std::fmax(we, std::fabs(gff::exp_reduced(x) - r) / ulp(r));
    const double p = (i & 1) ? up(g) : us(g);
    double s, c;
    gff::sincos_reduced(p, &s, &c);
    ws = std::fmax(ws, std::fabs(s - std::sin(p)) / ulp(1.0));
    wc = std::fmax(wc, std::fabs(c - std::cos(p)) / ulp(1.0));
  }
  double s0, c0;
  gff::sincos_reduced(0.0, &s0, &c0);
  printf("exp_ulp %.3f\nsin_ulp1 %.3f\ncos_ulp1 %.3f\n", we, ws, wc);
  printf("exp0 %.17g\nexp_m746 %.17g\nexp_709 %.17g\nsin0 %.17g\ncos0 %.17g\n", gff::exp_reduced(0.0),
         gff::exp_reduced(-746.5), gff::exp_reduced(709.0) / std::exp(709.0), s0, c0);
  // quadrants: x = q pi/2 + 0.3 for q = -5..5 against the library
  double wq = 0;
  for (int q = -5; q <= 5; q++) {
    const double x = q * 1.5707963267948966 + 0.3;
    double s, c;
    gff::sincos_reduced(x, &s, &c);
    wq = std::fmax(wq, std::fmax(std::fabs(s - std::sin(x)), std::fabs(c - std::cos(x))) / ulp(1.0));
  }
  printf("quadrant_ulp1 %.3f\n", wq);
  return 0;
}
