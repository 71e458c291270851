// coef.h — per-row and per-column coefficient tables of one model, computed
// in FP64 and rounded once to FP32.  Called per element by k_build_tables on
// the device; plain C++ so the same arithmetic can run on a CPU.
//
// Row j (depth j h) takes the layer containing that depth, an interface node
// belonging to the layer below (model.go:197-209); rhoZ is the arithmetic and
// muZ the harmonic mean with row min(j+1, nz-1) (model.go:182-195).  The
// C-PML profiles are newProfile's (pml.go:29-63) with kappa = 1.
#pragma once

#include <math.h>
#include <stdint.h>

#include "stencil.h"

namespace gfk {

// One layer, with its unrelaxed moduli and GSLS drives evaluated in FP64 on
// the host (visco.Moduli, visco.go:120-124).
struct LayerParams {
  double thickness;
  double rho;
  double mp;  // unrelaxed P modulus (lamU + 2 muU)
  double ms;  // unrelaxed S modulus (muU)
  double drive_p[kMaxMech];
  double drive_s[kMaxMech];
};

struct TableParams {
  int32_t nr, nz, pitch, nlayers;
  int32_t nmech_table;  // GSLS drive rows in the table (0 for the reference form)
  int32_t pad;
  double h, dt;
  double max_range, depth;
  double pml_thickness, d0, amax;  // newProfile's d0 and pi * dominant
};

// newProfile's per-index coefficients (pml.go:40-61).  pow(x, 2) is x*x,
// which is what Go's math.Pow returns for an exponent of two.
GF_HD void pml_coef(int i, double h, double offset, double inner, double thick, double d0, double amax, double dt,
                    double &a, double &b) {
#pragma clang fp contract(off)
  a = 0.0;
  b = 1.0;
  if (thick <= 0) return;
  double s = ((double)i + offset) * h - inner;
  if (s <= 0) return;
  double x = fmin(s / thick, 1.0);
  double d = d0 * (x * x);
  double alpha = amax * (1 - x);
  if (d + alpha <= 0) return;
  double bb = exp(-(d + alpha) * dt);
  a = d * (bb - 1) / (d + alpha);
  b = bb;
}

GF_HD int layer_index(const TableParams &P, const LayerParams *L, double z) {
#pragma clang fp contract(off)
  double top = 0;
  for (int i = 0; i < P.nlayers - 1; i++) {
    top += L[i].thickness;
    if (z < top) return i;
  }
  return P.nlayers - 1;
}

GF_HD double harmonic_mean(double a, double b) {
#pragma clang fp contract(off)
  if (a <= 0 || b <= 0) return 0;
  return 2 * a * b / (a + b);
}

// Row j of the [nz][kRowW] table.
GF_HD void build_row(const TableParams &P, const LayerParams *L, int j, float *rowc) {
#pragma clang fp contract(off)
  float *r = rowc + (size_t)j * kRowW;
  const int k = j + 1 < P.nz - 1 ? j + 1 : P.nz - 1;
  const LayerParams &lj = L[layer_index(P, L, (double)j * P.h)];
  const LayerParams &lk = L[layer_index(P, L, (double)k * P.h)];
  const double rho = lj.rho, mu = lj.ms, lam = lj.mp - 2 * lj.ms;
  const double rhoz = 0.5 * (lj.rho + lk.rho);
  const double muz = harmonic_mean(lj.ms, lk.ms);
  const double l2m = lam + 2 * mu;
  r[kRowCvr] = (float)(P.dt / rho);
  r[kRowCvz] = (float)(P.dt / rhoz);
  r[kRowLam] = (float)lam;
  r[kRowMu] = (float)mu;
  r[kRowL2m] = (float)l2m;
  r[kRowIl2m] = (float)(1.0 / l2m);
  r[kRowMuz] = (float)muz;
  double a, b;
  pml_coef(j, P.h, 0.0, P.depth, P.pml_thickness, P.d0, P.amax, P.dt, a, b);
  r[kRowAzi] = (float)a;
  r[kRowBzi] = (float)b;
  pml_coef(j, P.h, 0.5, P.depth, P.pml_thickness, P.d0, P.amax, P.dt, a, b);
  r[kRowAzh] = (float)a;
  r[kRowBzh] = (float)b;
  for (int q = 0; q < P.nmech_table; q++) {
    r[kRowN + q] = (float)lj.drive_p[q];
    r[kRowN + P.nmech_table + q] = (float)lj.drive_s[q];
  }
}

// Column c of the [kColN][pitch] table.
GF_HD void build_col(const TableParams &P, int c, float *colc) {
#pragma clang fp contract(off)
  const size_t pitch = (size_t)P.pitch;
  colc[kColIri * pitch + c] = c > 0 ? (float)(1.0 / ((double)c * P.h)) : 0.f;
  colc[kColIrh * pitch + c] = (float)(1.0 / (((double)c + 0.5) * P.h));
  double a = 0, b = 1;
  if (c < P.nr) pml_coef(c, P.h, 0.0, P.max_range, P.pml_thickness, P.d0, P.amax, P.dt, a, b);
  colc[kColAri * pitch + c] = (float)a;
  colc[kColBri * pitch + c] = (float)b;
  a = 0;
  b = 1;
  if (c < P.nr) pml_coef(c, P.h, 0.5, P.max_range, P.pml_thickness, P.d0, P.amax, P.dt, a, b);
  colc[kColArh * pitch + c] = (float)a;
  colc[kColBrh * pitch + c] = (float)b;
}

}  // namespace gfk
