// model.h — FP64 host arithmetic that turns a gf_model into the scalars the
// device tables are built from: validation, grid size, time step, relaxation
// constants, C-PML parameters.  Mirrors internal/fdtd/model.go, fdtd.go:45-115,
// pml.go and internal/visco/visco.go with Go's rounding (Smith complex
// division, cmplx.Sqrt, math.Hypot, no fused multiply-add), so that dt and the
// grid extents agree with the reference to the last bit.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "coef.h"

namespace gfm {

// Everything fdtd.New derives from a Model before allocating.
struct Resolved {
  std::vector<gf_layer> stack;
  int nr = 0, nz = 0, pml = 0;
  double h = 0, dt = 0, vmax = 0, dominant = 0, refl = 0;
  double max_range = 0, depth = 0;
  double src_area = 0;
  bool lossy = false;   // any memory variables
  bool scalar = true;   // the reference's single scalar SLS
  int nmech = 0;        // 0 elastic, else mechanisms per component
  double mem_decay = 0, mem_drive = 0;  // reference path
  double decay[gfk::kMaxMech] = {0};    // GSLS path
  std::vector<gfk::LayerParams> layers;
  int first_rint = 0, first_rhalf = 0, first_zint = 0, first_zhalf = 0;
  double d0 = 0, amax = 0, pml_thickness = 0;
  double lam0 = 0, mu0 = 0;  // unrelaxed moduli of the surface row
};

// Status and message in the reference's words; returns GF_OK or an error code.
int resolve(const gf_model &m, Resolved &out, std::string &err);

int validate_stack(const gf_layer *st, int n, std::string &err);
int visco_fit(double q, double f0, gf_sls &out, std::string &err);
int spacing_for(const gf_model &m, double max_freq, double cpw, double &out, std::string &err);
void ricker(double peak, double f0, double dt, int32_t n, double *out);
void ricker_spectrum(double peak, double f0, double f, double &re, double &im);

// Go's %g (strconv.FormatFloat(x, 'g', -1, 64)).
std::string go_g(double x);

}  // namespace gfm
