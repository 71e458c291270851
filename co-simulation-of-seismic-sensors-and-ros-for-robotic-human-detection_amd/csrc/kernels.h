// kernels.h — host-visible declarations of the gfx950 kernels' launchers and
// of the parameter blocks the coefficient builder takes.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "coef.h"
#include "stencil.h"

namespace gfk {

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_fill_f32(float *p, int64_t n, float v, hipStream_t st);
hipError_t launch_set_base(int32_t *p, int32_t v, hipStream_t st);
hipError_t launch_build_tables(const TableParams &P, const LayerParams *d_layers, float *rowc, float *colc,
                               hipStream_t st);
// waves: wavefronts (tiles) per workgroup, 1, 2 or 4.
hipError_t launch_step(int nmech, bool scalar, int th, int waves, const GridDesc *d, const uint16_t *tile_model,
                       int64_t tiles, int32_t n_off, const int32_t *nbase, int32_t probe, hipStream_t st);
hipError_t launch_dft_f32(const float *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st);
hipError_t launch_dft_f64(const double *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st);
hipError_t launch_surface_velocity(const GridDesc *d, int32_t buf, int32_t col, float *out, hipStream_t st);
hipError_t launch_max_speed(const GridDesc *d, int32_t buf, uint32_t *slot, int64_t cells, hipStream_t st);

}  // namespace gfk
