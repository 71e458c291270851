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
// Writes a captured graph's base step v and the buffers holding its state
// (wavefields buf, memory variables mbuf): p[0..2].
hipError_t launch_set_base(int32_t *p, int32_t v, int32_t buf, int32_t mbuf, hipStream_t st);
hipError_t launch_build_tables(const TableParams &P, const LayerParams *d_layers, float *rowc, float *colc,
                               hipStream_t st);
// One launch of a step kernel over `tiles` tiles of the descriptors d.
//   span 1: k_step, one time step; tiles of kOwn columns by th rows (th 1, 2,
//           4, 8: fully unrolled kernels; anything else: the row march).
//   span 2: k_step2, `count` (1 or 2) time steps; tiles of kOwn2 columns by
//           th rows.
// Every descriptor's tiling (tiles_r, tiles_z, th, tile_base) must be the
// one of the kernel launched.  The step is n_off (+ nbase[0]); the wavefields
// are read from buffer rb_off (^ nbase[1]) and written to the other one; the
// memory variables and stress psi are in buffer mb_off (^ nbase[2]), updated
// in place by span 1 and written to the other buffer by span 2.  waves:
// wavefronts (tiles) per workgroup, 1, 2 or 4.  tile_model maps each tile to
// its model (null: every tile is model 0), padded to whole groups of 4.
struct StepLaunch {
  int nmech = 0;
  bool scalar = true;
  int span = 1, th = 1, waves = 4;
  const GridDesc *d = nullptr;
  const uint16_t *tile_model = nullptr;
  int64_t tiles = 0;
  int32_t n_off = 0, rb_off = 0, mb_off = 0;
  const int32_t *nbase = nullptr;
  int32_t probe = 0, count = 1;
  // span 2 only: the tiles as two lists of tile indices, those for which
  // step2_interior holds (list_in, n_in) and the others (list_bd, n_bd), run
  // as two launches; null: tiles 0 .. tiles as one launch
  const uint32_t *list_in = nullptr, *list_bd = nullptr;
  int64_t n_in = 0, n_bd = 0;
};

// The lists pay only when the other tiles are enough to fill the GPU by
// themselves (their launch runs after the interior one): at least four
// wavefronts per CU.  Below that the one launch hides them among the interior
// tiles (a 1.2 GB single grid with ~900 of them: 691 us per two steps as one
// launch, 600 + 113 us as two).
constexpr int64_t kSplitMinBoundaryTiles = 4096;
hipError_t launch_step(const StepLaunch &L, hipStream_t st);
hipError_t launch_dft_f32(const float *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st);
hipError_t launch_dft_f64(const double *x, int64_t ld, int32_t rows, int32_t n, double dt, double t0,
                          const double *freqs, int32_t nbins, double *out, hipStream_t st);
hipError_t launch_surface_velocity(const GridDesc *d, int32_t buf, int32_t col, float *out, hipStream_t st);
hipError_t launch_max_speed(const GridDesc *d, int32_t buf, uint32_t *slot, int64_t cells, hipStream_t st);

}  // namespace gfk
