// layout.h — grid geometry and the sizes of every device array of one model.
//
// Shared by the device allocator (api.cpp) and by anything that needs to
// reproduce the exact addressing of the step kernel: every field is a padded
// plane with kGuardRows zero rows above and below and kGuardCols zero floats
// before and after, so that the halo lanes and rows of a tile read zeros
// outside the grid instead of branching.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <algorithm>

#include "model.h"
#include "stencil.h"

namespace gfl {

struct Layout {
  size_t field_floats = 0;   // one padded field plane
  size_t field_origin = 0;   // offset of element (0, 0) inside it
  size_t rs = 0;             // floats of one r-strip psi array
  size_t zs = 0;             // floats of one z-strip psi array
  size_t mem_floats = 0;     // one buffer of memory variables (4 components x nmech planes)
  size_t rowc_floats = 0;
  size_t colc_floats = 0;
  int nmech_table = 0;       // GSLS drive rows in the row table
};

// Owned columns of a tile of the step kernel of `span` time steps.
inline int tile_cols(int span) { return span == 2 ? gfk::kOwn2 : gfk::kOwn; }

// The descriptor's tiling for the kernel of `span` steps and `th` rows.
inline void set_tiling(gfk::GridDesc &g, int span, int th) {
  g.th = th;
  g.tiles_r = (g.nr + tile_cols(span) - 1) / tile_cols(span);
  g.tiles_z = (g.nz + th - 1) / th;
}

// Tile rows for a grid (see choose_tile_rows in api.cpp for the heuristic).
inline void plan(const gfm::Resolved &R, int span, int tile_rows, gfk::GridDesc &g, Layout &L) {
  g.nr = R.nr;
  g.nz = R.nz;
  set_tiling(g, span, tile_rows);
  // Halo lanes read columns up to tiles_r*kOwn (one-step tiles) or
  // tiles_r*kOwn2 + 1 (two-step tiles), and columns -2 and -1, which are the
  // last two of the row above: keep all of them inside a row and past nr, so
  // they read zeros, whichever kernel runs.
  const size_t w1 = (size_t)((R.nr + gfk::kOwn - 1) / gfk::kOwn) * gfk::kOwn;
  const size_t w2 = (size_t)((R.nr + gfk::kOwn2 - 1) / gfk::kOwn2) * gfk::kOwn2;
  g.pitch = (int)((std::max(w1, w2) + 2 + 63) / 64 * 64);
  g.first_rint = R.first_rint;
  g.first_rhalf = R.first_rhalf;
  g.first_zint = R.first_zint;
  g.first_zhalf = R.first_zhalf;
  g.r0 = std::min(R.first_rint, R.first_rhalf);
  g.rs_pitch = std::max(1, g.pitch - g.r0);
  g.z0 = std::min(R.first_zint, R.first_zhalf);
  g.nmech = R.nmech;
  g.scalar = R.scalar ? 1 : 0;
  g.lossy = R.lossy ? 1 : 0;
  g.invh = (float)(1.0 / R.h);
  g.dt = (float)R.dt;
  g.half_h = (float)(0.5 * R.h);
  g.svc = (float)(-R.lam0 / (R.lam0 + 2 * R.mu0));
  g.mem_decay = (float)R.mem_decay;
  g.mem_drive = (float)R.mem_drive;
  for (int q = 0; q < gfk::kMaxMech; q++) g.decay[q] = (float)R.decay[q];

  const size_t pitch = (size_t)g.pitch, nz = (size_t)g.nz;
  const size_t nzs = nz - std::min(nz, (size_t)std::max(0, g.z0));
  L.field_origin = gfk::kGuardCols + gfk::kGuardRows * pitch;
  L.field_floats = gfk::kGuardCols + (nz + 2 * gfk::kGuardRows) * pitch + gfk::kGuardCols;
  g.field_stride = (int64_t)L.field_floats;
  g.mem_stride = (int64_t)(nz * pitch);
  L.rs = std::max<size_t>(1, nz * (size_t)g.rs_pitch);
  L.zs = std::max<size_t>(1, nzs * pitch);
  L.nmech_table = R.scalar ? 0 : R.nmech;
  L.rowc_floats = (size_t)gfk::kRowW * nz;
  L.colc_floats = (size_t)gfk::kColN * pitch;
  L.mem_floats = R.lossy ? (size_t)4 * R.nmech * nz * pitch : 0;
}

}  // namespace gfl
