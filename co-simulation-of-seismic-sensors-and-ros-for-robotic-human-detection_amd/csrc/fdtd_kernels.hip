// fdtd_kernels.hip — gfx950 kernels of the axisymmetric viscoelastic FDTD.
//
//   k_build_tables   per-row and per-column coefficients, FP64 math rounded
//                    once to FP32 (model.go:157-209, pml.go:29-63)
//   k_step           one fused time step of every tile of every model
//                    (fdtd.go:149-307; receivers fdtd.go:330-342; the
//                    MaxSpeed divergence probe fdtd.go:351-360 folded in)
//   k_surface_velocity, k_max_speed, k_fill_*, k_set_base   small helpers
//
// The stencil is HBM-bound (about one flop per byte), so there is no MFMA
// here: the design goal is that each state array crosses HBM exactly once per
// step.  See stencil.h for the tile march.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "coef.h"
#include "kernels.h"
#include "stencil.h"

namespace gfk {

// ------------------------------------------------------------------ helpers

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f32(float *p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_set_base(int32_t *p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

// --------------------------------------------------------- coefficient build

__global__ void k_build_tables(TableParams P, const LayerParams *__restrict__ L, float *__restrict__ rowc,
                               float *__restrict__ colc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.nz) build_row(P, L, i, rowc);
  if (i < P.pitch) build_col(P, i, colc);
}

// ---------------------------------------------------------------- time step

// Neighbour lanes through DPP wave shifts: a VALU modifier, no LDS crossbar
// round trip.  wave_shr:1 hands lane i the value of lane i-1, wave_shl:1 the
// value of lane i+1; the lane that has no source gets 0 (its value is never
// used: lane 0 needs no left neighbour, lane 63 no right one).
__device__ __forceinline__ float lane_prev(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_next(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}

// Workgroups are dealt to the 8 XCDs round-robin.  Renumber them so that each
// XCD gets one contiguous run of tiles: a tile's halo rows and columns then
// belong to tiles on the same XCD and are shared through that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b & 7, j = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + j;
}

// One wavefront = one tile; one, two or four tiles per workgroup (the
// launch's choice).  Tiles are numbered model by model (GridDesc::tile_base),
// row-major within a model.
//
// TH > 0: the tile is TH rows tall (every model of the launch has g.th == TH)
// and every load of the tile -- TH + 2 rows of the six fields and TH rows of
// memory variables -- is issued before any arithmetic, so a wave pays one
// memory latency per step instead of one per row.  TH == 0: the tile height
// is g.th at run time and loads run one row ahead of the arithmetic.
template <int NM, bool SC, int TH>
__global__ __launch_bounds__(256) void k_step(const GridDesc *__restrict__ gds, const uint16_t *__restrict__ tile_model,
                                              int32_t n_off, const int32_t *__restrict__ nbase, int32_t probe) {
  const int lane = threadIdx.x & (kLanes - 1);
  // Wave-uniform by construction; readfirstlane lets the compiler keep the
  // descriptor reads scalar.
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int tile = __builtin_amdgcn_readfirstlane(blk * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));

  // A batch's tiles are numbered model by model; the host-built table gives
  // each tile's model in one load (a single model needs none).
  const int m = tile_model ? __builtin_amdgcn_readfirstlane((int)tile_model[tile]) : 0;
  const GridDesc &g = gds[m];
  const int n = n_off + (nbase ? nbase[0] : 0);
  if (n >= g.steps) return;
  // MaxSpeed every 256 steps (run.go:118-124), folded into the step.
  const bool do_max = probe && (n & 255) == 0;
  const int lt = tile - g.tile_base;
  if (lt < 0 || lt >= g.tiles_r * g.tiles_z) return;

  const int th = TH > 0 ? TH : g.th;
  const int tr = lt % g.tiles_r, tzi = lt / g.tiles_r;
  const int i0 = tr * kOwn, j0 = tzi * th;
  const int j1 = min(j0 + th, g.nz);
  const int c = i0 - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && c < g.nr;
  const int rb = n & 1, wb = rb ^ 1;
  const ColConst cc = load_col(g, c);
  GF_GLOBAL float *const *A = g.f[rb];
  GF_GLOBAL float *const *B = g.f[wb];

  // Carried state, named by row relative to the velocity row t being built.
  float trr_p = 0.f, tqq_p = 0.f;          // old trr, tqq at t-1
  float trz_p = ld(A[kTrz], g, j0 - 2, c); // old trz at t-1
  float tzz_p = 0.f;                       // old tzz at t-1
  float tzz_c = ld_tzz(A[kTzz], g, n, j0 - 1, c);  // old tzz at t
  float vrn_p = 0.f, vzn_p = 0.f, vzn_pp = 0.f;   // new vr(t-1), vz(t-1), vz(t-2)
  uint32_t mbits = 0;

  // One row of the march: new velocities of row t from `cur`, then new
  // stresses of row s = t-1 with that row's memory variables `mem_s`.
  auto row = [&](int t, const RowLoad &cur, MemVars<NM> &mem_s) {
    const float trr1 = lane_next(cur.trr), tqq1 = lane_next(cur.tqq), trzm = lane_prev(cur.trz);
    const bool row_own = t >= j0 && t < j1;
    float vrn, vzn;
    velocity(g, rb, t, c, row_own && own, cc, cur.trr, trr1, cur.tqq, tqq1, cur.trz, trzm, trz_p, tzz_c,
             cur.tzz_dn, cur.vr, cur.vz, vrn, vzn);
    if (row_own && own) {
      st_stream(at(B[kVr], g, t, c), vrn);
      st_stream(at(B[kVz], g, t, c), vzn);
      if (do_max) mbits = max(mbits, max(abs_bits(vrn), abs_bits(vzn)));
    }
    const float vrn_m = lane_prev(vrn);
    if (t == 0 && g.nrec > 0) {
      // Receivers sit on the surface row; the tile holding row 0 records them.
      const float sv = surface_velocity(g, c, cc, vrn, vrn_m, vzn);
      const int64_t slot = (int64_t)n - g.rec_base;
      if (slot >= 0 && slot < g.trace_cap) {
        for (int r = 0; r < g.nrec; r++) {
          const int col = g.rec_col[r];
          if (own && col == c) g.traces[(size_t)r * g.trace_cap + slot] = sv;
        }
      }
    }
    const int s = t - 1;
    if (s >= j0) {
      const float vrn_pm = lane_prev(vrn_p);  // new vr(s, c-1)
      const float vzn_p1 = lane_next(vzn_p);  // new vz(s, c+1)
      if (own) {
        float nrr = trr_p, nqq = tqq_p, nzz = tzz_p, nrz = trz_p;
        stress<NM, SC>(g, n, s, c, cc, vrn_p, vrn_pm, vrn, vzn_p, vzn_p1, vzn_pp, nrr, nqq, nzz, nrz, mem_s);
        st_stream(at(B[kTrr], g, s, c), nrr);
        st_stream(at(B[kTqq], g, s, c), nqq);
        st_stream(at(B[kTzz], g, s, c), nzz);
        st_stream(at(B[kTrz], g, s, c), nrz);
        store_mem<NM>(g, s, c, mem_s);
      }
    }
    trr_p = cur.trr;
    tqq_p = cur.tqq;
    trz_p = cur.trz;
    tzz_p = tzz_c;
    tzz_c = cur.tzz_dn;
    vzn_pp = vzn_p;
    vzn_p = vzn;
    vrn_p = vrn;
  };

  if constexpr (TH > 0) {
    RowLoad L[TH + 2];
    MemVars<NM> M[TH + 1];
#pragma unroll
    for (int k = 0; k < TH + 2; k++) {
      const int t = j0 - 1 + k;
      if (t <= j1) L[k] = load_row(A, g, n, t, c);
    }
#pragma unroll
    for (int k = 0; k < TH; k++) {
      const int t = j0 + k;
      if (own && t < j1) load_mem<NM>(g, t, c, M[k + 1]);
    }
#pragma unroll
    for (int k = 0; k < TH + 2; k++) {
      const int t = j0 - 1 + k;
      if (t > j1) break;
      row(t, L[k], M[k > 0 ? k - 1 : 0]);
    }
  } else {
    MemVars<NM> mem_p;  // memory variables of row t-1
    RowLoad cur = load_row(A, g, n, j0 - 1, c);
    for (int t = j0 - 1; t <= j1; t++) {
      // Next row's loads and this row's memory variables go out before any
      // arithmetic, so their latency hides behind it.
      RowLoad nxt;
      if (t < j1) nxt = load_row(A, g, n, t + 1, c);
      MemVars<NM> mem_c;
      if (own && t >= j0 && t < j1) load_mem<NM>(g, t, c, mem_c);
      row(t, cur, mem_p);
      mem_p = mem_c;
      cur = nxt;
    }
  }

  if (do_max) {
    for (int off = 32; off >= 1; off >>= 1) mbits = max(mbits, (uint32_t)__shfl_xor((int)mbits, off, kLanes));
    if (lane == 0) atomicMax(&g.maxbits[n >> 8], mbits);
  }
}

// Sim.SurfaceVelocity for one column of the buffer that holds step `cur`.
__global__ void k_surface_velocity(const GridDesc *__restrict__ gd, int32_t buf, int32_t col, float *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const GridDesc &g = *gd;
  GF_GLOBAL const float *vr = g.f[buf][kVr], *vz = g.f[buf][kVz];
  const ColConst cc = load_col(g, col);
  const float vrm = col > 0 ? vr[col - 1] : 0.f;
  out[0] = surface_velocity(g, col, cc, vr[col], vrm, vz[col]);
  out[1] = vr[col];
}

// Sim.MaxSpeed over both velocity fields of one buffer.
__global__ void k_max_speed(const GridDesc *__restrict__ gd, int32_t buf, uint32_t *slot) {
  const GridDesc &g = *gd;
  GF_GLOBAL const float *vr = g.f[buf][kVr], *vz = g.f[buf][kVz];
  const int64_t cells = (int64_t)g.nz * g.pitch;
  uint32_t m = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cells; i += (int64_t)gridDim.x * blockDim.x)
    m = max(m, max(abs_bits(vr[i]), abs_bits(vz[i])));
  for (int off = 32; off >= 1; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, kLanes));
  if ((threadIdx.x & (kLanes - 1)) == 0) atomicMax(slot, m);
}

// ------------------------------------------------------------------ launchers

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_set_base(int32_t *p, int32_t v, hipStream_t st) {
  hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

hipError_t launch_build_tables(const TableParams &P, const LayerParams *d_layers, float *rowc, float *colc,
                               hipStream_t st) {
  int n = P.nz > P.pitch ? P.nz : P.pitch;
  hipLaunchKernelGGL(k_build_tables, dim3((n + 255) / 256), dim3(256), 0, st, P, d_layers, rowc, colc);
  return hipGetLastError();
}

template <int NM, bool SC, int TH>
static hipError_t launch_step_t(int waves, const GridDesc *d, const uint16_t *tm, int64_t tiles, int32_t n_off,
                                const int32_t *nbase, int32_t probe, hipStream_t st) {
  unsigned blocks = (unsigned)((tiles + waves - 1) / waves);
  hipLaunchKernelGGL((k_step<NM, SC, TH>), dim3(blocks), dim3(64 * waves), 0, st, d, tm, n_off, nbase, probe);
  return hipGetLastError();
}

template <int NM, bool SC>
static hipError_t launch_step_th(int th, int waves, const GridDesc *d, const uint16_t *tm, int64_t tiles,
                                 int32_t n_off, const int32_t *nbase, int32_t probe, hipStream_t st) {
  switch (th) {
    case 1: return launch_step_t<NM, SC, 1>(waves, d, tm, tiles, n_off, nbase, probe, st);
    case 2: return launch_step_t<NM, SC, 2>(waves, d, tm, tiles, n_off, nbase, probe, st);
    case 4: return launch_step_t<NM, SC, 4>(waves, d, tm, tiles, n_off, nbase, probe, st);
    case 8: return launch_step_t<NM, SC, 8>(waves, d, tm, tiles, n_off, nbase, probe, st);
    default: return launch_step_t<NM, SC, 0>(waves, d, tm, tiles, n_off, nbase, probe, st);
  }
}

// `th` is the tile height every model of the launch was planned with; 1, 2,
// 4 and 8 select the fully unrolled kernels, anything else the row march.
// `waves` tiles per workgroup (1, 2 or 4).  `tile_model` maps each tile to
// its model (null: every tile is model 0); it is padded to whole groups of 4.
hipError_t launch_step(int nmech, bool scalar, int th, int waves, const GridDesc *d, const uint16_t *tile_model,
                       int64_t tiles, int32_t n_off, const int32_t *nbase, int32_t probe, hipStream_t st) {
  if (tiles <= 0) return hipSuccess;
  if (waves != 1 && waves != 2 && waves != 4) return hipErrorInvalidValue;
  if (nmech == 0) return launch_step_th<0, true>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
  if (scalar) return launch_step_th<1, true>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
  switch (nmech) {
    case 1: return launch_step_th<1, false>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
    case 2: return launch_step_th<2, false>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
    case 3: return launch_step_th<3, false>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
    case 4: return launch_step_th<4, false>(th, waves, d, tile_model, tiles, n_off, nbase, probe, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_surface_velocity(const GridDesc *d, int32_t buf, int32_t col, float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_surface_velocity, dim3(1), dim3(64), 0, st, d, buf, col, out);
  return hipGetLastError();
}

hipError_t launch_max_speed(const GridDesc *d, int32_t buf, uint32_t *slot, int64_t cells, hipStream_t st) {
  int blocks = (int)std::min<int64_t>((cells + 255) / 256, 2048);
  hipLaunchKernelGGL(k_max_speed, dim3(blocks), dim3(256), 0, st, d, buf, slot);
  return hipGetLastError();
}

}  // namespace gfk
