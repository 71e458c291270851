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

#include <type_traits>

#include "coef.h"
#include "kernels.h"
#include "march2.h"
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

// A captured graph's base step and the buffers that hold its first step's
// state (wavefields, memory variables).
__global__ void k_set_base(int32_t *p, int32_t v, int32_t buf, int32_t mbuf) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    p[0] = v;
    p[1] = buf;
    p[2] = mbuf;
  }
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
//
// The step is n_off (+ nbase[0] when replayed from a graph); the wavefields
// are read from buffer rb_off (^ nbase[1]) and written to the other, the
// memory variables updated in place in buffer mb_off (^ nbase[2]).
template <int NM, bool SC, int TH, bool IN>
__device__ __forceinline__ void step_tile(const GridDesc &g, int n, int rb, int mb, bool do_max, int th, int i0,
                                          int j0, int j1, int lane) {
  const int wb = rb ^ 1;
  const int c = i0 - 1 + lane;
  const bool own = lane >= 1 && lane <= kOwn && (IN || c < g.nr);
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
    velocity<IN>(g, rb, t, c, row_own && own, cc, cur.trr, trr1, cur.tqq, tqq1, cur.trz, trzm, trz_p, tzz_c,
                 cur.tzz_dn, cur.vr, cur.vz, vrn, vzn);
    if (row_own && own) {
      st_stream(at(B[kVr], g, t, c), vrn);
      st_stream(at(B[kVz], g, t, c), vzn);
      if (do_max) mbits = max(mbits, max(abs_bits(vrn), abs_bits(vzn)));
    }
    const float vrn_m = lane_prev(vrn);
    if (!IN && t == 0 && g.nrec > 0) {
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
        stress<NM, SC, IN>(g, mb, n, s, c, cc, vrn_p, vrn_pm, vrn, vzn_p, vzn_p1, vzn_pp, nrr, nqq, nzz, nrz, mem_s);
        st_stream(at(B[kTrr], g, s, c), nrr);
        st_stream(at(B[kTqq], g, s, c), nqq);
        st_stream(at(B[kTzz], g, s, c), nzz);
        st_stream(at(B[kTrz], g, s, c), nrz);
        store_mem<NM>(g, mb, s, c, mem_s);
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
      if (own && t < j1) load_mem<NM>(g, mb, t, c, M[k + 1]);
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
      if (own && t >= j0 && t < j1) load_mem<NM>(g, mb, t, c, mem_c);
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

template <int NM, bool SC, int TH>
__global__ __launch_bounds__(256) void k_step(const GridDesc *__restrict__ gds, const uint16_t *__restrict__ tile_model,
                                              int32_t n_off, int32_t rb_off, int32_t mb_off,
                                              const int32_t *__restrict__ nbase, int32_t probe) {
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
  const int rb = (rb_off ^ (nbase ? nbase[1] : 0)) & 1;
  const int mb = (mb_off ^ (nbase ? nbase[2] : 0)) & 1;
  // Velocities are formed at columns i0-1 .. i0+62 and rows j0-1 .. j1.
  if (interior(g, i0 - 1, i0 + kOwn, j0 - 1, j1))
    step_tile<NM, SC, TH, true>(g, n, rb, mb, do_max, th, i0, j0, j1, lane);
  else
    step_tile<NM, SC, TH, false>(g, n, rb, mb, do_max, th, i0, j0, j1, lane);
}

// Two time steps per launch (march2.h): the tile is kOwn2 columns by g.th
// rows; stage 1 (step n) marches rows j0-3 .. j1+2 from buffers rb (fields)
// and mb (memory variables, stress psi), stage 2 (step n+1) follows two rows
// behind and stores rows [j0, j1) to buffers rb ^ 1 and mb ^ 1.  With
// count == 1, or when step n is a model's last, stage 1's owned rows are
// stored instead (one step).  Loads run one march row ahead;
// every per-row array is indexed by the march row modulo 3 at compile time,
// so the rows in flight stay in registers without copies.
template <int NM, bool SC, bool IN>
__device__ __forceinline__ void step2_tile(const GridDesc &g, int n, bool two, int rb, int mb, bool max1, bool max2,
                                           int i0, int j0, int j1, int lane) {
  const int wb = rb ^ 1, mw = mb ^ 1;
  const int c = i0 - 2 + lane;
  const bool own = lane >= 2 && lane < 2 + kOwn2 && (IN || c < g.nr);
  const ColConst cc = load_col(g, c);
  GF_GLOBAL float *const *A = g.f[rb];
  GF_GLOBAL float *const *B = g.f[wb];
  uint32_t mbits = 0;
  // Receivers sit on row 0; the tile that owns it records them.
  const bool rec = !IN && j0 == 0 && g.nrec > 0;
  auto record = [&](int64_t step, float sv) {
    const int64_t slot = step - g.rec_base;
    if (slot >= 0 && slot < g.trace_cap) {
      for (int r = 0; r < g.nrec; r++) {
        const int col = g.rec_col[r];
        if (own && col == c) g.traces[(size_t)r * g.trace_cap + slot] = sv;
      }
    }
  };

  const int t0 = j0 - 3, t_end = j1 + 2;
  Stage S1 = stage_start(ld(A[kTrz], g, j0 - 4, c), ld_tzz(A[kTzz], g, n, j0 - 3, c));
  Stage S2 = stage_start(0.f, 0.f);
  // stage 1's inputs (loaded a row ahead) and outputs, by march row mod 3
  RowLoad ROW[3];
  VPsi IVP[3];
  SPsi ISP[3];
  MemVars<NM> IMEM[3];
  VelOut QV[3];       // velocity half of row t
  StrOut<NM> QS[3];   // stress half of row t - 1

  auto fetch = [&](auto P, int t) {
    constexpr int p = decltype(P)::value;
    ROW[p] = load_row(A, g, n, t, c);
    IVP[p] = VPsi{0.f, 0.f, 0.f, 0.f};
    if (!IN && vel_cell(g, t, c)) IVP[p] = load_vpsi(g, rb, t, c);
    const int s = t - 1;
    ISP[p] = SPsi{0.f, 0.f, 0.f, 0.f};
    if (IN) {
      if (s >= j0 - 2) load_mem<NM>(g, mb, s, c, IMEM[p]);
    } else {
      IMEM[p] = MemVars<NM>{};
      if (s >= j0 - 2 && grid_cell(g, s, c)) {
        ISP[p] = load_spsi(g, mb, s, c);
        load_mem<NM>(g, mb, s, c, IMEM[p]);
      }
    }
  };

  auto body = [&](auto P, int t) {
    constexpr int p = decltype(P)::value;
    constexpr int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
    if (t + 1 <= t_end) fetch(std::integral_constant<int, p1>(), t + 1);
    const RowLoad &in = ROW[p];

    // stage 1: velocities of row t
    VelOut &v1 = QV[p];
    v1.vp = IVP[p];
    stage_velocity<IN>(g, t, c, cc, S1, in, lane_next(in.trr), lane_next(in.tqq), lane_prev(in.trz), v1.vp, v1.vr,
                       v1.vz);
    const bool vown = own && t >= j0 && t < j1;
    if (vown) {
      if (!two) {
        st_stream(at(B[kVr], g, t, c), v1.vr);
        st_stream(at(B[kVz], g, t, c), v1.vz);
        if (!IN && vel_cell(g, t, c)) store_vpsi(g, wb, t, c, v1.vp);
      }
      if (max1) mbits = max(mbits, max(abs_bits(v1.vr), abs_bits(v1.vz)));
    }
    if (rec && t == 0) record(n, surface_velocity(g, c, cc, v1.vr, lane_prev(v1.vr), v1.vz));

    // stage 1: stresses of row t - 1
    const int s = t - 1;
    if (s >= j0 - 2) {
      const float vrn_pm = lane_prev(S1.vrn_p), vzn_p1 = lane_next(S1.vzn_p);
      StrOut<NM> &o1 = QS[p];
      o1.sp = ISP[p];
      o1.mem = IMEM[p];
      stage_stress<NM, SC, IN>(g, n, s, c, cc, S1, v1.vr, vrn_pm, vzn_p1, o1.st, o1.sp, o1.mem);
      if (!two && own && s >= j0 && s < j1) {
        st_stream(at(B[kTrr], g, s, c), o1.st[0]);
        st_stream(at(B[kTqq], g, s, c), o1.st[1]);
        st_stream(at(B[kTzz], g, s, c), o1.st[2]);
        st_stream(at(B[kTrz], g, s, c), o1.st[3]);
        if (!IN) store_spsi(g, mw, s, c, o1.sp);
        store_mem<NM>(g, mw, s, c, o1.mem);
      }
    }
    stage_carry(S1, in, v1.vr, v1.vz);
    if (!two) return;

    // stage 2 on row u = t - 2: its velocity half was formed at march row
    // t - 2 (slot p1), its stresses at t - 1 (slot p2), the stresses of row
    // u + 1 just now (slot p) and those of row u - 1 at t - 2 (slot p1).
    const int u = t - 2;
    if (u < j0 - 1) return;
    const VelOut &vu = QV[p1];
    const StrOut<NM> &su = QS[p2];
    if (u == j0 - 1) S2 = stage_start(QS[p1].st[3], IN ? su.st[2] : tzz_for_step(g, n + 1, u, c, su.st[2]));
    RowLoad in2;
    in2.trr = su.st[0];
    in2.tqq = su.st[1];
    in2.trz = su.st[3];
    in2.tzz_dn = IN ? QS[p].st[2] : tzz_for_step(g, n + 1, u + 1, c, QS[p].st[2]);
    in2.vr = vu.vr;
    in2.vz = vu.vz;
    VPsi vp2 = vu.vp;
    float vr2, vz2;
    stage_velocity<IN>(g, u, c, cc, S2, in2, lane_next(in2.trr), lane_next(in2.tqq), lane_prev(in2.trz), vp2, vr2,
                       vz2);
    if (own && u >= j0 && u < j1) {
      st_stream(at(B[kVr], g, u, c), vr2);
      st_stream(at(B[kVz], g, u, c), vz2);
      if (!IN && vel_cell(g, u, c)) store_vpsi(g, wb, u, c, vp2);
      if (max2) mbits = max(mbits, max(abs_bits(vr2), abs_bits(vz2)));
    }
    if (rec && u == 0) record((int64_t)n + 1, surface_velocity(g, c, cc, vr2, lane_prev(vr2), vz2));
    const int s2 = u - 1;
    if (s2 >= j0) {
      const float vrn_pm = lane_prev(S2.vrn_p), vzn_p1 = lane_next(S2.vzn_p);
      SPsi sp2 = QS[p1].sp;
      MemVars<NM> mem2 = QS[p1].mem;
      float o2[4];
      stage_stress<NM, SC, IN>(g, n + 1, s2, c, cc, S2, vr2, vrn_pm, vzn_p1, o2, sp2, mem2);
      if (own) {
        st_stream(at(B[kTrr], g, s2, c), o2[0]);
        st_stream(at(B[kTqq], g, s2, c), o2[1]);
        st_stream(at(B[kTzz], g, s2, c), o2[2]);
        st_stream(at(B[kTrz], g, s2, c), o2[3]);
        if (!IN) store_spsi(g, mw, s2, c, sp2);
        store_mem<NM>(g, mw, s2, c, mem2);
      }
    }
    stage_carry(S2, in2, vr2, vz2);
  };

  fetch(std::integral_constant<int, 0>(), t0);
  for (int t = t0; t <= t_end; t += 3) {
    body(std::integral_constant<int, 0>(), t);
    if (t + 1 > t_end) break;
    body(std::integral_constant<int, 1>(), t + 1);
    if (t + 2 > t_end) break;
    body(std::integral_constant<int, 2>(), t + 2);
  }

  if (max1 || max2) {
    for (int off = 32; off >= 1; off >>= 1) mbits = max(mbits, (uint32_t)__shfl_xor((int)mbits, off, kLanes));
    if (lane == 0) atomicMax(&g.maxbits[(max1 ? n : n + 1) >> 8], mbits);
  }
}

// PART 0: the launch's tiles are 0 .. (the grid's tiles), each taking the
// interior or the general path as it finds itself.  PART 1 and 2: the tiles
// are the n_list entries of tile_list, all interior (1) or all not (2), so
// that the interior kernel carries none of the general path's registers (155
// against 199 VGPRs for three mechanisms: three wavefronts per SIMD instead
// of two).
template <int NM, bool SC, int PART>
__global__ __launch_bounds__(256) void k_step2(const GridDesc *__restrict__ gds, const uint16_t *__restrict__ tile_model,
                                               int32_t n_off, int32_t rb_off, int32_t mb_off,
                                               const int32_t *__restrict__ nbase, int32_t probe, int32_t count,
                                               const uint32_t *__restrict__ tile_list, int32_t n_list) {
  const int lane = threadIdx.x & (kLanes - 1);
  const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
  int tile = __builtin_amdgcn_readfirstlane(blk * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
  if constexpr (PART != 0) {
    if (tile >= n_list) return;
    tile = __builtin_amdgcn_readfirstlane((int)tile_list[tile]);
  }
  const int m = tile_model ? __builtin_amdgcn_readfirstlane((int)tile_model[tile]) : 0;
  const GridDesc &g = gds[m];
  const int n = n_off + (nbase ? nbase[0] : 0);
  if (n >= g.steps) return;
  const int lt = tile - g.tile_base;
  if (lt < 0 || lt >= g.tiles_r * g.tiles_z) return;
  // A model whose run ends after step n takes one step.
  const bool two = count >= 2 && n + 1 < g.steps;
  const int rb = (rb_off ^ (nbase ? nbase[1] : 0)) & 1;
  const int mb = (mb_off ^ (nbase ? nbase[2] : 0)) & 1;

  const Tile2 tt = step2_tile_at(g, lt);
  const int i0 = tt.i0, j0 = tt.j0, j1 = tt.j1;
  // MaxSpeed every 256 steps; at most one of the two steps is a multiple.
  const bool max1 = probe && (n & 255) == 0;
  const bool max2 = two && probe && ((n + 1) & 255) == 0;
  // Stage 1 forms columns i0-2 .. i0+61, velocity rows j0-3 .. j1+2.
  if constexpr (PART == 1) {
    step2_tile<NM, SC, true>(g, n, two, rb, mb, max1, max2, i0, j0, j1, lane);
  } else if constexpr (PART == 2) {
    step2_tile<NM, SC, false>(g, n, two, rb, mb, max1, max2, i0, j0, j1, lane);
  } else {
    if (step2_interior(g, i0, j0, j1))
      step2_tile<NM, SC, true>(g, n, two, rb, mb, max1, max2, i0, j0, j1, lane);
    else
      step2_tile<NM, SC, false>(g, n, two, rb, mb, max1, max2, i0, j0, j1, lane);
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

hipError_t launch_set_base(int32_t *p, int32_t v, int32_t buf, int32_t mbuf, hipStream_t st) {
  hipLaunchKernelGGL(k_set_base, dim3(1), dim3(64), 0, st, p, v, buf, mbuf);
  return hipGetLastError();
}

hipError_t launch_build_tables(const TableParams &P, const LayerParams *d_layers, float *rowc, float *colc,
                               hipStream_t st) {
  int n = P.nz > P.pitch ? P.nz : P.pitch;
  hipLaunchKernelGGL(k_build_tables, dim3((n + 255) / 256), dim3(256), 0, st, P, d_layers, rowc, colc);
  return hipGetLastError();
}

template <int NM, bool SC, int TH>
static hipError_t launch_step_t(const StepLaunch &L, hipStream_t st) {
  unsigned blocks = (unsigned)((L.tiles + L.waves - 1) / L.waves);
  hipLaunchKernelGGL((k_step<NM, SC, TH>), dim3(blocks), dim3(64 * L.waves), 0, st, L.d, L.tile_model, L.n_off,
                     L.rb_off, L.mb_off, L.nbase, L.probe);
  return hipGetLastError();
}

template <int NM, bool SC>
static hipError_t launch_step_v(const StepLaunch &L, hipStream_t st) {
  if (L.span == 2) {
    if (L.list_in || L.list_bd) {
      // the interior tiles and the others as two launches; both read buffer
      // rb and write the other, so their order does not matter
      if (L.n_in > 0) {
        unsigned blocks = (unsigned)((L.n_in + L.waves - 1) / L.waves);
        hipLaunchKernelGGL((k_step2<NM, SC, 1>), dim3(blocks), dim3(64 * L.waves), 0, st, L.d, L.tile_model, L.n_off,
                           L.rb_off, L.mb_off, L.nbase, L.probe, L.count, L.list_in, (int32_t)L.n_in);
      }
      if (L.n_bd > 0) {
        unsigned blocks = (unsigned)((L.n_bd + L.waves - 1) / L.waves);
        hipLaunchKernelGGL((k_step2<NM, SC, 2>), dim3(blocks), dim3(64 * L.waves), 0, st, L.d, L.tile_model, L.n_off,
                           L.rb_off, L.mb_off, L.nbase, L.probe, L.count, L.list_bd, (int32_t)L.n_bd);
      }
      return hipGetLastError();
    }
    unsigned blocks = (unsigned)((L.tiles + L.waves - 1) / L.waves);
    hipLaunchKernelGGL((k_step2<NM, SC, 0>), dim3(blocks), dim3(64 * L.waves), 0, st, L.d, L.tile_model, L.n_off,
                       L.rb_off, L.mb_off, L.nbase, L.probe, L.count, (const uint32_t *)nullptr, 0);
    return hipGetLastError();
  }
  switch (L.th) {
    case 1: return launch_step_t<NM, SC, 1>(L, st);
    case 2: return launch_step_t<NM, SC, 2>(L, st);
    case 4: return launch_step_t<NM, SC, 4>(L, st);
    case 8: return launch_step_t<NM, SC, 8>(L, st);
    default: return launch_step_t<NM, SC, 0>(L, st);
  }
}

hipError_t launch_step(const StepLaunch &L, hipStream_t st) {
  if (L.tiles <= 0) return hipSuccess;
  if (L.waves != 1 && L.waves != 2 && L.waves != 4) return hipErrorInvalidValue;
  if (L.span != 1 && L.span != 2) return hipErrorInvalidValue;
  if (L.count < 1 || L.count > L.span) return hipErrorInvalidValue;
  if (L.nmech == 0) return launch_step_v<0, true>(L, st);
  if (L.scalar) return launch_step_v<1, true>(L, st);
  switch (L.nmech) {
    case 1: return launch_step_v<1, false>(L, st);
    case 2: return launch_step_v<2, false>(L, st);
    case 3: return launch_step_v<3, false>(L, st);
    case 4: return launch_step_v<4, false>(L, st);
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
