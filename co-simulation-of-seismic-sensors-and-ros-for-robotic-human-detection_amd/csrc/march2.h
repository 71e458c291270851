// march2.h — the per-lane phases of the two-step (temporally blocked) kernel.
//
// k_step2 advances every tile by two time steps per launch, so each state
// array crosses HBM once per two steps instead of once per step.  A wavefront
// owns kOwn2 = 60 columns (lanes 2..61; lane L holds column i0 - 2 + L) and
// `th` rows [j0, j1), and runs two one-step row marches in lockstep:
//
//   stage 1 (step n) reads the old state of rows j0-4 .. j1+3 from buffer rb
//     and, at march row t, forms the new velocities of row t and the new
//     stresses (with C-PML auxiliaries and memory variables) of row t-1, for
//     every lane, halo lanes included;
//   stage 2 (step n+1) takes stage 1's rows as its input two rows behind
//     (row u = t - 2 once stage 1 has produced the stresses of row u + 1) and
//     stores the owned rows and lanes of step n+2's state to buffer rb ^ 1.
//
// The dependence cone of one step is one column and one row each way (per
// field: vr(c) reads the stresses of c and c+1, vz(c) those of c-1 and c, a
// stress at c the velocities vr(c-1), vr(c), vz(c), vz(c+1); likewise rows),
// so two steps need two halo columns each side of the 60 owned ones and the
// old rows j0-4 .. j1+3.  Values stage 1 forms at lanes or rows whose inputs
// lie outside the wavefront are wrong but feed nothing stage 2 stores.  Both
// stages use the one-step kernel's per-cell arithmetic (stencil.h), so the
// results are those of two one-step launches.
//
// Nothing a launch reads is written by it (the halo of one tile is the owned
// region of another): the wavefields and velocity psi are double-buffered
// for every launch, and a two-step launch also reads the memory variables
// and stress psi from one buffer and writes them to the other (a one-step
// launch updates those in place).
//
// Plain C++ so the CPU emulation runs the same phases lane by lane.
#pragma once

#include "stencil.h"

namespace gfk {

// The carried values of one stage's row march, named by row relative to the
// velocity row t being formed: the stage's input stresses of rows t-1
// (trr_p, tqq_p, trz_p, tzz_p) and t (tzz_c), and its new velocities of rows
// t-1 (vrn_p, vzn_p) and t-2 (vzn_pp).
struct Stage {
  float trr_p, tqq_p, trz_p, tzz_p, tzz_c;
  float vrn_p, vzn_p, vzn_pp;
};

GF_HD Stage stage_start(float trz_above, float tzz_row) {
  Stage S = {0.f, 0.f, trz_above, 0.f, tzz_row, 0.f, 0.f, 0.f};
  return S;
}

// New velocities of row t from the stage's input row `in` (its stresses and
// velocities at t, tzz at t+1) and the neighbour lanes' trr(c+1), tqq(c+1),
// trz(c-1).
template <bool IN = false>
GF_HD void stage_velocity(const GridDesc &g, int t, int c, const ColConst &cc, const Stage &S, const RowLoad &in,
                          float trr1, float tqq1, float trzm, VPsi &vp, float &vrn, float &vzn) {
  velocity_core<IN>(g, t, c, cc, in.trr, trr1, in.tqq, tqq1, in.trz, trzm, S.trz_p, S.tzz_c, in.tzz_dn, in.vr, in.vz,
                    vp, vrn, vzn);
}

// New stresses of row s = t - 1 of time step n (out: trr, tqq, tzz, trz),
// given the new velocities of row t (vrn) and the neighbour lanes' vr(s, c-1)
// and vz(s, c+1).  Cells outside the grid keep their input values.
template <int NM, bool SC, bool IN = false>
GF_HD void stage_stress(const GridDesc &g, int n, int s, int c, const ColConst &cc, const Stage &S, float vrn,
                        float vrn_pm, float vzn_p1, float out[4], SPsi &sp, MemVars<NM> &mem) {
  out[0] = S.trr_p;
  out[1] = S.tqq_p;
  out[2] = S.tzz_p;
  out[3] = S.trz_p;
  if (IN || grid_cell(g, s, c))
    stress_core<NM, SC, IN>(g, n, s, c, cc, S.vrn_p, vrn_pm, vrn, S.vzn_p, vzn_p1, S.vzn_pp, out[0], out[1], out[2],
                            out[3], sp, mem);
}

// Advance the carried values past row t.
GF_HD void stage_carry(Stage &S, const RowLoad &in, float vrn, float vzn) {
  S.trr_p = in.trr;
  S.tqq_p = in.tqq;
  S.trz_p = in.trz;
  S.tzz_p = S.tzz_c;
  S.tzz_c = in.tzz_dn;
  S.vzn_pp = S.vzn_p;
  S.vzn_p = vzn;
  S.vrn_p = vrn;
}

// tzz as step n reads it: the prescribed surface traction at the axis cell
// (fdtd.go:152), the stored value elsewhere.
GF_HD float tzz_for_step(const GridDesc &g, int n, int t, int c, float stored) {
  if (t == 0 && c == 0) return n < g.nsrc ? g.src_tzz[n] : 0.f;
  return stored;
}

// Tile lt of a model's two-step tiling: owned columns from i0 (kOwn2 of them,
// fewer at the grid's edge), rows [j0, j1).
struct Tile2 {
  int i0, j0, j1;
};

GF_HD Tile2 step2_tile_at(const GridDesc &g, int lt) {
  const int tr = lt % g.tiles_r, tzi = lt / g.tiles_r;
  Tile2 t;
  t.i0 = tr * kOwn2;
  t.j0 = tzi * g.th;
  t.j1 = t.j0 + g.th < g.nz ? t.j0 + g.th : g.nz;
  return t;
}

// Whether a two-step tile is clear of the axis, the surface and the C-PML
// strips: stage 1 forms columns i0-2 .. i0+kOwn2+1 and velocity rows j0-3 ..
// j1+2.  The host sorts tiles by this test, with the kernel's own function.
GF_HD bool step2_interior(const GridDesc &g, int i0, int j0, int j1) {
  return interior(g, i0 - 2, i0 + kOwn2 + 1, j0 - 3, j1 + 2);
}

// Stage 1's outputs as stage 2 reads them: the velocity half of a row
// (formed at march row t) and the stress half (formed at march row t + 1).
struct VelOut {
  float vr, vz;  // new velocities of the row
  VPsi vp;       // its advanced velocity auxiliaries
};

template <int NM>
struct StrOut {
  float st[4];      // new trr, tqq, tzz, trz of the row
  SPsi sp;          // advanced stress auxiliaries
  MemVars<NM> mem;  // advanced memory variables
};

}  // namespace gfk
