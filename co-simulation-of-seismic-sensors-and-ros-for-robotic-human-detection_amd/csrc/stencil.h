// stencil.h — the per-lane arithmetic of one FDTD time step on gfx950.
//
// One time step of the reference's Sim.Step (internal/fdtd/fdtd.go:149-307)
// is one kernel launch.  Each wavefront owns a tile of kOwn = 62 grid columns
// by `th` rows and marches down it row by row.  Lane L holds column
// c = i0 - 1 + L, so lanes 0 and 63 are halo columns: they recompute the new
// velocities their neighbours' stresses need (vr one column left, vz one
// column right), and the march starts one row above and ends one row below
// the tile for the same reason.  Recomputing a one-cell velocity halo lets
// velocities and stresses fuse into a single pass: every state array is read
// once and written once per step, which is the whole roofline budget of a
// memory-bound stencil.
//
// Because halo lanes and rows read neighbours' old values while neighbours
// write their new ones, everything a halo recompute reads is double-buffered:
// the six wavefield arrays and the four C-PML psi arrays of the velocity
// derivatives are read from buffer rb and written to buffer rb ^ 1 by every
// launch.  The memory variables and the psi arrays of the stress derivatives
// are touched only by the owning lane in a one-step launch, which updates
// them in place in buffer mb; the two-step kernel of march2.h recomputes a
// neighbour's stresses in its halo, so it reads them from buffer mb and
// writes buffer mb ^ 1.  (Doubling them for every launch would grow a
// one-step launch's working set by half and push a 2048 x 1024 GSLS grid,
// 305 MB of state, further out of the 256 MB last-level cache.)  The caller
// tracks both buffer indices.
//
// The functions below are the per-lane phases of that march; the kernel
// (fdtd_kernels.hip) strings them together with cross-lane shuffles.  They
// are plain C++ so that the same arithmetic can be run lane by lane on a CPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GF_HD __host__ __device__ __forceinline__
#else
#define GF_HD inline
#endif

// Device code sees every array the descriptor points to as global memory, so
// the compiler emits global_load/global_store rather than flat accesses (a
// flat access also counts against the scalar-load counter, which serialises
// the kernel's vector loads behind every scalar wait).  Host code sees plain
// pointers; the layout is the same.
#if defined(__HIP_DEVICE_COMPILE__)
#define GF_GLOBAL __attribute__((address_space(1)))
#else
#define GF_GLOBAL
#endif

namespace gfk {

constexpr int kLanes = 64;   // CDNA wavefront
constexpr int kOwn = 62;     // owned columns per tile (two halo lanes)
constexpr int kOwn2 = 60;    // owned columns per tile of the two-step kernel
constexpr int kMaxMech = 8;  // GSLS mechanisms
constexpr int kGuardRows = 4;   // zero rows above and below every field
constexpr int kGuardCols = 64;  // zero floats before row -kGuardRows

// Per-row coefficient table, float32, one row of kRowW floats per grid row:
// [t][kRow*], then the GSLS drives at kRowN + q (P) and kRowN + nmech + q
// (S).  A row's coefficients are contiguous, so a wavefront reads them with
// a few wide scalar loads from one address per row.
constexpr int kRowW = 32;  // >= kRowN + 2 * kMaxMech
enum RowCoef {
  kRowCvr = 0,   // dt / rho            (vr update, fdtd.go:186)
  kRowCvz,       // dt / rhoZ           (vz update, fdtd.go:214)
  kRowLam,       // lamU
  kRowMu,        // muU
  kRowL2m,       // lamU + 2 muU
  kRowIl2m,      // 1 / (lamU + 2 muU)  (free-surface dvz, fdtd.go:256)
  kRowMuz,       // harmonic muU at the half row (fdtd.go:288)
  kRowAzi,       // C-PML z integer profile a, b   (pml.go:29-63)
  kRowBzi,
  kRowAzh,       // C-PML z half profile a, b
  kRowBzh,
  kRowN
};

// Per-column coefficient table, float32, [kCol*][pitch].
enum ColCoef {
  kColIri = 0,  // 1 / (i h)        (vrr and radial terms; 0 on the axis)
  kColIrh,      // 1 / ((i+1/2) h)  (hoop term, fdtd.go:185)
  kColAri,      // C-PML r integer profile a, b
  kColBri,
  kColArh,      // C-PML r half profile a, b
  kColBrh,
  kColN
};

// Everything one model's step kernel needs.  Lives in device memory; one per
// model of a batch, ordered by tile_base.
struct GridDesc {
  int32_t nr, nz, pitch, th;
  int32_t tiles_r, tiles_z, tile_base, steps;
  int32_t first_rint, first_rhalf, first_zint, first_zhalf;
  int32_t r0, rs_pitch, z0, nrec;
  int32_t nsrc, nmech, scalar, lossy;
  float invh, dt, half_h, svc;  // svc = -lamU0 / (lamU0 + 2 muU0)
  float mem_decay, mem_drive;   // reference single-SLS constants
  float pad0, pad1;
  float decay[kMaxMech];        // GSLS per-mechanism decay
  int64_t trace_cap, rec_base;  // traces[r * trace_cap + (n - rec_base)]
  int64_t field_stride;         // floats per padded field allocation
  int64_t mem_stride;           // floats per memory-variable plane (nz*pitch)
  GF_GLOBAL float *f[2][6];        // [buffer][vr, vz, trr, tqq, tzz, trz] at (0,0)
  GF_GLOBAL float *vpsi_r[2][2];   // [buffer][psiTrrR (r half), psiTrzR (r int)]
  GF_GLOBAL float *vpsi_z[2][2];   // [buffer][psiTrzZ (z int), psiTzzZ (z half)]
  GF_GLOBAL float *spsi_r[2][2];   // [buffer][psiVrR (r int), psiVzR (r half)]
  GF_GLOBAL float *spsi_z[2][2];   // [buffer][psiVzZ (z int), psiVrZ (z half)]
  GF_GLOBAL float *mem[2];         // [buffer] -> [4 components][nmech][nz][pitch]
  GF_GLOBAL const float *rowc;     // [nz][kRowW]
  GF_GLOBAL const float *colc;     // [kColN][pitch]
  GF_GLOBAL const float *src_tzz;  // -F[n] / A          [nsrc]
  GF_GLOBAL const float *src_dtdt; // (F[n+1]-F[n]) / dt  [nsrc]
  GF_GLOBAL const int32_t *rec_col;  // [nrec]
  GF_GLOBAL float *traces;         // [nrec][trace_cap]
  GF_GLOBAL uint32_t *maxbits;     // running-max slots, one per 256 steps
};

enum Field { kVr = 0, kVz, kTrr, kTqq, kTzz, kTrz };

// Stores of state that is not read again within the step.  With
// GF_NT_STORES the device build marks them non-temporal, so they do not
// displace the halo rows neighbouring tiles are about to re-read from L2.
template <typename T>
GF_HD void st_stream(GF_GLOBAL T *p, T v) {
#if defined(__HIP_DEVICE_COMPILE__) && defined(GF_NT_STORES)
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Column constants a lane keeps in registers for the whole march.
struct ColConst {
  float iri, irh, ari, bri, arh, brh;
};

GF_HD ColConst load_col(const GridDesc &g, int c) {
  ColConst cc = {0.f, 0.f, 0.f, 1.f, 0.f, 1.f};
  if (c >= 0 && c < g.pitch) {
    GF_GLOBAL const float *t = g.colc;
    size_t p = (size_t)g.pitch;
    cc.iri = t[kColIri * p + c];
    cc.irh = t[kColIrh * p + c];
    cc.ari = t[kColAri * p + c];
    cc.bri = t[kColBri * p + c];
    cc.arh = t[kColArh * p + c];
    cc.brh = t[kColBrh * p + c];
  }
  return cc;
}

// Address of element (row t, column c) of a plane of pitch g.pitch.  On the
// device the plane's origin (row -kGuardRows, column -2) is wave-uniform and
// does not change along a march, so it stays in scalar registers, and row
// and column go into one 32-bit byte offset (t >= -kGuardRows, c >= -2: the
// offset is non-negative), the same for every plane: the access takes the
// scalar-base + vector-offset form and a march row costs one offset update,
// not 64-bit scalar address arithmetic per plane.
GF_HD GF_GLOBAL float *at(GF_GLOBAL float *base, const GridDesc &g, int t, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
  GF_GLOBAL char *org = (GF_GLOBAL char *)(base - ((ptrdiff_t)kGuardRows * g.pitch + 2));
  return (GF_GLOBAL float *)(org + (uint32_t)((t + kGuardRows) * g.pitch + c + 2) * 4u);
#else
  return base + ((ptrdiff_t)t * g.pitch + c);
#endif
}

// Field element at (row t, column c); rows -4..nz+3 and columns -2..pitch-1
// are inside the padded allocation and read as zero outside the grid
// (columns -2 and -1 of a row are the last two of the row above, which lie
// past the grid's last column, nr <= pitch - 2).
GF_HD float ld(GF_GLOBAL const float *base, const GridDesc &g, int t, int c) {
  return *at(const_cast<GF_GLOBAL float *>(base), g, t, c);
}

// tzz with the prescribed surface traction of step n at the axis cell: the
// reference writes tzz[0] = -F[n]/A before the velocity update (fdtd.go:152).
GF_HD float ld_tzz(GF_GLOBAL const float *base, const GridDesc &g, int n, int t, int c) {
  if (t == 0 && c == 0) return n < g.nsrc ? g.src_tzz[n] : 0.f;
  return ld(base, g, t, c);
}

// The six old field values of one cell that the velocity update of row t
// reads (tzz one row down, for dtzz).
struct RowLoad {
  float trr, tqq, trz, tzz_dn, vr, vz;
};

GF_HD RowLoad load_row(GF_GLOBAL float *const *A, const GridDesc &g, int n, int t, int c) {
  RowLoad L;
  L.trr = ld(A[kTrr], g, t, c);
  L.tqq = ld(A[kTqq], g, t, c);
  L.trz = ld(A[kTrz], g, t, c);
  L.tzz_dn = ld_tzz(A[kTzz], g, n, t + 1, c);
  L.vr = ld(A[kVr], g, t, c);
  L.vz = ld(A[kVz], g, t, c);
  return L;
}

// The C-PML auxiliaries one cell's velocity update reads and advances
// (zero, and untouched, outside their strips).
struct VPsi {
  float trr_r;  // psiTrrR (r half)
  float trz_r;  // psiTrzR (r int)
  float trz_z;  // psiTrzZ (z int)
  float tzz_z;  // psiTzzZ (z half)
};

// The C-PML auxiliaries of one cell's stress update.
struct SPsi {
  float vr_r;  // psiVrR (r int)
  float vz_r;  // psiVzR (r half)
  float vz_z;  // psiVzZ (z int)
  float vr_z;  // psiVrZ (z half)
};

// Cells whose velocities the reference updates (fdtd.go:158-160), and cells
// of the grid (whose stresses it updates, fdtd.go:219-221).
GF_HD bool vel_cell(const GridDesc &g, int t, int c) { return t >= 0 && t <= g.nz - 2 && c >= 0 && c <= g.nr - 2; }
GF_HD bool grid_cell(const GridDesc &g, int s, int c) { return s >= 0 && s < g.nz && c >= 0 && c < g.nr; }

// Whether a tile's every computed cell is an updated cell of the interior:
// columns [clo, chi] and rows [tlo, thi] off the axis and the surface row,
// inside the updated range and clear of every C-PML strip.  Such a tile runs
// the case-free form of the per-cell arithmetic (IN = true), which computes
// the same values.
GF_HD bool interior(const GridDesc &g, int clo, int chi, int tlo, int thi) {
  return clo >= 1 && chi <= g.nr - 2 && chi < g.r0 && tlo >= 1 && thi <= g.nz - 2 && thi < g.z0;
}

// Byte offsets into an r strip (every row, columns from r0) and a z strip
// (rows from z0, every column), for cells inside the strip; 32-bit, like
// at()'s, so an access is a scalar base plus a vector offset.
GF_HD uint32_t rs_at(const GridDesc &g, int t, int c) { return (uint32_t)(t * g.rs_pitch + (c - g.r0)) * 4u; }
GF_HD uint32_t zs_at(const GridDesc &g, int t, int c) { return (uint32_t)((t - g.z0) * g.pitch + c) * 4u; }
GF_HD GF_GLOBAL float &strip(GF_GLOBAL float *base, uint32_t off) {
  return *(GF_GLOBAL float *)((GF_GLOBAL char *)base + off);
}

// A row of the coefficient table.
GF_HD GF_GLOBAL const float *row_coef(const GridDesc &g, int t) { return g.rowc + (size_t)t * kRowW; }

// Which auxiliaries exist at a velocity cell and at a stress cell (the
// strips, and for psiVzZ, psiVzR, psiVrZ the rows and columns whose update
// uses them).
GF_HD bool vp_trr_r(const GridDesc &g, int, int c) { return c >= g.first_rhalf; }
GF_HD bool vp_trz_r(const GridDesc &g, int, int c) { return c != 0 && c >= g.first_rint; }
GF_HD bool vp_trz_z(const GridDesc &g, int t, int) { return t >= g.first_zint; }
GF_HD bool vp_tzz_z(const GridDesc &g, int t, int) { return t >= g.first_zhalf; }
GF_HD bool sp_vr_r(const GridDesc &g, int, int c) { return c >= g.first_rint; }
GF_HD bool sp_vz_z(const GridDesc &g, int s, int) { return s > 0 && s >= g.first_zint; }
GF_HD bool sp_vz_r(const GridDesc &g, int s, int c) { return s <= g.nz - 2 && c <= g.nr - 2 && c >= g.first_rhalf; }
GF_HD bool sp_vr_z(const GridDesc &g, int s, int c) { return s <= g.nz - 2 && c <= g.nr - 2 && s >= g.first_zhalf; }

// Velocity auxiliaries of cell (t, c), a velocity cell, from buffer b.
GF_HD VPsi load_vpsi(const GridDesc &g, int b, int t, int c) {
  VPsi p = {0.f, 0.f, 0.f, 0.f};
  if (vp_trr_r(g, t, c)) p.trr_r = strip(g.vpsi_r[b][0], rs_at(g, t, c));
  if (vp_trz_r(g, t, c)) p.trz_r = strip(g.vpsi_r[b][1], rs_at(g, t, c));
  if (vp_trz_z(g, t, c)) p.trz_z = strip(g.vpsi_z[b][0], zs_at(g, t, c));
  if (vp_tzz_z(g, t, c)) p.tzz_z = strip(g.vpsi_z[b][1], zs_at(g, t, c));
  return p;
}

GF_HD void store_vpsi(const GridDesc &g, int b, int t, int c, const VPsi &p) {
  if (vp_trr_r(g, t, c)) strip(g.vpsi_r[b][0], rs_at(g, t, c)) = p.trr_r;
  if (vp_trz_r(g, t, c)) strip(g.vpsi_r[b][1], rs_at(g, t, c)) = p.trz_r;
  if (vp_trz_z(g, t, c)) strip(g.vpsi_z[b][0], zs_at(g, t, c)) = p.trz_z;
  if (vp_tzz_z(g, t, c)) strip(g.vpsi_z[b][1], zs_at(g, t, c)) = p.tzz_z;
}

// Stress auxiliaries of grid cell (s, c) from buffer b.
GF_HD SPsi load_spsi(const GridDesc &g, int b, int s, int c) {
  SPsi p = {0.f, 0.f, 0.f, 0.f};
  if (sp_vr_r(g, s, c)) p.vr_r = strip(g.spsi_r[b][0], rs_at(g, s, c));
  if (sp_vz_r(g, s, c)) p.vz_r = strip(g.spsi_r[b][1], rs_at(g, s, c));
  if (sp_vz_z(g, s, c)) p.vz_z = strip(g.spsi_z[b][0], zs_at(g, s, c));
  if (sp_vr_z(g, s, c)) p.vr_z = strip(g.spsi_z[b][1], zs_at(g, s, c));
  return p;
}

GF_HD void store_spsi(const GridDesc &g, int b, int s, int c, const SPsi &p) {
  if (sp_vr_r(g, s, c)) st_stream(&strip(g.spsi_r[b][0], rs_at(g, s, c)), p.vr_r);
  if (sp_vz_r(g, s, c)) st_stream(&strip(g.spsi_r[b][1], rs_at(g, s, c)), p.vz_r);
  if (sp_vz_z(g, s, c)) st_stream(&strip(g.spsi_z[b][0], zs_at(g, s, c)), p.vz_z);
  if (sp_vr_z(g, s, c)) st_stream(&strip(g.spsi_z[b][1], zs_at(g, s, c)), p.vr_z);
}

// New vr and vz at (t, c): Sim.velocities (fdtd.go:158-217) for one cell.
// Inputs are the old stresses around the cell (tzz_dn = tzz(t+1, c),
// trz_up = trz(t-1, c), *1 = column c+1, trzm = trz(t, c-1)); psi carries the
// cell's velocity auxiliaries in and their advanced values out.  Outside the
// updated cells the velocities are returned unchanged.
template <bool IN = false>
GF_HD void velocity_core(const GridDesc &g, int t, int c, const ColConst &cc, float trr, float trr1, float tqq,
                         float tqq1, float trz, float trzm, float trz_up, float tzz, float tzz_dn, float vr_old,
                         float vz_old, VPsi &psi, float &vr_new, float &vz_new) {
#pragma clang fp contract(on)
  vr_new = vr_old;
  vz_new = vz_old;
  if (!IN && !vel_cell(g, t, c)) return;
  const float invh = g.invh;
  GF_GLOBAL const float *rc = row_coef(g, t);

  float dtrr = (trr1 - trr) * invh;
  if (!IN && vp_trr_r(g, t, c)) {
    psi.trr_r = cc.brh * psi.trr_r + cc.arh * dtrr;
    dtrr += psi.trr_r;
  }
  float dtrz = !IN && t == 0 ? 2.f * trz * invh : (trz - trz_up) * invh;
  if (!IN && vp_trz_z(g, t, c)) {
    psi.trz_z = rc[kRowBzi] * psi.trz_z + rc[kRowAzi] * dtrz;
    dtrz += psi.trz_z;
  }
  float hoop = 0.5f * ((trr - tqq) + (trr1 - tqq1)) * cc.irh;
  vr_new = vr_old + rc[kRowCvr] * (dtrr + dtrz + hoop);

  float dtzz = (tzz_dn - tzz) * invh;
  if (!IN && vp_tzz_z(g, t, c)) {
    psi.tzz_z = rc[kRowBzh] * psi.tzz_z + rc[kRowAzh] * dtzz;
    dtzz += psi.tzz_z;
  }
  float radial;
  if (!IN && c == 0) {
    radial = 4.f * trz * invh;
  } else {
    float d = (trz - trzm) * invh;
    if (!IN && vp_trz_r(g, t, c)) {
      psi.trz_r = cc.bri * psi.trz_r + cc.ari * d;
      d += psi.trz_r;
    }
    radial = d + 0.5f * (trz + trzm) * cc.iri;
  }
  vz_new = vz_old + rc[kRowCvz] * (dtzz + radial);
}

// One-step form: the auxiliaries are read from buffer rb and, when `write`,
// the advanced ones stored to buffer rb ^ 1.  IN: the caller knows the cell
// is an updated cell outside every C-PML strip, off the axis and off the
// surface row (a tile of the interior), so no case is tested.
template <bool IN = false>
GF_HD void velocity(const GridDesc &g, int rb, int t, int c, bool write, const ColConst &cc, float trr,
                    float trr1, float tqq, float tqq1, float trz, float trzm, float trz_up, float tzz,
                    float tzz_dn, float vr_old, float vz_old, float &vr_new, float &vz_new) {
  VPsi psi = {0.f, 0.f, 0.f, 0.f};
  const bool cell = !IN && vel_cell(g, t, c);
  if (cell) psi = load_vpsi(g, rb, t, c);
  velocity_core<IN>(g, t, c, cc, trr, trr1, tqq, tqq1, trz, trzm, trz_up, tzz, tzz_dn, vr_old, vz_old, psi, vr_new,
                    vz_new);
  if (cell && write) store_vpsi(g, rb ^ 1, t, c, psi);
}

// Crank-Nicolson memory variable (fdtd.go:315-320), drive already formed.
GF_HD float relax_cn(float decay, float &r, float drive_term) {
#pragma clang fp contract(on)
  float next = decay * r - drive_term;
  float avg = 0.5f * (r + next);
  r = next;
  return avg;
}

// Memory variables of one cell, [component rr, qq, zz, rz][mechanism].  The
// kernel loads them a row ahead of use so their latency overlaps arithmetic.
template <int NM>
struct MemVars {
  float r[4][NM > 0 ? NM : 1];
};

template <int NM>
GF_HD void load_mem(const GridDesc &g, int b, int s, int c, MemVars<NM> &m) {
  if (NM == 0) return;
  const size_t ms = (size_t)g.mem_stride;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < NM; q++) m.r[k][q] = *at(g.mem[b] + (size_t)(k * NM + q) * ms, g, s, c);
}

template <int NM>
GF_HD void store_mem(const GridDesc &g, int b, int s, int c, const MemVars<NM> &m) {
  if (NM == 0) return;
  const size_t ms = (size_t)g.mem_stride;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < NM; q++) st_stream(at(g.mem[b] + (size_t)(k * NM + q) * ms, g, s, c), m.r[k][q]);
}

// New stresses at grid cell (s, c): Sim.stresses (fdtd.go:219-307) for one
// cell.  vr/vz are this step's new velocities: vrm = vr(s, c-1),
// vr_dn = vr(s+1, c), vz1 = vz(s, c+1), vz_up = vz(s-1, c).  trr..trz carry
// the old values in and the new ones out, psi and mem the cell's C-PML
// auxiliaries and memory variables.  NM = 0 is elastic; SC selects the
// reference's scalar relaxation (NM = 1, one drive for P and S alike),
// otherwise separate P and S drives per mechanism.
template <int NM, bool SC, bool IN = false>
GF_HD void stress_core(const GridDesc &g, int n, int s, int c, const ColConst &cc, float vr, float vrm, float vr_dn,
                       float vz, float vz1, float vz_up, float &trr, float &tqq, float &tzz, float &trz, SPsi &psi,
                       MemVars<NM> &mem) {
#pragma clang fp contract(on)
  const float invh = g.invh, dt = g.dt;
  GF_GLOBAL const float *rc = row_coef(g, s);
  const float lam = rc[kRowLam], l2m = rc[kRowL2m];

  float dvr, vrr;
  if (!IN && c == 0) {
    dvr = 2.f * vr * invh;
    vrr = dvr;
  } else {
    dvr = (vr - vrm) * invh;
    vrr = 0.5f * (vr + vrm) * cc.iri;
  }
  if (!IN && sp_vr_r(g, s, c)) {
    psi.vr_r = cc.bri * psi.vr_r + cc.ari * dvr;
    dvr += psi.vr_r;
  }
  float dvz;
  if (!IN && s == 0) {
    // F6: the applied rate enters unscaled by -1/A, exactly as fdtd.go:251-256.
    float applied = (c == 0 && n < g.nsrc) ? g.src_dtdt[n] : 0.f;
    dvz = (applied - lam * (dvr + vrr)) * rc[kRowIl2m];
  } else {
    dvz = (vz - vz_up) * invh;
    if (!IN && sp_vz_z(g, s, c)) {
      psi.vz_z = rc[kRowBzi] * psi.vz_z + rc[kRowAzi] * dvz;
      dvz += psi.vz_z;
    }
  }
  const float err = l2m * dvr + lam * (vrr + dvz);
  const float eqq = lam * (dvr + dvz) + l2m * vrr;

  if (NM == 0) {
    trr += dt * err;
    tqq += dt * eqq;
  } else if (SC) {
    trr += dt * (err + relax_cn(g.mem_decay, mem.r[0][0], g.mem_drive * err));
    tqq += dt * (eqq + relax_cn(g.mem_decay, mem.r[1][0], g.mem_drive * eqq));
  } else {
    const float mu2 = 2.f * rc[kRowMu];
    const float pb = l2m * (dvr + vrr + dvz);
    float sr = 0.f, sq = 0.f;
#pragma unroll
    for (int q = 0; q < NM; q++) {
      const float dP = rc[(kRowN + q)], dS = rc[(kRowN + NM + q)];
      sr += relax_cn(g.decay[q], mem.r[0][q], dP * pb - dS * (mu2 * (vrr + dvz)));
      sq += relax_cn(g.decay[q], mem.r[1][q], dP * pb - dS * (mu2 * (dvr + dvz)));
    }
    trr += dt * (err + sr);
    tqq += dt * (eqq + sq);
  }

  if (IN || s > 0) {
    const float ezz = lam * (dvr + vrr) + l2m * dvz;
    if (NM == 0) {
      tzz += dt * ezz;
    } else if (SC) {
      tzz += dt * (ezz + relax_cn(g.mem_decay, mem.r[2][0], g.mem_drive * ezz));
    } else {
      const float mu2 = 2.f * rc[kRowMu];
      const float pb = l2m * (dvr + vrr + dvz);
      float sz = 0.f;
#pragma unroll
      for (int q = 0; q < NM; q++) {
        const float dP = rc[(kRowN + q)], dS = rc[(kRowN + NM + q)];
        sz += relax_cn(g.decay[q], mem.r[2][q], dP * pb - dS * (mu2 * (dvr + vrr)));
      }
      tzz += dt * (ezz + sz);
    }
  } else if (c == 0) {
    tzz = n < g.nsrc ? g.src_tzz[n] : 0.f;
  }

  if (IN || (s <= g.nz - 2 && c <= g.nr - 2)) {
    float dvzdr = (vz1 - vz) * invh;
    if (!IN && sp_vz_r(g, s, c)) {
      psi.vz_r = cc.brh * psi.vz_r + cc.arh * dvzdr;
      dvzdr += psi.vz_r;
    }
    float dvrdz = (vr_dn - vr) * invh;
    if (!IN && sp_vr_z(g, s, c)) {
      psi.vr_z = rc[kRowBzh] * psi.vr_z + rc[kRowAzh] * dvrdz;
      dvrdz += psi.vr_z;
    }
    const float e = rc[kRowMuz] * (dvzdr + dvrdz);
    if (NM == 0) {
      trz += dt * e;
    } else if (SC) {
      trz += dt * (e + relax_cn(g.mem_decay, mem.r[3][0], g.mem_drive * e));
    } else {
      float sz = 0.f;
#pragma unroll
      for (int q = 0; q < NM; q++) sz += relax_cn(g.decay[q], mem.r[3][q], rc[(kRowN + NM + q)] * e);
      trz += dt * (e + sz);
    }
  }
}

// One-step form: auxiliaries updated in place in buffer mb; the memory
// variables come in (loaded ahead by the caller) and go out advanced, for
// the caller to store.
template <int NM, bool SC, bool IN = false>
GF_HD void stress(const GridDesc &g, int mb, int n, int s, int c, const ColConst &cc, float vr, float vrm,
                  float vr_dn, float vz, float vz1, float vz_up, float &trr, float &tqq, float &tzz, float &trz,
                  MemVars<NM> &mem) {
  SPsi psi = {0.f, 0.f, 0.f, 0.f};
  if (!IN) psi = load_spsi(g, mb, s, c);
  stress_core<NM, SC, IN>(g, n, s, c, cc, vr, vrm, vr_dn, vz, vz1, vz_up, trr, tqq, tzz, trz, psi, mem);
  if (!IN) store_spsi(g, mb, s, c, psi);
}

// Sim.SurfaceVelocity (fdtd.go:330-342) from this step's row-0 velocities.
GF_HD float surface_velocity(const GridDesc &g, int c, const ColConst &cc, float vr, float vrm, float vz) {
#pragma clang fp contract(on)
  float dvr, vrr;
  if (c == 0) {
    dvr = 2.f * vr * g.invh;
    vrr = dvr;
  } else {
    dvr = (vr - vrm) * g.invh;
    vrr = 0.5f * (vr + vrm) * cc.iri;
  }
  return vz - g.half_h * (g.svc * (dvr + vrr));
}

// |x| as ordered bits: every NaN orders above +Inf, which orders above every
// finite value, so an unsigned max propagates NaN the way math.Max does.
GF_HD uint32_t abs_bits(float x) {
  union {
    float f;
    uint32_t u;
  } v;
  v.f = x;
  return v.u & 0x7fffffffu;
}

}  // namespace gfk
