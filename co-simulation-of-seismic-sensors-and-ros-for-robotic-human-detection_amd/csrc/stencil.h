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
// derivatives.  Memory variables and the psi arrays of the stress
// derivatives are touched only by the owning lane and are updated in place.
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
constexpr int kMaxMech = 8;  // GSLS mechanisms
constexpr int kGuardRows = 2;   // zero rows above and below every field
constexpr int kGuardCols = 64;  // zero floats before row -kGuardRows

// Per-row coefficient table, float32, [kRow*][nz] then the GSLS drives
// [kRowN + q][nz] (P) and [kRowN + nmech + q][nz] (S).
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
  GF_GLOBAL float *spsi_r[2];      // psiVrR (r int), psiVzR (r half); in place
  GF_GLOBAL float *spsi_z[2];      // psiVzZ (z int), psiVrZ (z half); in place
  GF_GLOBAL float *mem;            // [4 components][nmech][nz][pitch]
  GF_GLOBAL const float *rowc;     // [kRowN + 2 nmech][nz]
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
// device the row start (less one column) is wave-uniform and lives in
// scalar registers, and the column is a 32-bit byte offset (c >= -1, so
// (c + 1) * 4 is non-negative): the access takes the scalar-base +
// vector-offset form and needs no 64-bit address arithmetic per lane.
GF_HD GF_GLOBAL float *at(GF_GLOBAL float *base, const GridDesc &g, int t, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
  GF_GLOBAL char *row = (GF_GLOBAL char *)(base + ((ptrdiff_t)t * g.pitch - 1));
  return (GF_GLOBAL float *)(row + (uint32_t)(c + 1) * 4u);
#else
  return base + ((ptrdiff_t)t * g.pitch + c);
#endif
}

// Field element at (row t, column c); rows -2..nz+1 and columns -1..pitch-1
// are inside the padded allocation and read as zero outside the grid.
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

// New vr and vz at (t, c): Sim.velocities (fdtd.go:158-217) for one cell.
// Inputs are the old stresses around the cell (tzz_dn = tzz(t+1, c),
// trz_up = trz(t-1, c), *1 = column c+1, trzm = trz(t, c-1)).  Velocity psi
// are read from buffer rb and, when `write`, stored to buffer rb ^ 1.
GF_HD void velocity(const GridDesc &g, int rb, int t, int c, bool write, const ColConst &cc, float trr,
                    float trr1, float tqq, float tqq1, float trz, float trzm, float trz_up, float tzz,
                    float tzz_dn, float vr_old, float vz_old, float &vr_new, float &vz_new) {
  vr_new = vr_old;
  vz_new = vz_old;
  if (t < 0 || t > g.nz - 2 || c < 0 || c > g.nr - 2) return;
  const float invh = g.invh;
  const size_t nz = (size_t)g.nz;
  const int wb = rb ^ 1;

  float dtrr = (trr1 - trr) * invh;
  if (c >= g.first_rhalf) {
    size_t k = (size_t)t * g.rs_pitch + (size_t)(c - g.r0);
    float psi = cc.brh * g.vpsi_r[rb][0][k] + cc.arh * dtrr;
    if (write) g.vpsi_r[wb][0][k] = psi;
    dtrr += psi;
  }
  float dtrz = t == 0 ? 2.f * trz * invh : (trz - trz_up) * invh;
  if (t >= g.first_zint) {
    size_t k = (size_t)(t - g.z0) * g.pitch + (size_t)c;
    float psi = g.rowc[kRowBzi * nz + t] * g.vpsi_z[rb][0][k] + g.rowc[kRowAzi * nz + t] * dtrz;
    if (write) g.vpsi_z[wb][0][k] = psi;
    dtrz += psi;
  }
  float hoop = 0.5f * ((trr - tqq) + (trr1 - tqq1)) * cc.irh;
  vr_new = vr_old + g.rowc[kRowCvr * nz + t] * (dtrr + dtrz + hoop);

  float dtzz = (tzz_dn - tzz) * invh;
  if (t >= g.first_zhalf) {
    size_t k = (size_t)(t - g.z0) * g.pitch + (size_t)c;
    float psi = g.rowc[kRowBzh * nz + t] * g.vpsi_z[rb][1][k] + g.rowc[kRowAzh * nz + t] * dtzz;
    if (write) g.vpsi_z[wb][1][k] = psi;
    dtzz += psi;
  }
  float radial;
  if (c == 0) {
    radial = 4.f * trz * invh;
  } else {
    float d = (trz - trzm) * invh;
    if (c >= g.first_rint) {
      size_t k = (size_t)t * g.rs_pitch + (size_t)(c - g.r0);
      float psi = cc.bri * g.vpsi_r[rb][1][k] + cc.ari * d;
      if (write) g.vpsi_r[wb][1][k] = psi;
      d += psi;
    }
    radial = d + 0.5f * (trz + trzm) * cc.iri;
  }
  vz_new = vz_old + g.rowc[kRowCvz * nz + t] * (dtzz + radial);
}

// Crank-Nicolson memory variable (fdtd.go:315-320), drive already formed.
GF_HD float relax_cn(float decay, float &r, float drive_term) {
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
GF_HD void load_mem(const GridDesc &g, int s, int c, MemVars<NM> &m) {
  if (NM == 0) return;
  const size_t ms = (size_t)g.mem_stride;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < NM; q++) m.r[k][q] = *at(g.mem + (size_t)(k * NM + q) * ms, g, s, c);
}

template <int NM>
GF_HD void store_mem(const GridDesc &g, int s, int c, const MemVars<NM> &m) {
  if (NM == 0) return;
  const size_t ms = (size_t)g.mem_stride;
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int q = 0; q < NM; q++) st_stream(at(g.mem + (size_t)(k * NM + q) * ms, g, s, c), m.r[k][q]);
}

// New stresses at (s, c): Sim.stresses (fdtd.go:219-307) for one cell.
// vr/vz are this step's new velocities: vrm = vr(s, c-1), vr_dn = vr(s+1, c),
// vz1 = vz(s, c+1), vz_up = vz(s-1, c).  trr..trz carry the old values in and
// the new ones out, mem the memory variables.  NM = 0 is elastic; SC selects
// the reference's scalar relaxation (NM = 1, one drive for P and S alike),
// otherwise separate P and S drives per mechanism.
template <int NM, bool SC>
GF_HD void stress(const GridDesc &g, int n, int s, int c, const ColConst &cc, float vr, float vrm, float vr_dn,
                  float vz, float vz1, float vz_up, float &trr, float &tqq, float &tzz, float &trz,
                  MemVars<NM> &mem) {
  const float invh = g.invh, dt = g.dt;
  const size_t nz = (size_t)g.nz;
  GF_GLOBAL const float *rc = g.rowc;
  const float lam = rc[kRowLam * nz + s], l2m = rc[kRowL2m * nz + s];

  float dvr, vrr;
  if (c == 0) {
    dvr = 2.f * vr * invh;
    vrr = dvr;
  } else {
    dvr = (vr - vrm) * invh;
    vrr = 0.5f * (vr + vrm) * cc.iri;
  }
  if (c >= g.first_rint) {
    GF_GLOBAL float *p = g.spsi_r[0] + (size_t)s * g.rs_pitch + (size_t)(c - g.r0);
    float psi = cc.bri * *p + cc.ari * dvr;
    *p = psi;
    dvr += psi;
  }
  float dvz;
  if (s == 0) {
    // F6: the applied rate enters unscaled by -1/A, exactly as fdtd.go:251-256.
    float applied = (c == 0 && n < g.nsrc) ? g.src_dtdt[n] : 0.f;
    dvz = (applied - lam * (dvr + vrr)) * rc[kRowIl2m * nz + s];
  } else {
    dvz = (vz - vz_up) * invh;
    if (s >= g.first_zint) {
      GF_GLOBAL float *p = g.spsi_z[0] + (size_t)(s - g.z0) * g.pitch + (size_t)c;
      float psi = rc[kRowBzi * nz + s] * *p + rc[kRowAzi * nz + s] * dvz;
      *p = psi;
      dvz += psi;
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
    const float mu2 = 2.f * rc[kRowMu * nz + s];
    const float pb = l2m * (dvr + vrr + dvz);
    float sr = 0.f, sq = 0.f;
#pragma unroll
    for (int q = 0; q < NM; q++) {
      const float dP = rc[(kRowN + q) * nz + s], dS = rc[(kRowN + NM + q) * nz + s];
      sr += relax_cn(g.decay[q], mem.r[0][q], dP * pb - dS * (mu2 * (vrr + dvz)));
      sq += relax_cn(g.decay[q], mem.r[1][q], dP * pb - dS * (mu2 * (dvr + dvz)));
    }
    trr += dt * (err + sr);
    tqq += dt * (eqq + sq);
  }

  if (s > 0) {
    const float ezz = lam * (dvr + vrr) + l2m * dvz;
    if (NM == 0) {
      tzz += dt * ezz;
    } else if (SC) {
      tzz += dt * (ezz + relax_cn(g.mem_decay, mem.r[2][0], g.mem_drive * ezz));
    } else {
      const float mu2 = 2.f * rc[kRowMu * nz + s];
      const float pb = l2m * (dvr + vrr + dvz);
      float sz = 0.f;
#pragma unroll
      for (int q = 0; q < NM; q++) {
        const float dP = rc[(kRowN + q) * nz + s], dS = rc[(kRowN + NM + q) * nz + s];
        sz += relax_cn(g.decay[q], mem.r[2][q], dP * pb - dS * (mu2 * (dvr + vrr)));
      }
      tzz += dt * (ezz + sz);
    }
  } else if (c == 0) {
    tzz = n < g.nsrc ? g.src_tzz[n] : 0.f;
  }

  if (s <= g.nz - 2 && c <= g.nr - 2) {
    float dvzdr = (vz1 - vz) * invh;
    if (c >= g.first_rhalf) {
      GF_GLOBAL float *p = g.spsi_r[1] + (size_t)s * g.rs_pitch + (size_t)(c - g.r0);
      float psi = cc.brh * *p + cc.arh * dvzdr;
      *p = psi;
      dvzdr += psi;
    }
    float dvrdz = (vr_dn - vr) * invh;
    if (s >= g.first_zhalf) {
      GF_GLOBAL float *p = g.spsi_z[1] + (size_t)(s - g.z0) * g.pitch + (size_t)c;
      float psi = rc[kRowBzh * nz + s] * *p + rc[kRowAzh * nz + s] * dvrdz;
      *p = psi;
      dvrdz += psi;
    }
    const float e = rc[kRowMuz * nz + s] * (dvzdr + dvrdz);
    if (NM == 0) {
      trz += dt * e;
    } else if (SC) {
      trz += dt * (e + relax_cn(g.mem_decay, mem.r[3][0], g.mem_drive * e));
    } else {
      float sz = 0.f;
#pragma unroll
      for (int q = 0; q < NM; q++) sz += relax_cn(g.decay[q], mem.r[3][q], rc[(kRowN + NM + q) * nz + s] * e);
      trz += dt * (e + sz);
    }
  }
}

// Sim.SurfaceVelocity (fdtd.go:330-342) from this step's row-0 velocities.
GF_HD float surface_velocity(const GridDesc &g, int c, const ColConst &cc, float vr, float vrm, float vz) {
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
