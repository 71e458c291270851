// sampler.cpp — seeded soil models for batched bank generation.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): a counter-based generator, so model `index` is a
// pure function of (seed, index) with no shared stream state, and any subset
// of a batch can be regenerated exactly.  Counter (index, block, 0, 0) with
// key (seed low, seed high); uniforms are (x + 0.5) / 2^32.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "../../include/geofdtd.h"

namespace {

constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;  // Philox round multipliers
constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;  // key schedule (Weyl) constants

void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)kM0 * c0, p1 = (uint64_t)kM1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += kW0;
    k1 += kW1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

struct Draws {
  uint32_t key[2];
  uint32_t index;
  uint32_t stream = 0;  // separates the soil families drawn from one seed
  void block(uint32_t b, double u[4]) const {
    const uint32_t ctr[4] = {index, b, stream, 0};
    uint32_t x[4];
    philox(ctr, key, x);
    for (int k = 0; k < 4; k++) u[k] = ((double)x[k] + 0.5) / 4294967296.0;
  }
};

}  // namespace

extern "C" void gf_philox4x32(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
  philox(counter, key, out);
}

extern "C" int gf_soil_sample(uint64_t seed, uint32_t index, gf_soil *out) {
  if (!out) return GF_ERR_ARGUMENT;
  memset(out, 0, sizeof *out);
  Draws d{{(uint32_t)seed, (uint32_t)(seed >> 32)}, index};
  double u[4];
  d.block(0, u);
  const int above = 1 + std::min(2, (int)(u[0] * 3));  // layers over the half-space
  out->n_layers = above + 1;
  double vs_min = 1e30;
  for (int l = 0; l < out->n_layers; l++) {
    d.block(1 + l, u);
    gf_layer &L = out->layers[l];
    L.vs_mps = 120 + 280 * u[0];
    L.vp_mps = L.vs_mps * (1.7 + 0.8 * u[1]);
    L.density_kgm3 = 1600 + 500 * u[2];
    L.qs = 10 + 70 * u[3];
    L.qp = 2 * L.qs;
    d.block(5 + l, u);
    L.thickness_m = l < above ? 0.5 + 4.5 * u[0] : 0.0;
    vs_min = std::min(vs_min, L.vs_mps);
  }
  const double h = 0.9 * vs_min / (100.0 * 10.0);
  const int pml = 20;
  const int nr = std::max(600, std::min(2048, (int)lround(218.0 / h) + 1 + pml));
  const int nz = std::max(300, std::min(1024, (int)lround(108.0 / h) + 1 + pml));
  out->spacing_m = h;
  out->pml_cells = pml;
  out->max_range_m = (double)(nr - 1 - pml) * h;
  out->depth_m = (double)(nz - 1 - pml) * h;
  return GF_OK;
}

extern "C" int gf_soil_sample_layered(uint64_t seed, uint32_t index, gf_soil *out, double params[10]) {
  if (!out || !params) return GF_ERR_ARGUMENT;
  memset(out, 0, sizeof *out);
  Draws d{{(uint32_t)seed, (uint32_t)(seed >> 32)}, index, 1};
  double u[4], w[4], q[4];
  d.block(0, u);
  d.block(1, w);
  d.block(2, q);
  const double vs[3] = {120 + 80 * u[0], 200 + 120 * u[1], 320 + 160 * u[2]};
  const double qs[3] = {10 + 30 * w[0], 20 + 40 * w[1], 40 + 60 * w[2]};
  const double th[2] = {0.5 + 2.5 * q[0], 2.0 + 4.0 * q[1]};
  const double ratio = 1.7 + 0.8 * q[2], rho0 = 1600 + 200 * q[3];
  out->n_layers = 3;
  for (int l = 0; l < 3; l++) {
    gf_layer &L = out->layers[l];
    L.thickness_m = l < 2 ? th[l] : 0.0;
    L.vs_mps = vs[l];
    L.vp_mps = ratio * vs[l];
    L.density_kgm3 = rho0 + 150.0 * l;
    L.qs = qs[l];
    L.qp = 2 * qs[l];
  }
  const double h = 0.9 * vs[0] / (100.0 * 10.0);
  const int pml = 20;
  // 80 m of interior for receivers out to 50 m: surface waves that reach the
  // C-PML excite a slowly growing mode there in soils with a thick soft top
  // layer, and 30 m of distance keeps it below 1e-3 of the signal at 50 m
  // over a 1 s run
  const int nr = (int)lround(80.0 / h) + 1 + pml, nz = (int)lround(30.0 / h) + 1 + pml;
  out->spacing_m = h;
  out->pml_cells = pml;
  out->max_range_m = (double)(nr - 1 - pml) * h;
  out->depth_m = (double)(nz - 1 - pml) * h;
  const double p[10] = {vs[0], vs[1], vs[2], qs[0], qs[1], qs[2], th[0], th[1], ratio, rho0};
  memcpy(params, p, sizeof p);
  return GF_OK;
}
