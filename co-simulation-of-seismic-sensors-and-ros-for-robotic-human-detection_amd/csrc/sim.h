// sim.h — the object behind a gf_sim handle (one model on one GPU).
#pragma once

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "kernels.h"
#include "model.h"

namespace gfs {

// Step kernels per captured graph (the time-step block a graph replays).
constexpr int kGraphSteps = 100;

int hip_fail(hipError_t e, const char *where);
int check_device(std::string &why);
// Launch shape when nothing better is known: time steps per launch (1 or 2,
// GEOFDTD_STEPS_PER_LAUNCH overrides) and the tile height for it
// (GEOFDTD_TILE_ROWS overrides), from the grid columns and rows to cover.
int choose_span(int fallback = 1);
int64_t split_min_boundary_tiles();
int choose_tile_rows(int span, int nr, int nz);
int choose_tile_rows_cells(int span, int64_t tiles_r_times_nz);

// Ordered |v| bits back to a speed; NaN bits stay NaN (math.Max semantics).
inline double bits_to_speed(uint32_t b) {
  float f;
  memcpy(&f, &b, sizeof f);
  return (double)f;
}

// Algorithmic HBM bytes of one fused step: every wavefield and memory
// variable read and written once in float32, plus 4 psi read and written per
// cell of each C-PML strip.
inline double step_bytes(const gfm::Resolved &R) {
  const double cells = (double)R.nr * R.nz;
  const int nm = R.lossy ? R.nmech : 0;
  const int r0 = R.first_rint < R.first_rhalf ? R.first_rint : R.first_rhalf;
  const int z0 = R.first_zint < R.first_zhalf ? R.first_zint : R.first_zhalf;
  const double strip = (double)R.nz * (R.nr - (r0 < R.nr ? r0 : R.nr)) + (double)R.nr * (R.nz - (z0 < R.nz ? z0 : R.nz));
  return 4.0 * 2.0 * (6 + 4 * nm) * cells + 32.0 * strip;
}

struct Sim {
  gfm::Resolved R;
  gfk::GridDesc g;  // host copy of the device descriptor
  int waves = 4;    // tiles (wavefronts) per step-kernel workgroup
  int span = 1;     // time steps per step-kernel launch (1 or 2)
  int buf = 0;      // the buffer holding the current wavefields
  int mbuf = 0;     // the buffer holding the memory variables and stress psi
  gfk::GridDesc *d_desc = nullptr;
  // span 2: the tiles split by gfk::step2_interior ([interior][the others])
  uint32_t *d_tile_list = nullptr;
  int64_t tile_list_cap = 0, n_interior = 0, n_boundary = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *arena = nullptr;
  size_t arena_bytes = 0;
  size_t state_bytes = 0;  // leading part of the arena that holds state
  float *rowc = nullptr, *colc = nullptr;
  float *d_src = nullptr;
  int32_t *d_rec = nullptr;
  float *d_traces = nullptr;
  int64_t rec_cap_cols = 0, trace_floats = 0;
  uint32_t *d_max = nullptr;
  int64_t max_slots = 0;
  void *d_scratch = nullptr;  // [0] graph base step, then small outputs
  void *h_scratch = nullptr;  // pinned mirror of small outputs
  int32_t *nbase_dev = nullptr;
  std::vector<double> force;
  int64_t step = 0;
  // timing of the step kernels since the last gf_sim_sync_timing
  bool timing_open = false;
  int64_t launches = 0;
  // graph of kGraphSteps steps
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_probe = false;

  ~Sim();
  int init(const gf_model &m, int32_t steps_limit);
  int upload_desc();
  int reset_state();
  int set_force(const double *force, int64_t n);
  int set_receivers(const int32_t *cols, int32_t ncols, int64_t cap, int64_t rec_base);
  int reset_max_slots(int64_t steps);
  gfk::StepLaunch step_launch() const;
  int launch_one(int64_t n, int count, bool probe);
  int advance(int64_t n, bool use_graph, bool probe);
  int ensure_graph(bool probe);
  int set_launch(int span, int th);
  int set_tile_rows(int th) { return set_launch(span, th); }
  int autotune(int trial_steps, int *chosen);
};

}  // namespace gfs
