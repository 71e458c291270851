// api.cpp — the C ABI of libgeofdtd (include/geofdtd.h).
//
// A gf_sim owns one device arena holding the double-buffered wavefields, the
// C-PML strips, the memory variables, the coefficient tables and the receiver
// buffers of one model, plus a HIP stream at default priority.  gf_run and
// gf_run_batch are built on the same pieces.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "kernels.h"
#include "layout.h"
#include "march2.h"
#include "model.h"
#include "sim.h"
#include "tune.h"

using gfk::GridDesc;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int ok() {
  g_err.clear();
  return GF_OK;
}

}  // namespace

namespace gfs {

int hip_fail(hipError_t e, const char *where) {
  return set_err(GF_ERR_DEVICE, std::string("geofdtd: HIP error in ") + where + ": " + hipGetErrorString(e));
}

#define HIPCHK(expr)                                         \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr);   \
  } while (0)

int check_device(std::string &why) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    why = "geofdtd: no HIP device is visible";
    return GF_ERR_DEVICE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
    why = "geofdtd: cannot query the HIP device";
    return GF_ERR_DEVICE;
  }
  if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    why = std::string("geofdtd: built for gfx950 only, found ") + p.gcnArchName;
    return GF_ERR_DEVICE;
  }
  return GF_OK;
}

// Time steps per launch when nothing better is known: `fallback`, unless
// GEOFDTD_STEPS_PER_LAUNCH asks for one or two.
int choose_span(int fallback) {
  if (const char *e = getenv("GEOFDTD_STEPS_PER_LAUNCH")) {
    const int v = atoi(e);
    if (v == 1 || v == 2) return v;
  }
  return fallback;
}

// Two-step launches are split by tile kind when at least this many of their
// tiles are not interior (gfk::kSplitMinBoundaryTiles);
// GEOFDTD_SPLIT_MIN_TILES overrides it (0 always splits).
int64_t split_min_boundary_tiles() {
  if (const char *e = getenv("GEOFDTD_SPLIT_MIN_TILES")) return atoll(e);
  return gfk::kSplitMinBoundaryTiles;
}

// Tile height when nothing better is known.  One-step kernel: the tallest of
// the unrolled variants (8, 4, 2, 1 rows) that still puts about a dozen waves
// on each of the 1024 SIMDs, since the step is bound by how many loads are
// in flight.  Two-step kernel: the tallest of 32, 16, 8 rows that gives four
// waves per SIMD (its march holds a row of loads in flight; shorter tiles
// repeat more of the six halo rows).  GEOFDTD_TILE_ROWS overrides.
int choose_tile_rows_cells(int span, int64_t tiles_r_times_nz) {
  if (const char *e = getenv("GEOFDTD_TILE_ROWS")) {
    int v = atoi(e);
    if (v >= 1 && v <= 4096) return v;
  }
  if (span == 2) {
    for (int th = 32; th > 8; th /= 2)
      if (tiles_r_times_nz / th >= 1024 * 4) return th;
    return 8;
  }
  const int64_t want_waves = 1024 * 12;
  for (int th = 8; th > 1; th /= 2)
    if (tiles_r_times_nz / th >= want_waves) return th;
  return 1;
}

int choose_tile_rows(int span, int nr, int nz) {
  const int tiles_r = (nr + gfl::tile_cols(span) - 1) / gfl::tile_cols(span);
  return choose_tile_rows_cells(span, (int64_t)tiles_r * nz);
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Sim::~Sim() {
  if (stream) (void)hipStreamSynchronize(stream);  // nothing of ours may still run when it is freed
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (arena) hipFree(arena);
  if (d_desc) hipFree(d_desc);
  if (d_tile_list) hipFree(d_tile_list);
  if (d_scratch) hipFree(d_scratch);
  if (d_src) hipFree(d_src);
  if (d_rec) hipFree(d_rec);
  if (d_traces) hipFree(d_traces);
  if (d_max) hipFree(d_max);
  if (h_scratch) hipHostFree(h_scratch);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (stream) hipStreamDestroy(stream);
}

int Sim::set_launch(int sp, int th) {
  if (th < 1) return set_err(GF_ERR_ARGUMENT, "geofdtd: tile rows must be positive");
  if (sp != 1 && sp != 2) return set_err(GF_ERR_ARGUMENT, "geofdtd: steps per launch must be 1 or 2");
  // launches in flight read the graph and the descriptor about to change
  HIPCHK(hipStreamSynchronize(stream));
  span = sp;
  gfl::set_tiling(g, span, th);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  return upload_desc();
}

// Time each candidate launch shape (tile height, tiles per workgroup) on
// this model, keep the fastest, and put the model back to step 0 with zero
// state.  *chosen receives the tile height.
int Sim::autotune(int trial_steps, int *chosen) {
  const int64_t s0 = step;
  gft::Key key{0, g.lossy ? g.nmech : 0, g.scalar, 1, (int64_t)log2((double)R.nr * R.nz)};
  int err = GF_OK;
  auto prepare = [&](int code) {
    const gft::Shape sh = gft::Shape::of(code);
    waves = sh.waves;
    err = set_launch(sh.span, sh.th);
    return err == GF_OK;
  };
  // trial state is discarded, but each trial reads the buffers the one
  // before wrote
  int tb = buf, tm = mbuf;
  auto launch = [&](int code, int k) {
    gfk::StepLaunch L = step_launch();
    L.n_off = (int32_t)(s0 + (int64_t)k * gft::Shape::of(code).span);
    L.rb_off = tb;
    L.mb_off = tm;
    L.count = L.span;
    tb ^= 1;
    if (L.span == 2) tm ^= 1;
    return gfk::launch_step(L, stream);
  };
  auto work = [&](int code) { return gft::Shape::of(code).span * step_bytes(R); };
  int code = gft::cached(key);
  const int sp0 = choose_span();
  const int preferred = gft::Shape{choose_tile_rows(sp0, R.nr, R.nz), 4, sp0}.code();
  if (!code)
    code = gft::tune(key, stream, trial_steps, gft::shape_candidates(), preferred, prepare, launch, work, nullptr);
  if (!code) return err ? err : set_err(GF_ERR_DEVICE, "geofdtd: autotune found no working tile height");
  const int best = gft::Shape::of(code).th;
  waves = gft::Shape::of(code).waves;
  int rc = set_launch(gft::Shape::of(code).span, best);
  if (rc) return rc;
  if ((rc = reset_state())) return rc;
  if (d_traces && trace_floats) HIPCHK(gfk::launch_fill_f32(d_traces, trace_floats, 0.f, stream));
  if (d_max && max_slots) HIPCHK(gfk::launch_fill_u32(d_max, max_slots, 0u, stream));
  HIPCHK(hipStreamSynchronize(stream));
  if (chosen) *chosen = best;
  return GF_OK;
}

// Zero the state part of the arena (fields, psi strips, memory variables) with
// a fill kernel on the solver's stream; tables and buffers stay.
int Sim::reset_state() {
  HIPCHK(gfk::launch_fill_f32((float *)arena, (int64_t)(state_bytes / sizeof(float)), 0.f, stream));
  step = 0;
  buf = 0;
  mbuf = 0;
  return GF_OK;
}

int Sim::upload_desc() {
  HIPCHK(hipStreamSynchronize(stream));
  HIPCHK(hipMemcpy(d_desc, &g, sizeof g, hipMemcpyHostToDevice));
  n_interior = n_boundary = 0;
  if (span == 2) {
    // the interior tiles and the others as two lists (two launches)
    std::vector<uint32_t> in, bd;
    for (int lt = 0; lt < g.tiles_r * g.tiles_z; lt++) {
      const gfk::Tile2 t = gfk::step2_tile_at(g, lt);
      (gfk::step2_interior(g, t.i0, t.j0, t.j1) ? in : bd).push_back((uint32_t)(g.tile_base + lt));
    }
    n_interior = (int64_t)in.size();
    n_boundary = (int64_t)bd.size();
    const int64_t n = n_interior + n_boundary;
    if (n > tile_list_cap) {
      if (d_tile_list) (void)hipFree(d_tile_list);
      d_tile_list = nullptr;
      HIPCHK(hipMalloc(&d_tile_list, sizeof(uint32_t) * (size_t)n));
      tile_list_cap = n;
    }
    in.insert(in.end(), bd.begin(), bd.end());
    if (n > 0) HIPCHK(hipMemcpy(d_tile_list, in.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice));
  }
  return GF_OK;
}

int Sim::init(const gf_model &m, int32_t steps_limit) {
  std::string err;
  int rc = gfm::resolve(m, R, err);
  if (rc) return set_err(rc, err);
  if ((rc = check_device(err))) return set_err(rc, err);
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&ev0));
  HIPCHK(hipEventCreate(&ev1));

  memset(&g, 0, sizeof g);
  gfl::Layout L;
  span = choose_span();
  gfl::plan(R, span, choose_tile_rows(span, R.nr, R.nz), g, L);
  g.tile_base = 0;
  g.steps = steps_limit;
  const size_t rs = L.rs, zs = L.zs;
  const int nmt = L.nmech_table;
  struct Part {
    size_t floats;
    size_t off;
  };
  std::vector<Part> parts;
  auto add = [&](size_t floats) {
    parts.push_back({floats, 0});
    return parts.size() - 1;
  };
  size_t p_fields = add(12 * L.field_floats);
  size_t p_vr = add(2 * 2 * rs), p_vz = add(2 * 2 * zs), p_sr = add(2 * 2 * rs), p_sz = add(2 * 2 * zs);
  size_t p_mem = add(std::max<size_t>(2 * L.mem_floats, 1)), p_rowc = add(L.rowc_floats), p_colc = add(L.colc_floats);
  size_t total = 0;
  for (auto &p : parts) {
    p.off = total;
    total += align_up(p.floats * sizeof(float), 256);
  }
  arena_bytes = total;
  state_bytes = parts[p_rowc].off;  // fields, strips and memory come first
  HIPCHK(hipMalloc(&arena, total));
  HIPCHK(hipMemsetAsync(arena, 0, total, stream));
  char *base = (char *)arena;
  auto at = [&](size_t i) { return (float *)(base + parts[i].off); };
  for (int b = 0; b < 2; b++)
    for (int f = 0; f < 6; f++)
      g.f[b][f] = at(p_fields) + (size_t)(b * 6 + f) * L.field_floats + L.field_origin;
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 2; k++) {
      g.vpsi_r[b][k] = at(p_vr) + (size_t)(b * 2 + k) * rs;
      g.vpsi_z[b][k] = at(p_vz) + (size_t)(b * 2 + k) * zs;
    }
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 2; k++) {
      g.spsi_r[b][k] = at(p_sr) + (size_t)(b * 2 + k) * rs;
      g.spsi_z[b][k] = at(p_sz) + (size_t)(b * 2 + k) * zs;
    }
  for (int b = 0; b < 2; b++) g.mem[b] = R.lossy ? at(p_mem) + (size_t)b * L.mem_floats : nullptr;
  rowc = at(p_rowc);
  colc = at(p_colc);
  g.rowc = rowc;
  g.colc = colc;

  // ---- coefficient tables, FP64 on the device ----
  gfk::TableParams P;
  memset(&P, 0, sizeof P);
  P.nr = g.nr;
  P.nz = g.nz;
  P.pitch = g.pitch;
  P.nlayers = (int32_t)R.layers.size();
  P.nmech_table = nmt;
  P.h = R.h;
  P.dt = R.dt;
  P.max_range = R.max_range;
  P.depth = R.depth;
  P.pml_thickness = R.pml_thickness;
  P.d0 = R.d0;
  P.amax = R.amax;
  gfk::LayerParams *d_layers = nullptr;
  HIPCHK(hipMalloc(&d_layers, sizeof(gfk::LayerParams) * R.layers.size()));
  HIPCHK(hipMemcpyAsync(d_layers, R.layers.data(), sizeof(gfk::LayerParams) * R.layers.size(),
                        hipMemcpyHostToDevice, stream));
  HIPCHK(gfk::launch_build_tables(P, d_layers, rowc, colc, stream));
  HIPCHK(hipStreamSynchronize(stream));
  hipFree(d_layers);

  HIPCHK(hipMalloc(&d_desc, sizeof(GridDesc)));
  HIPCHK(hipMalloc(&d_scratch, 64));
  HIPCHK(hipHostMalloc(&h_scratch, 64, hipHostMallocDefault));
  nbase_dev = (int32_t *)d_scratch;
  HIPCHK(hipMemsetAsync(d_scratch, 0, 64, stream));
  return upload_desc();
}

// Replace the source series: tzz(0,0) = -F[n]/A and the applied rate
// (F[n+1]-F[n])/dt, both formed in FP64 (fdtd.go:152-154) and rounded once.
int Sim::set_force(const double *force, int64_t n) {
  this->force.assign(force, force + std::max<int64_t>(n, 0));
  const int64_t ns = (int64_t)this->force.size();
  std::vector<float> tzz(std::max<int64_t>(ns, 1)), rate(std::max<int64_t>(ns, 1));
  for (int64_t i = 0; i < ns; i++) {
    const double f0 = this->force[i], f1 = i + 1 < ns ? this->force[i + 1] : 0.0;
    tzz[i] = (float)(-f0 / R.src_area);
    rate[i] = (float)((f1 - f0) / R.dt);
  }
  HIPCHK(hipStreamSynchronize(stream));
  if (d_src) hipFree(d_src);
  d_src = nullptr;
  HIPCHK(hipMalloc(&d_src, 2 * sizeof(float) * tzz.size()));
  HIPCHK(hipMemcpy(d_src, tzz.data(), sizeof(float) * tzz.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_src + tzz.size(), rate.data(), sizeof(float) * rate.size(), hipMemcpyHostToDevice));
  g.src_tzz = d_src;
  g.src_dtdt = d_src + tzz.size();
  g.nsrc = (int32_t)std::min<int64_t>(ns, INT32_MAX);
  return upload_desc();
}

int Sim::set_receivers(const int32_t *cols, int32_t ncols, int64_t cap, int64_t rec_base) {
  HIPCHK(hipStreamSynchronize(stream));
  if (ncols > 0) {
    if ((int64_t)ncols > rec_cap_cols || cap * ncols > trace_floats) {
      if (d_rec) hipFree(d_rec);
      if (d_traces) hipFree(d_traces);
      d_rec = nullptr;
      d_traces = nullptr;
      HIPCHK(hipMalloc(&d_rec, sizeof(int32_t) * ncols));
      HIPCHK(hipMalloc(&d_traces, sizeof(float) * (size_t)cap * ncols));
      rec_cap_cols = ncols;
      trace_floats = cap * ncols;
    }
    HIPCHK(hipMemcpy(d_rec, cols, sizeof(int32_t) * ncols, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(d_traces, 0, sizeof(float) * (size_t)cap * ncols));
  }
  g.rec_col = d_rec;
  g.traces = d_traces;
  g.nrec = ncols;
  g.trace_cap = cap;
  g.rec_base = rec_base;
  return upload_desc();
}

// Divergence probe slots: one uint per 256 steps, cleared by a fill kernel.
int Sim::reset_max_slots(int64_t steps) {
  const int64_t slots = steps / 256 + 2;
  HIPCHK(hipStreamSynchronize(stream));
  if (slots > max_slots) {
    if (d_max) hipFree(d_max);
    d_max = nullptr;
    HIPCHK(hipMalloc(&d_max, sizeof(uint32_t) * slots));
    max_slots = slots;
  }
  HIPCHK(gfk::launch_fill_u32(d_max, max_slots, 0u, stream));
  g.maxbits = d_max;
  return upload_desc();
}

gfk::StepLaunch Sim::step_launch() const {
  gfk::StepLaunch L;
  L.nmech = g.lossy ? g.nmech : 0;
  L.scalar = g.scalar != 0;
  L.span = span;
  L.th = g.th;
  L.waves = waves;
  L.d = d_desc;
  L.tiles = (int64_t)g.tiles_r * g.tiles_z;
  if (span == 2 && d_tile_list && n_interior + n_boundary == L.tiles && n_boundary >= split_min_boundary_tiles()) {
    L.list_in = d_tile_list;
    L.n_in = n_interior;
    L.list_bd = d_tile_list + n_interior;
    L.n_bd = n_boundary;
  }
  return L;
}

// One launch of `count` steps from step n, reading the current buffer.
int Sim::launch_one(int64_t n, int count, bool probe) {
  gfk::StepLaunch L = step_launch();
  L.n_off = (int32_t)n;
  L.rb_off = buf;
  L.mb_off = mbuf;
  L.probe = probe ? 1 : 0;
  L.count = count;
  HIPCHK(gfk::launch_step(L, stream));
  buf ^= 1;
  if (span == 2) mbuf ^= 1;
  return GF_OK;
}

// Steps [step, step + n) as plain launches or graph replays.  With `probe`,
// the kernel runs the divergence probe on steps that are multiples of 256.
int Sim::advance(int64_t n, bool use_graph, bool probe) {
  if (n <= 0) return GF_OK;
  if (step + n > (int64_t)INT32_MAX) return set_err(GF_ERR_ARGUMENT, "geofdtd: step counter would overflow");
  if (probe && !g.maxbits) return set_err(GF_ERR_ARGUMENT, "geofdtd: probe slots not set");
  HIPCHK(hipEventRecord(timing_open ? ev1 : ev0, stream));
  if (!timing_open) timing_open = true;
  int64_t done = 0;
  if (use_graph && n >= kGraphSteps) {
    int rc = ensure_graph(probe);
    if (rc) return rc;
    // a block is an even number of launches: the current buffers are the
    // same after it
    while (n - done >= kGraphSteps) {
      const int64_t n0 = step + done;
      HIPCHK(gfk::launch_set_base(nbase_dev, (int32_t)n0, buf, mbuf, stream));
      HIPCHK(hipGraphLaunch(exec, stream));
      done += kGraphSteps;
      launches += kGraphSteps / span;
    }
  }
  while (done < n) {
    const int count = (int)std::min<int64_t>(span, n - done);
    int rc = launch_one(step + done, count, probe);
    if (rc) return rc;
    done += count;
    launches++;
  }
  step += n;
  HIPCHK(hipEventRecord(ev1, stream));
  return GF_OK;
}

// A graph of kGraphSteps steps (kGraphSteps / span launches); node k runs
// step nbase[0] + k * span from wavefield buffer nbase[1] ^ (k & 1) and
// memory-variable buffer nbase[2] (^ (k & 1) for two-step launches), where
// nbase is set by a one-thread kernel before each replay (kernel arguments of
// a captured node are fixed, the step number and the buffers are not).
int Sim::ensure_graph(bool probe) {
  if (exec && graph_probe == probe) return GF_OK;
  // the old graph may still be running: wait before destroying it
  HIPCHK(hipStreamSynchronize(stream));
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kGraphSteps / span; k++) {
    gfk::StepLaunch L = step_launch();
    L.n_off = k * span;
    L.rb_off = k & 1;
    L.mb_off = span == 2 ? (k & 1) : 0;
    L.nbase = nbase_dev;
    L.probe = probe ? 1 : 0;
    L.count = span;
    hipError_t e = gfk::launch_step(L, stream);
    if (e != hipSuccess) {
      hipGraph_t tmp;
      hipStreamEndCapture(stream, &tmp);
      if (tmp) hipGraphDestroy(tmp);
      return hip_fail(e, "graph capture");
    }
  }
  HIPCHK(hipStreamEndCapture(stream, &graph));
  HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  graph_probe = probe;
  return GF_OK;
}

}  // namespace gfs

using gfs::Sim;

// =================================================================== C ABI

extern "C" {

const char *gf_last_error(void) { return g_err.c_str(); }
int gf_abi_version(void) { return GF_ABI_VERSION; }

int gf_visco_fit(double q, double f0, gf_sls *out) {
  if (!out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null output");
  std::string err;
  int rc = gfm::visco_fit(q, f0, *out, err);
  return rc ? set_err(rc, err) : ok();
}

int gf_spacing_for(const gf_model *m, double max_freq, double cpw, double *out) {
  if (!m || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  std::string err;
  int rc = gfm::spacing_for(*m, max_freq, cpw, *out, err);
  return rc ? set_err(rc, err) : ok();
}

int gf_ricker(double peak, double f0, double dt, int32_t n, double *out) {
  if (f0 <= 0 || dt <= 0 || n <= 0)
    return set_err(GF_ERR_INVALID, "fdtd: Ricker needs a positive frequency, step and length");
  if (!out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null output");
  gfm::ricker(peak, f0, dt, n, out);
  return ok();
}

int gf_ricker_spectrum(double peak, double f0, double f, double *re, double *im) {
  if (!re || !im) return set_err(GF_ERR_ARGUMENT, "geofdtd: null output");
  gfm::ricker_spectrum(peak, f0, f, *re, *im);
  return ok();
}

int gf_stack_validate(const gf_layer *stack, int32_t n) {
  std::string err;
  int rc = gfm::validate_stack(stack, n, err);
  return rc ? set_err(rc, err) : ok();
}

int gf_model_grid(const gf_model *m, double *dt, int32_t *nr, int32_t *nz) {
  if (!m) return set_err(GF_ERR_ARGUMENT, "geofdtd: null model");
  gfm::Resolved R;
  std::string err;
  int rc = gfm::resolve(*m, R, err);
  if (rc) return set_err(rc, err);
  if (dt) *dt = R.dt;
  if (nr) *nr = R.nr;
  if (nz) *nz = R.nz;
  return ok();
}

int gf_sim_reset(gf_sim *sp) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  int rc = s->reset_state();
  return rc ? rc : ok();
}

int gf_sim_autotune(gf_sim *sp, int32_t trial_steps, int32_t *tile_rows) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || trial_steps < 1) return set_err(GF_ERR_ARGUMENT, "geofdtd: bad autotune arguments");
  int chosen = 0;
  int rc = s->autotune(trial_steps, &chosen);
  if (rc) return rc;
  if (tile_rows) *tile_rows = chosen;
  return ok();
}

int gf_sim_set_tile_rows(gf_sim *sp, int32_t tile_rows) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  int rc = s->set_tile_rows(tile_rows);
  return rc ? rc : ok();
}

int gf_sim_set_launch(gf_sim *sp, int32_t steps_per_launch, int32_t tile_rows) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  int rc = s->set_launch(steps_per_launch, tile_rows);
  return rc ? rc : ok();
}

int gf_sim_launch_shape(const gf_sim *sp, int32_t *steps_per_launch, int32_t *tile_rows, int32_t *waves) {
  const Sim *s = reinterpret_cast<const Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  if (steps_per_launch) *steps_per_launch = s->span;
  if (tile_rows) *tile_rows = s->g.th;
  if (waves) *waves = s->waves;
  return ok();
}

int gf_sim_create(const gf_model *m, gf_sim **out) {
  if (!m || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = nullptr;
  Sim *s = new Sim();
  int rc = s->init(*m, INT32_MAX);
  if (rc) {
    std::string keep = g_err;
    delete s;
    return set_err(rc, keep);
  }
  *out = reinterpret_cast<gf_sim *>(s);
  return ok();
}

void gf_sim_destroy(gf_sim *s) { delete reinterpret_cast<Sim *>(s); }

int gf_sim_info(const gf_sim *sp, double *dt, int32_t *nr, int32_t *nz) {
  const Sim *s = reinterpret_cast<const Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  if (dt) *dt = s->R.dt;
  if (nr) *nr = s->R.nr;
  if (nz) *nz = s->R.nz;
  return ok();
}

int gf_sim_time(const gf_sim *sp, double *out) {
  const Sim *s = reinterpret_cast<const Sim *>(sp);
  if (!s || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = ((double)s->step - 0.5) * s->R.dt;
  return ok();
}

int gf_sim_drive(gf_sim *sp, const double *force, int64_t n) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || (n > 0 && !force)) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  int rc = s->set_force(force, n);
  return rc ? rc : ok();
}

int gf_sim_column(const gf_sim *sp, double r, int32_t *col, double *snapped) {
  const Sim *s = reinterpret_cast<const Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  int i = (int)round(r / s->R.h);
  i = std::max(0, std::min(i, s->R.nr - 2));
  if (col) *col = i;
  if (snapped) *snapped = (double)i * s->R.h;
  return ok();
}

int gf_sim_step(gf_sim *sp) { return gf_sim_step_n(sp, 1, nullptr, 0, nullptr); }

int gf_sim_step_n(gf_sim *sp, int32_t n, const int32_t *cols, int32_t ncols, double *out) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  if (n < 0 || ncols < 0 || (ncols > 0 && (!cols || !out)))
    return set_err(GF_ERR_ARGUMENT, "geofdtd: bad step_n arguments");
  for (int32_t i = 0; i < ncols; i++)
    if (cols[i] < 0 || cols[i] > s->R.nr - 2) return set_err(GF_ERR_ARGUMENT, "geofdtd: column out of range");
  int rc = s->set_receivers(cols, ncols, std::max<int64_t>(n, 1), s->step);
  if (rc) return rc;
  if ((rc = s->advance(n, false, false))) return rc;
  if (ncols > 0) {
    std::vector<float> tmp((size_t)n * ncols);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(tmp.data(), s->d_traces, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < tmp.size(); i++) out[i] = tmp[i];
  }
  if ((rc = s->set_receivers(nullptr, 0, 0, 0))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return ok();
}

int gf_sim_surface_velocity(gf_sim *sp, int32_t col, double *out) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  if (col < 0 || col > s->R.nr - 1) return set_err(GF_ERR_ARGUMENT, "geofdtd: column out of range");
  HIPCHK(gfk::launch_surface_velocity(s->d_desc, (int32_t)s->buf, col, (float *)s->d_scratch + 4, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_scratch, (float *)s->d_scratch + 4, 2 * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *out = ((float *)s->h_scratch)[0];
  return ok();
}

int gf_sim_radial_velocity(gf_sim *sp, int32_t col, double *out) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  if (col < 0 || col > s->R.nr - 1) return set_err(GF_ERR_ARGUMENT, "geofdtd: column out of range");
  HIPCHK(gfk::launch_surface_velocity(s->d_desc, (int32_t)s->buf, col, (float *)s->d_scratch + 4, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_scratch, (float *)s->d_scratch + 4, 2 * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *out = ((float *)s->h_scratch)[1];
  return ok();
}

int gf_sim_max_speed(gf_sim *sp, double *out) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || !out) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  uint32_t *slot = (uint32_t *)s->d_scratch + 8;
  HIPCHK(gfk::launch_fill_u32(slot, 1, 0u, s->stream));
  HIPCHK(gfk::launch_max_speed(s->d_desc, (int32_t)s->buf, slot, (int64_t)s->R.nz * s->g.pitch, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_scratch, slot, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *out = gfs::bits_to_speed(((uint32_t *)s->h_scratch)[0]);
  return ok();
}

int gf_sim_set_receivers(gf_sim *sp, const int32_t *cols, int32_t ncols, int32_t capacity_steps) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || ncols < 0 || capacity_steps < 0 || (ncols > 0 && !cols))
    return set_err(GF_ERR_ARGUMENT, "geofdtd: bad receiver arguments");
  for (int32_t i = 0; i < ncols; i++)
    if (cols[i] < 0 || cols[i] > s->R.nr - 2) return set_err(GF_ERR_ARGUMENT, "geofdtd: column out of range");
  int rc = s->set_receivers(cols, ncols, std::max<int32_t>(capacity_steps, 1), s->step);
  if (rc) return rc;
  if ((rc = s->reset_max_slots(s->step + std::max<int32_t>(capacity_steps, 1)))) return rc;
  return ok();
}

int gf_sim_advance(gf_sim *sp, int32_t n, int32_t use_graph) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s || n < 0) return set_err(GF_ERR_ARGUMENT, "geofdtd: bad advance arguments");
  // The divergence probe runs only when slots cover the steps ahead.
  const bool probe = s->g.maxbits && (s->step + n) / 256 + 1 < s->max_slots;
  int rc = s->advance(n, use_graph != 0, probe);
  return rc ? rc : ok();
}

int gf_sim_sync_timing(gf_sim *sp, double *kernel_ms, int64_t *kernel_launches) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  float ms = 0.f;
  if (s->timing_open) HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  if (kernel_ms) *kernel_ms = ms;
  if (kernel_launches) *kernel_launches = s->launches;
  s->timing_open = false;
  s->launches = 0;
  return ok();
}

int gf_sim_read_traces(gf_sim *sp, double *out, int32_t *steps_recorded) {
  Sim *s = reinterpret_cast<Sim *>(sp);
  if (!s) return set_err(GF_ERR_ARGUMENT, "geofdtd: null handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  const int64_t rec = std::max<int64_t>(0, std::min<int64_t>(s->step - s->g.rec_base, s->g.trace_cap));
  if (steps_recorded) *steps_recorded = (int32_t)rec;
  if (out && s->g.nrec > 0) {
    std::vector<float> tmp((size_t)s->g.trace_cap * s->g.nrec);
    HIPCHK(hipMemcpy(tmp.data(), s->d_traces, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
    for (int32_t r = 0; r < s->g.nrec; r++)
      for (int64_t k = 0; k < rec; k++) out[(size_t)r * rec + k] = tmp[(size_t)r * s->g.trace_cap + k];
  }
  return ok();
}

int gf_sim_step_bytes(const gf_sim *sp, double *bytes) {
  const Sim *s = reinterpret_cast<const Sim *>(sp);
  if (!s || !bytes) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *bytes = gfs::step_bytes(s->R);
  return ok();
}

int gf_run(const gf_shot *shot, gf_result *res) {
  if (!shot || !res || !shot->model) return set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  return gf_run_batch(shot, 1, res, nullptr);
}

}  // extern "C"

namespace gfs {
int api_set_err(int code, const std::string &msg) { return code ? set_err(code, msg) : ok(); }
}  // namespace gfs
