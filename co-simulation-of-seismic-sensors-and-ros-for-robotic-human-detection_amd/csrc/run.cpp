// run.cpp — fdtd.Run (internal/fdtd/run.go:88-127) for one shot or many.
//
// Many shots advance together: every step is one launch over the tiles of
// every model still running, so uneven grids share the GPU instead of
// queueing behind one another.  Models are ordered by step count, longest
// first, so the models still running are always a prefix of the tile list
// and the launch shrinks as short runs retire.  Blocks of kGraphSteps steps
// are replayed from captured HIP graphs.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>

#include <cmath>
#include <string.h>

#include <stdlib.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "kernels.h"
#include "model.h"
#include "batch.h"
#include "layout.h"
#include "march2.h"
#include "sim.h"
#include "tune.h"

namespace gfs {
int api_set_err(int code, const std::string &msg);
}

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

using gfs::Sim;
using gfb::Group;

namespace {

void free_group(Group &G) {
  if (G.exec) hipGraphExecDestroy(G.exec);
  if (G.graph) hipGraphDestroy(G.graph);
  if (G.d_descs) hipFree(G.d_descs);
  if (G.d_nbase) hipFree(G.d_nbase);
  if (G.d_tile_model) hipFree(G.d_tile_model);
  G.d_tile_model = nullptr;
  G.tile_model_cap = 0;
  if (G.d_tile_list) hipFree(G.d_tile_list);
  G.d_tile_list = nullptr;
  G.tile_list_cap = 0;
  G.in_prefix.clear();
  G.bd_prefix.clear();
  G.descs_cap = 0;
  G.exec = nullptr;
  G.graph = nullptr;
  G.d_descs = nullptr;
  G.d_nbase = nullptr;
}

// Models of the group still running at step n.
int active_models(const Group &G, const std::vector<int32_t> &steps, int64_t n) {
  int k = 0;
  while (k < (int)G.order.size() && steps[G.order[k]] > n) k++;
  return k;
}

// Descriptors of the group's models for tile height th, tile bases in run
// order, uploaded to the group's device array (allocated on first use).
// Destroys a group's graph only once the stream it was launched on has
// finished with it: an executable graph must not be destroyed while a launch
// of it may still be running.
int drop_graph(Group &G, hipStream_t st) {
  if (G.exec || G.graph) HIPCHK(hipStreamSynchronize(st));
  if (G.exec) hipGraphExecDestroy(G.exec);
  if (G.graph) hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
  return GF_OK;
}

int plan_group(Group &G, std::vector<std::unique_ptr<Sim>> &sims, int span, int th, hipStream_t st) {
  G.span = span;
  G.th = th;
  std::vector<gfk::GridDesc> descs;
  G.prefix.assign(1, 0);
  for (int i : G.order) {
    gfk::GridDesc d = sims[i]->g;
    gfl::set_tiling(d, span, th);
    d.tile_base = (int32_t)G.prefix.back();
    descs.push_back(d);
    G.prefix.push_back(G.prefix.back() + (int64_t)d.tiles_r * d.tiles_z);
  }
  if (G.prefix.back() > (int64_t)INT32_MAX) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: batch has too many tiles");
  if ((int64_t)descs.size() > G.descs_cap) {
    if (G.d_descs) (void)hipFree(G.d_descs);
    G.d_descs = nullptr;
    HIPCHK(hipMalloc(&G.d_descs, sizeof(gfk::GridDesc) * descs.size()));
    G.descs_cap = (int64_t)descs.size();
  }
  if (!G.d_nbase) HIPCHK(hipMalloc(&G.d_nbase, 64));
  HIPCHK(hipMemcpy(G.d_descs, descs.data(), sizeof(gfk::GridDesc) * descs.size(), hipMemcpyHostToDevice));
  {
    // Built for every plan, one model included: a group re-planned with
    // fewer models (the budget tuner's trial chunks) must not keep the table
    // of an earlier plan.
    if (descs.size() > 65535) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: too many models in one batch");
    // Padded to whole workgroups of 4 tiles: the launch's last workgroup
    // can hold up to three tiles past the end, which read the last model and
    // exit on its tile count.
    const int64_t tiles = (G.prefix.back() + 3) / 4 * 4;
    std::vector<uint16_t> tm((size_t)tiles, (uint16_t)(descs.size() - 1));
    for (size_t m = 0; m < descs.size(); m++)
      std::fill(tm.begin() + G.prefix[m], tm.begin() + G.prefix[m + 1], (uint16_t)m);
    if (tiles > G.tile_model_cap) {
      if (G.d_tile_model) (void)hipFree(G.d_tile_model);
      G.d_tile_model = nullptr;
      HIPCHK(hipMalloc(&G.d_tile_model, sizeof(uint16_t) * (size_t)tiles));
      G.tile_model_cap = tiles;
    }
    HIPCHK(hipMemcpy(G.d_tile_model, tm.data(), sizeof(uint16_t) * tm.size(), hipMemcpyHostToDevice));
  }
  G.in_prefix.clear();
  G.bd_prefix.clear();
  if (span == 2) {
    // the two-step kernel's interior tiles and the others, as separate lists
    std::vector<uint32_t> in, bd;
    G.in_prefix.assign(1, 0);
    G.bd_prefix.assign(1, 0);
    for (size_t m = 0; m < descs.size(); m++) {
      const gfk::GridDesc &d = descs[m];
      for (int lt = 0; lt < d.tiles_r * d.tiles_z; lt++) {
        const gfk::Tile2 t = gfk::step2_tile_at(d, lt);
        (gfk::step2_interior(d, t.i0, t.j0, t.j1) ? in : bd).push_back((uint32_t)(G.prefix[m] + lt));
      }
      G.in_prefix.push_back((int64_t)in.size());
      G.bd_prefix.push_back((int64_t)bd.size());
    }
    const int64_t n = (int64_t)(in.size() + bd.size());
    if (n > G.tile_list_cap) {
      if (G.d_tile_list) (void)hipFree(G.d_tile_list);
      G.d_tile_list = nullptr;
      HIPCHK(hipMalloc(&G.d_tile_list, sizeof(uint32_t) * (size_t)n));
      G.tile_list_cap = n;
    }
    in.insert(in.end(), bd.begin(), bd.end());
    HIPCHK(hipMemcpy(G.d_tile_list, in.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice));
  }
  int rc = drop_graph(G, st);
  if (rc) return rc;
  G.graph_models = -1;
  return GF_OK;
}

// A launch of the group's first `models` models.
gfk::StepLaunch group_launch(const Group &G, int models) {
  gfk::StepLaunch L;
  L.nmech = G.nmech;
  L.scalar = G.scalar;
  L.span = G.span;
  L.th = G.th;
  L.waves = G.waves;
  L.d = G.d_descs;
  L.tile_model = G.d_tile_model;
  L.tiles = G.prefix[models];
  L.count = G.span;
  if (G.span == 2 && !G.in_prefix.empty() && G.bd_prefix[models] >= gfs::split_min_boundary_tiles()) {
    L.list_in = G.d_tile_list;
    L.n_in = G.in_prefix[models];
    L.list_bd = G.d_tile_list + G.in_prefix.back();
    L.n_bd = G.bd_prefix[models];
  }
  return L;
}

// kGraphSteps steps (kGraphSteps / span launches, an even number) of the
// group's first `models` models; node k runs step nbase[0] + k * span from
// wavefield buffer nbase[1] ^ (k & 1) and memory-variable buffer nbase[2]
// (^ (k & 1) for two-step launches).
int capture(Group &G, int models, hipStream_t st) {
  int rc = drop_graph(G, st);
  if (rc) return rc;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < gfs::kGraphSteps / G.span; k++) {
    gfk::StepLaunch L = group_launch(G, models);
    L.n_off = k * G.span;
    L.rb_off = k & 1;
    L.mb_off = G.span == 2 ? (k & 1) : 0;
    L.nbase = G.d_nbase;
    L.probe = 1;
    hipError_t e = gfk::launch_step(L, st);
    if (e != hipSuccess) {
      hipGraph_t tmp = nullptr;
      hipStreamEndCapture(st, &tmp);
      if (tmp) hipGraphDestroy(tmp);
      return gfs::hip_fail(e, "graph capture");
    }
  }
  HIPCHK(hipStreamEndCapture(st, &G.graph));
  HIPCHK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
  G.graph_models = models;
  return GF_OK;
}

}  // namespace

namespace gfb {

int Batch::fail_shot(int i, int code, const std::string &m) {
  st[i] = code;
  msg[i] = m;
  if (!first_err) {
    first_err = code;
    first_msg = m;
  }
  sims[i].reset();
  return code;
}

// Per-shot set-up in the reference's order of checks (run.go:88-108):
// New's validation, the step count, then each receiver against MaxRange.
// `ranges[i]` (n_ranges[i] entries) are shot i's receiver ranges.
void Batch::setup(const gf_shot *shots, int32_t n, const std::vector<std::vector<double>> &ranges) {
  sims.resize(n);
  st.assign(n, GF_OK);
  steps.assign(n, 0);
  msg.assign(n, std::string());
  cols.assign(n, {});
  snapped.assign(n, {});
  for (int i = 0; i < n; i++) {
    const gf_shot &sh = shots[i];
    if (!sh.model || (sh.n_force > 0 && !sh.force)) {
      fail_shot(i, GF_ERR_ARGUMENT, "geofdtd: null argument in shot");
      continue;
    }
    sims[i].reset(new Sim());
    Sim &s = *sims[i];
    int rc = s.init(*sh.model, std::max(sh.steps, 0));
    if (rc) {
      fail_shot(i, rc, gf_last_error());
      continue;
    }
    if (sh.steps <= 0) {
      fail_shot(i, GF_ERR_INVALID, "fdtd: a shot needs a positive number of steps");
      continue;
    }
    if ((rc = s.set_force(sh.force, sh.n_force))) {
      fail_shot(i, rc, gf_last_error());
      continue;
    }
    bool bad = false;
    for (double r : ranges[i]) {
      int c = (int)round(r / s.R.h);
      c = std::max(0, std::min(c, s.R.nr - 2));
      const double sn = (double)c * s.R.h;
      if (sn >= sh.model->max_range_m) {
        fail_shot(i, GF_ERR_INVALID,
                  "fdtd: receiver at " + gfm::go_g(r) + " m is outside the interior (" +
                      gfm::go_g(sh.model->max_range_m) + " m)");
        bad = true;
        break;
      }
      cols[i].push_back(c);
      snapped[i].push_back(sn);
    }
    if (bad) continue;
    if ((rc = s.set_receivers(cols[i].data(), (int32_t)cols[i].size(), sh.steps, 0)) ||
        (rc = s.reset_max_slots(sh.steps))) {
      fail_shot(i, rc, gf_last_error());
      continue;
    }
    steps[i] = sh.steps;
  }
}

// State bytes of one model: every array a step reads (the other half of the
// double buffers is written): the figure that decides what stays in cache.
double state_bytes(const Sim &s) { return gfs::step_bytes(s.R) / 2; }

// Models of one kernel variant are advanced either all together or in chunks
// whose combined state fits the GPU's last-level cache (256 MB Infinity
// Cache), a chunk running its steps while its arrays are served from cache
// step after step.  Soils are independent, so the order changes nothing in any
// result.  A variant whose state exceeds kLockstepStateBytes is advanced all
// together by the two-step kernel, which moves the state once per two steps:
// on whole 1 s runs of 64 and 256 seeded layered soils (600-2048 x 300-1024
// cells, GSLS; 6.5 and 26 GB) that beat the best cache-sized chunks with
// either kernel by 7-15 %.  Smaller variants are
// chunked, the budget tuned (candidates 96-224 MB).  GEOFDTD_CACHE_MB fixes
// the budget, 0 advancing every model together.
constexpr double kLockstepStateBytes = 1e9;

Batch::~Batch() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (Chunk &C : chunks) free_group(C.G);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
}

bool Batch::finished() const {
  for (const Chunk &C : chunks)
    if (C.done < C.max_steps) return false;
  return true;
}

// Chunks: models of one kernel variant sorted by step count (the running set
// is a prefix), cut where the chunk budget would be exceeded; each variant's
// launch shape (tile height, tiles per workgroup) and chunk budget picked by
// the tuner on long runs; every chunk's tables uploaded.
int Batch::plan() {
  if (planned) return GF_OK;
  planned = true;
  const int n = (int)sims.size();
  std::vector<Group> variants;
  for (int i = 0; i < n; i++) {
    if (!sims[i]) continue;
    const Sim &s = *sims[i];
    const int nm = s.g.lossy ? s.g.nmech : 0;
    const bool sc = s.g.scalar != 0;
    auto it = std::find_if(variants.begin(), variants.end(),
                           [&](const Group &G) { return G.nmech == nm && G.scalar == sc; });
    if (it == variants.end()) {
      variants.emplace_back();
      variants.back().nmech = nm;
      variants.back().scalar = sc;
      it = variants.end() - 1;
    }
    it->order.push_back(i);
  }
  if (variants.empty()) return GF_OK;
  stream = sims[variants[0].order[0]]->stream;
  for (auto &sp : sims)
    if (sp) {
      hipError_t e = hipStreamSynchronize(sp->stream);
      if (e != hipSuccess) return gfs::hip_fail(e, "batch set-up");
    }
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return gfs::api_set_err(GF_ERR_DEVICE, "geofdtd: event set-up failed");

  const char *budget_env = getenv("GEOFDTD_CACHE_MB");
  const char *rows_env = getenv("GEOFDTD_TILE_ROWS");
  const char *tune_env = getenv("GEOFDTD_TUNE");
  int rc = GF_OK;
  for (size_t v = 0; v < variants.size(); v++) {
    Group &V = variants[v];
    std::stable_sort(V.order.begin(), V.order.end(), [&](int a, int b) { return steps[a] > steps[b]; });
    int64_t cells = 0, cell_steps = 0;
    for (int i : V.order) {
      cells += (int64_t)sims[i]->g.nr * sims[i]->g.nz;
      cell_steps += (int64_t)sims[i]->g.nr * sims[i]->g.nz * steps[i];
    }
    const bool want_tune = tune_env ? atoi(tune_env) != 0 : cell_steps >= 5000000000LL;
    const int64_t bucket = (int64_t)log2((double)std::max<int64_t>(1, cells));
    double total_state = 0;
    for (int i : V.order) total_state += state_bytes(*sims[i]);
    const bool lockstep = budget_env ? atof(budget_env) <= 0 : total_state > kLockstepStateBytes;
    // the models of a chunk under a budget (all of them for 0), cut from
    // position `from` of the longest-first order
    auto prefix_for = [&](double mb, size_t from = 0) {
      std::vector<int> out;
      double acc = 0;
      for (size_t p = from; p < V.order.size(); p++) {
        const int i = V.order[p];
        const double b = state_bytes(*sims[i]);
        if (!out.empty() && mb > 0 && acc + b > mb * 1e6) break;
        out.push_back(i);
        acc += b;
      }
      return out;
    };
    double mb = budget_env ? atof(budget_env) : lockstep ? 0.0 : 160.0;
    Group T;  // trial group: the first chunk
    T.nmech = V.nmech;
    T.scalar = V.scalar;
    T.order = prefix_for(mb);
    const int span0 = gfs::choose_span(lockstep ? 2 : 1);
    int64_t cols_rows = 0;
    for (int i : T.order)
      cols_rows += (int64_t)((sims[i]->g.nr + gfl::tile_cols(span0) - 1) / gfl::tile_cols(span0)) * sims[i]->g.nz;
    gft::Shape shape{gfs::choose_tile_rows_cells(span0, cols_rows), 4, span0};
    std::vector<int> touched;
    // Trial launches read the buffers the launch before wrote.  Each
    // candidate starts from cleared state (the trial state is discarded
    // afterwards), so its first launch may read either buffer.
    auto clear = [&](const std::vector<int> &models) {
      for (int i : models) {
        if (int e = sims[i]->reset_state()) return e;
        // the fill ran on the model's own stream; the trials run on the batch's
        hipError_t e = hipStreamSynchronize(sims[i]->stream);
        if (e != hipSuccess) return gfs::hip_fail(e, "trial reset");
      }
      return (int)GF_OK;
    };
    int tb = 0, tm = 0;
    auto trial_launch = [&](int k) {
      gfk::StepLaunch L = group_launch(T, (int)T.order.size());
      L.n_off = k * T.span;
      L.rb_off = tb;
      L.mb_off = tm;
      tb ^= 1;
      if (T.span == 2) tm ^= 1;
      return gfk::launch_step(L, stream);
    };
    auto trial_bytes = [&](const std::vector<int> &models, int span) {
      double bytes = 0;
      for (int i : models) bytes += gfs::step_bytes(sims[i]->R);
      return bytes * span;
    };
    if (want_tune && !rows_env) {
      // all together: the two-step kernel's shapes only (the one-step kernel
      // streams the whole state every step)
      gft::Key key{lockstep ? 2 : 0, V.nmech, V.scalar ? 1 : 0, (int)V.order.size(), bucket};
      std::vector<int> cands = gft::shape_candidates();
      if (lockstep)
        cands.erase(std::remove_if(cands.begin(), cands.end(), [](int c) { return gft::Shape::of(c).span != 2; }),
                    cands.end());
      auto prepare = [&](int code) {
        const gft::Shape sh = gft::Shape::of(code);
        T.waves = sh.waves;
        return plan_group(T, sims, sh.span, sh.th, stream) == GF_OK && clear(T.order) == GF_OK;
      };
      auto launch = [&](int, int k) { return trial_launch(k); };
      auto work = [&](int code) { return trial_bytes(T.order, gft::Shape::of(code).span); };
      int code = gft::cached(key);
      if (!code) code = gft::tune(key, stream, 8, cands, shape.code(), prepare, launch, work, nullptr);
      if (code) shape = gft::Shape::of(code);
      touched = T.order;
    }
    if (want_tune && !budget_env && !lockstep) {
      // chunk budget: the chunk each candidate cuts from the median model on
      // (the first chunk often holds a single large model whatever the
      // budget), timed at the chosen shape; the rate is its algorithmic
      // bytes per second
      gft::Key key{1, V.nmech, V.scalar ? 1 : 0, (int)V.order.size(), bucket};
      std::vector<int> cands(std::begin(gft::kBudgetsMB), std::end(gft::kBudgetsMB));
      const size_t mid = V.order.size() / 2;
      auto prepare = [&](int b) {
        T.order = prefix_for(b, mid);
        for (int i : T.order)
          if (std::find(touched.begin(), touched.end(), i) == touched.end()) touched.push_back(i);
        T.waves = shape.waves;
        return plan_group(T, sims, shape.span, shape.th, stream) == GF_OK && clear(T.order) == GF_OK;
      };
      auto launch = [&](int, int k) { return trial_launch(k); };
      auto work = [&](int b) { return trial_bytes(prefix_for(b, mid), shape.span); };
      int best = gft::cached(key);
      if (!best) best = gft::tune(key, stream, 48, cands, (int)mb, prepare, launch, work, nullptr);
      if (best) mb = best;
    }
    if ((rc = drop_graph(T, stream))) return rc;
    free_group(T);
    for (int i : touched)
      if ((rc = sims[i]->reset_state())) return rc;
    // cut the variant into chunks under the budget
    double acc = 0;
    bool first = true;
    for (int i : V.order) {
      const double b = state_bytes(*sims[i]);
      if (first || (mb > 0 && acc + b > mb * 1e6)) {
        chunks.emplace_back();
        chunks.back().variant = (int)v;
        chunks.back().G.nmech = V.nmech;
        chunks.back().G.scalar = V.scalar;
        chunks.back().G.waves = shape.waves;
        acc = 0;
        first = false;
      }
      chunks.back().G.order.push_back(i);
      acc += b;
    }
    for (Chunk &C : chunks) {
      if (C.variant != (int)v) continue;
      C.max_steps = steps[C.G.order.front()];
      if ((rc = plan_group(C.G, sims, shape.span, shape.th, stream))) return rc;
    }
    tile_rows = shape.th;
    waves = shape.waves;
    span = shape.span;
    budget_mb = mb;
  }
  for (auto &sp : sims)
    if (sp) {
      hipError_t e = hipStreamSynchronize(sp->stream);
      if (e != hipSuccess) return gfs::hip_fail(e, "batch set-up");
    }
  return GF_OK;
}

int Batch::advance(int64_t num, int64_t den) {
  if (den <= 0 || num < 0) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad batch fraction");
  int rc = plan();
  if (rc) return rc;
  for (Chunk &C : chunks) {
    Group &G = C.G;
    const int64_t target = num >= den ? C.max_steps : std::min(C.max_steps, (C.max_steps * num + den - 1) / den);
    if (target <= C.done) continue;
    // algorithmic bytes of the real steps this call runs
    for (int i : G.order) {
      const int64_t a = std::min<int64_t>(C.done, steps[i]), b = std::min<int64_t>(target, steps[i]);
      bytes += gfs::step_bytes(sims[i]->R) * (double)(b - a);
      // launches start at C.done and take G.span steps (the last maybe fewer)
      pass_bytes += gfs::step_bytes(sims[i]->R) * (double)((b - a + G.span - 1) / G.span);
    }
    if (hipEventRecord(e0, stream) != hipSuccess) return gfs::api_set_err(GF_ERR_DEVICE, "geofdtd: event record failed");
    int64_t k = C.done;
    while (k < target) {
      const int act = active_models(G, steps, k);
      if (act == 0) break;
      // A graph block may carry models past their last step (their tiles
      // return at once) but never past this call's target.
      const bool graph_block = k + gfs::kGraphSteps <= target;
      if (graph_block) {
        // Re-capture when the running set has shrunk by a tenth of its tiles.
        if (!G.exec || G.graph_models < act ||
            (G.prefix[act] < G.prefix[G.graph_models] - G.prefix[G.graph_models] / 10)) {
          if ((rc = capture(G, act, stream))) return rc;
        }
        hipError_t e = gfk::launch_set_base(G.d_nbase, (int32_t)k, C.buf, C.mbuf, stream);
        if (e == hipSuccess) e = hipGraphLaunch(G.exec, stream);
        if (e != hipSuccess) return gfs::hip_fail(e, "graph launch");
        // an even number of launches: the chunk's buffers are the same after
        k += gfs::kGraphSteps;
        launches += gfs::kGraphSteps / G.span;
      } else {
        gfk::StepLaunch L = group_launch(G, act);
        L.n_off = (int32_t)k;
        L.rb_off = C.buf;
        L.mb_off = C.mbuf;
        L.probe = 1;
        L.count = (int)std::min<int64_t>(G.span, target - k);
        hipError_t e = gfk::launch_step(L, stream);
        if (e != hipSuccess) return gfs::hip_fail(e, "step launch");
        C.buf ^= 1;
        if (G.span == 2) C.mbuf ^= 1;
        k += L.count;
        launches += 1;
      }
    }
    C.done = target;
    hipError_t e = hipEventRecord(e1, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) return gfs::hip_fail(e, "batch run");
    kernel_ms += ms;
    // a finished chunk's launch tables and graph are no longer needed
    if (C.done >= C.max_steps) free_group(G);
  }
  record_stats(*this);
  return GF_OK;
}

int Batch::reset() {
  int rc = plan();
  if (rc) return rc;
  for (const Chunk &C : chunks)
    for (int i : C.G.order)
      if (!sims[i]) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: a batch with a failed shot cannot be reset");
  for (Chunk &C : chunks) {
    if (C.done == 0) continue;
    if (C.done >= C.max_steps) {
      // its tables were released when it finished: plan it again
      if ((rc = plan_group(C.G, sims, C.G.span, C.G.th, stream))) return rc;
    }
    for (int i : C.G.order)
      if ((rc = sims[i]->reset_state()) || (rc = sims[i]->reset_max_slots(steps[i]))) return rc;
    C.done = 0;
    C.buf = 0;
    C.mbuf = 0;
  }
  for (auto &sp : sims)
    if (sp) {
      hipError_t e = hipStreamSynchronize(sp->stream);
      if (e != hipSuccess) return gfs::hip_fail(e, "batch reset");
    }
  kernel_ms = 0;
  launches = 0;
  bytes = 0;
  pass_bytes = 0;
  return GF_OK;
}

thread_local gf_batch_stats_t g_stats;

void record_stats(const Batch &b) {
  g_stats.kernel_ms = b.kernel_ms;
  g_stats.step_launches = b.launches;
  g_stats.algorithmic_bytes = b.bytes;
  g_stats.pass_bytes = b.pass_bytes;
  g_stats.tile_rows = b.tile_rows;
  g_stats.waves = b.waves;
  g_stats.steps_per_launch = b.span;
  g_stats.budget_mb = b.budget_mb;
  int models = 0;
  for (auto &s : b.sims) models += s ? 1 : 0;
  g_stats.models = models;
}

// The divergence probes of shot i (run.go:118-124): its peak, and whether a
// probe saw NaN or Inf (then the shot fails as the reference's Run does).
void Batch::probes(int i, double *peak_out) {
  Sim &s = *sims[i];
  const int64_t slots = (steps[i] - 1) / 256 + 1;
  std::vector<uint32_t> bits(slots);
  if (hipMemcpy(bits.data(), s.d_max, sizeof(uint32_t) * slots, hipMemcpyDeviceToHost) != hipSuccess) {
    fail_shot(i, GF_ERR_DEVICE, "geofdtd: reading divergence probes failed");
    return;
  }
  double peak = 0;
  for (int64_t k = 0; k < slots; k++) {
    const double p = gfs::bits_to_speed(bits[k]);
    peak = std::isnan(p) || std::isnan(peak) ? NAN : std::max(peak, p);
    if (std::isnan(p) || std::isinf(p)) {
      char buf[128];
      snprintf(buf, sizeof buf, "fdtd: diverged at step %lld of %d", (long long)(k * 256), steps[i]);
      if (peak_out) *peak_out = peak;
      fail_shot(i, GF_ERR_DIVERGED, buf);
      return;
    }
  }
  if (peak_out) *peak_out = peak;
}

}  // namespace gfb

extern "C" int gf_batch_stats(gf_batch_stats_t *out) {
  if (!out) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = gfb::g_stats;
  return gfs::api_set_err(GF_OK, "");
}

namespace gfb {

// Per-shot geometry of a set-up batch into the caller's results.
void describe_results(const Batch &B, gf_result *results) {
  for (size_t i = 0; i < B.sims.size(); i++) {
    if (!B.sims[i]) continue;
    results[i].dt = B.sims[i]->R.dt;
    results[i].t0 = 0.5 * B.sims[i]->R.dt;
    results[i].nr = B.sims[i]->R.nr;
    results[i].nz = B.sims[i]->R.nz;
    results[i].peak = 0;
    for (size_t r = 0; r < B.snapped[i].size(); r++)
      if (results[i].snapped) results[i].snapped[r] = B.snapped[i][r];
  }
}

// What fdtd.Run returns for each shot of a finished batch (run.go:104-127):
// the divergence probes, then the traces widened to float64.
int finish_results(Batch &B, gf_result *results, int32_t *statuses) {
  const int n = (int)B.sims.size();
  for (int i = 0; i < n; i++) {
    if (!B.sims[i]) continue;
    B.probes(i, &results[i].peak);
    if (!B.sims[i]) continue;
    Sim &s = *B.sims[i];
    const size_t nrec = B.cols[i].size();
    if (nrec > 0) {
      std::vector<float> tmp((size_t)B.steps[i] * nrec);
      if (hipMemcpy(tmp.data(), s.d_traces, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        B.fail_shot(i, GF_ERR_DEVICE, "geofdtd: reading traces failed");
        continue;
      }
      for (size_t k = 0; k < tmp.size(); k++) results[i].traces[k] = tmp[k];
    }
  }
  for (int i = 0; i < n; i++) {
    if (statuses) statuses[i] = B.st[i];
    results[i].status = B.st[i];
    snprintf(results[i].error, sizeof results[i].error, "%s", B.msg[i].c_str());
  }
  if (B.first_err) return gfs::api_set_err(B.first_err, B.first_msg);
  return gfs::api_set_err(GF_OK, "");
}

}  // namespace gfb

extern "C" int gf_run_batch(const gf_shot *shots, int32_t n_shots, gf_result *results, int32_t *statuses) {
  if (!shots || !results || n_shots <= 0) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad batch arguments");
  std::vector<std::vector<double>> ranges(n_shots);
  for (int i = 0; i < n_shots; i++) {
    if (shots[i].n_ranges > 0 && (!shots[i].ranges || !results[i].traces))
      return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null ranges or traces in shot");
    if (shots[i].n_ranges > 0) ranges[i].assign(shots[i].ranges, shots[i].ranges + shots[i].n_ranges);
  }
  gfb::Batch B;
  B.setup(shots, n_shots, ranges);
  gfb::describe_results(B, results);
  int rc = B.run();
  if (rc) return rc;
  return gfb::finish_results(B, results, statuses);
}
