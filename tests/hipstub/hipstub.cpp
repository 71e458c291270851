// hipstub.cpp — a recording stand-in for the HIP runtime and for the step
// library's kernel launchers, so the host side of the batched driver
// (csrc/run.cpp, csrc/api.cpp, csrc/tune.cpp) runs on the CPU under
// AddressSanitizer.  Test infrastructure only.
//
// Device memory is host heap memory, so ASan checks every copy and fill the
// host code asks for against the allocation it names, and flags a free of
// memory still in use.  On top of that the stub models what ASan cannot see:
//   * ordering: every stream keeps the operations enqueued on it that no
//     synchronization has retired yet, with the allocations each reads and
//     writes.  A synchronous hipMemcpy/hipMemset (the null stream, which does
//     not wait for non-blocking streams) or an operation on another stream
//     that touches an allocation a pending operation uses is a race;
//   * launches: a step launch (direct or replayed from a captured graph) is
//     checked against the descriptor array, the tile-to-model table and every
//     array each descriptor points at: live allocations of the sizes the
//     device layout gives them, the step inside the trace buffer and the
//     probe slot inside the probe array;
//   * capture: nothing but kernels may be recorded into a graph, and a graph
//     replays the launches with the descriptors as they are at replay time.
// Any violation prints "hipstub: ..." and aborts.
#include <hip/hip_runtime_api.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "kernels.h"
#include "layout.h"
#include "march2.h"

namespace {

[[noreturn]] void fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "hipstub: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  fflush(stderr);
  abort();
}

std::recursive_mutex g_mu;
std::map<uintptr_t, size_t> g_live;  // device allocations: base -> bytes
std::set<uintptr_t> g_host;          // pinned host allocations

struct Stats {
  long mallocs = 0, frees = 0, copies = 0, launches = 0, replays = 0, syncs = 0, races_checked = 0;
} g_stats;

// The allocation holding [p, p + n), or 0 when none does.
uintptr_t owner(const void *p, size_t n) {
  const uintptr_t a = (uintptr_t)p;
  auto it = g_live.upper_bound(a);
  if (it == g_live.begin()) return 0;
  --it;
  if (a >= it->first && a + n <= it->first + it->second) return it->first;
  return 0;
}

uintptr_t need(const void *p, size_t n, const char *what) {
  uintptr_t o = owner(p, n);
  if (!o) fail("%s: [%p, +%zu) is not inside a live device allocation", what, p, n);
  return o;
}

// ------------------------------------------------------------ stream model

struct Op {
  long seq;
  std::set<uintptr_t> reads, writes;
  const char *what;
};

}  // namespace

struct ihipStream_t {
  bool nonblocking = false;
  bool capturing = false;
  std::vector<Op> pending;
};
struct ihipEvent_t {
  hipStream_t stream = nullptr;
  long seq = 0;
  bool recorded = false;
};

namespace {

long g_seq = 0;

// A step launch as the kernel would see it.
struct Launch {
  int kind = 0;  // 0 step, 1 set_base
  int nmech = 0, th = 0, waves = 4, span = 1, count = 1;
  bool scalar = false;
  const gfk::GridDesc *d = nullptr;
  const uint16_t *tm = nullptr;
  int64_t tiles = 0;
  int32_t n_off = 0, rb_off = 0, mb_off = 0;
  const int32_t *nbase = nullptr;
  int32_t probe = 0;
  const uint32_t *list_in = nullptr, *list_bd = nullptr;  // two-step tiles split by kind
  int64_t n_in = 0, n_bd = 0;
  int32_t *base_ptr = nullptr;
  int32_t base_val = 0, base_buf = 0, base_mbuf = 0;
};

std::vector<ihipStream_t *> g_streams;
// per model (keyed by its first field plane, and by its first stress psi
// strip): the buffer its wavefields and its memory variables are in
std::map<uintptr_t, int> g_state_buf;
ihipStream_t g_null;  // the null stream (synchronous calls)

void retire_all() {
  for (auto *s : g_streams) s->pending.clear();
  g_stats.syncs++;
}

// A conflict between an operation about to run on stream `s` (or the null
// stream, s == nullptr) and anything still pending on another stream.
void check_race(ihipStream_t *s, const std::set<uintptr_t> &reads, const std::set<uintptr_t> &writes,
                const char *what) {
  g_stats.races_checked++;
  for (auto *o : g_streams) {
    if (o == s) continue;
    if (s == nullptr && !o->nonblocking) continue;  // the null stream waits for blocking streams
    for (const Op &op : o->pending) {
      for (uintptr_t w : writes)
        if (op.reads.count(w) || op.writes.count(w))
          fail("race: %s writes allocation %#lx while '%s' on another stream may still use it", what, (unsigned long)w,
               op.what);
      for (uintptr_t r : reads)
        if (op.writes.count(r))
          fail("race: %s reads allocation %#lx while '%s' on another stream may still write it", what,
               (unsigned long)r, op.what);
    }
  }
}

void enqueue(ihipStream_t *s, std::set<uintptr_t> reads, std::set<uintptr_t> writes, const char *what) {
  check_race(s, reads, writes, what);
  if (s) s->pending.push_back(Op{++g_seq, std::move(reads), std::move(writes), what});
}

// ------------------------------------------------------- launch validation

void check_span(const void *p, size_t floats, const char *what, std::set<uintptr_t> &rw) {
  rw.insert(need(p, floats * sizeof(float), what));
}

void check_desc(const gfk::GridDesc &g, int span, int th, int count, int64_t step, bool probe,
                std::set<uintptr_t> &reads, std::set<uintptr_t> &writes) {
  if (g.th != th) fail("descriptor planned for %d tile rows launched with %d", g.th, th);
  if (g.tiles_z != (g.nz + g.th - 1) / g.th) fail("descriptor tiles_z %d for nz %d at %d rows", g.tiles_z, g.nz, g.th);
  if (g.tiles_r != (g.nr + gfl::tile_cols(span) - 1) / gfl::tile_cols(span))
    fail("descriptor tiles_r %d for nr %d launched with the %d-step kernel", g.tiles_r, g.nr, span);
  // every column a halo lane reads stays inside its row and past nr
  const int64_t widest = std::max<int64_t>((int64_t)((g.nr + gfk::kOwn - 1) / gfk::kOwn) * gfk::kOwn,
                                           (int64_t)((g.nr + gfk::kOwn2 - 1) / gfk::kOwn2) * gfk::kOwn2 + 1);
  if (widest > g.pitch - 1 || g.nr > g.pitch - 2) fail("pitch %d too narrow for nr %d", g.pitch, g.nr);
  const size_t pitch = (size_t)g.pitch, nz = (size_t)g.nz;
  const size_t origin = gfk::kGuardCols + gfk::kGuardRows * pitch;
  const size_t field = (size_t)g.field_stride;
  if (field != gfk::kGuardCols + (nz + 2 * gfk::kGuardRows) * pitch + gfk::kGuardCols)
    fail("field stride %zu does not match the grid", field);
  for (int b = 0; b < 2; b++)
    for (int c = 0; c < 6; c++) check_span(g.f[b][c] - origin, field, "field plane", writes);
  const size_t rs = std::max<size_t>(1, nz * (size_t)g.rs_pitch);
  const size_t nzs = nz - std::min(nz, (size_t)std::max(0, g.z0));
  const size_t zs = std::max<size_t>(1, nzs * pitch);
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 2; k++) {
      check_span(g.vpsi_r[b][k], rs, "velocity psi_r strip", writes);
      check_span(g.vpsi_z[b][k], zs, "velocity psi_z strip", writes);
    }
  for (int b = 0; b < 2; b++)
    for (int k = 0; k < 2; k++) {
      check_span(g.spsi_r[b][k], rs, "stress psi_r strip", writes);
      check_span(g.spsi_z[b][k], zs, "stress psi_z strip", writes);
    }
  if (g.lossy)
    for (int b = 0; b < 2; b++) check_span(g.mem[b], (size_t)4 * g.nmech * nz * pitch, "memory variables", writes);
  check_span(g.rowc, (size_t)gfk::kRowW * nz, "row table", reads);
  check_span(g.colc, (size_t)gfk::kColN * pitch, "column table", reads);
  // A model whose run has ended may still be in a launch (a graph block
  // captured for a longer running set): its tiles return at once.
  if (step < 0) fail("negative step %lld", (long long)step);
  if (step < g.steps) {
    if (g.nsrc > 0) {
      check_span(g.src_tzz, g.nsrc, "source tzz", reads);
      check_span(g.src_dtdt, g.nsrc, "source rate", reads);
    }
    // the steps this launch takes for this model
    const int64_t last = std::min<int64_t>(step + count, g.steps);
    for (int64_t n = step; n < last; n++) {
      if (g.nrec > 0) {
        check_span(g.rec_col, g.nrec, "receiver columns", reads);
        const int64_t slot = n - g.rec_base;
        if (slot < 0 || slot >= g.trace_cap)
          fail("step %lld writes trace slot %lld of %lld", (long long)n, (long long)slot, (long long)g.trace_cap);
        check_span(g.traces, (size_t)g.nrec * g.trace_cap, "traces", writes);
      }
      if (probe && n % 256 == 0) check_span(g.maxbits + n / 256, 1, "probe slot", writes);
    }
  }
}

void run_launch(ihipStream_t *s, const Launch &L) {
  g_stats.launches++;
  std::set<uintptr_t> reads, writes;
  if (L.kind == 1) {
    writes.insert(need(L.base_ptr, 12, "graph base"));
    enqueue(s, reads, writes, "k_set_base");
    L.base_ptr[0] = L.base_val;
    L.base_ptr[1] = L.base_buf;
    L.base_ptr[2] = L.base_mbuf;
    return;
  }
  const uintptr_t dd = need(L.d, sizeof(gfk::GridDesc), "descriptor array");
  reads.insert(dd);
  const size_t ndesc = (g_live[dd] - ((uintptr_t)L.d - dd)) / sizeof(gfk::GridDesc);
  int32_t base = 0, bbuf = 0, mbuf = 0;
  if (L.nbase) {
    reads.insert(need(L.nbase, 12, "graph base"));
    base = L.nbase[0];
    bbuf = L.nbase[1];
    mbuf = L.nbase[2];
    if (bbuf != 0 && bbuf != 1) fail("graph base buffer %d", bbuf);
    if (mbuf != 0 && mbuf != 1) fail("graph base memory buffer %d", mbuf);
  }
  const int rb = (L.rb_off ^ bbuf) & 1, mb = (L.mb_off ^ mbuf) & 1;
  const int64_t padded = (L.tiles + L.waves - 1) / L.waves * L.waves;
  std::vector<int64_t> seen(ndesc, 0);
  if (L.tm) {
    reads.insert(need(L.tm, sizeof(uint16_t) * padded, "tile-to-model table"));
    for (int64_t t = 0; t < padded; t++) {
      const unsigned m = L.tm[t];
      if (m >= ndesc) fail("tile %lld maps to model %u of %zu descriptors", (long long)t, m, ndesc);
      const gfk::GridDesc &g = L.d[m];
      const int64_t own = (int64_t)g.tiles_r * g.tiles_z;
      if (t < L.tiles && (t < g.tile_base || t >= g.tile_base + own))
        fail("tile %lld maps to model %u whose tiles are [%d, %lld)", (long long)t, m, g.tile_base,
             (long long)(g.tile_base + own));
      if (t < L.tiles) seen[m]++;
    }
  } else {
    if (L.tiles > (int64_t)L.d[0].tiles_r * L.d[0].tiles_z) fail("launch of %lld tiles over one model", (long long)L.tiles);
    seen[0] = L.tiles;
  }
  if (L.list_in || L.list_bd) {
    // the interior and the other tiles of a two-step launch: both lists
    // live, every entry a tile of this launch of the right kind, together
    // every tile exactly once
    if (L.span != 2) fail("tile lists on a %d-step launch", L.span);
    if (L.n_in + L.n_bd != L.tiles)
      fail("tile lists of %lld + %lld entries for %lld tiles", (long long)L.n_in, (long long)L.n_bd, (long long)L.tiles);
    std::vector<char> hit((size_t)L.tiles, 0);
    for (int kind = 0; kind < 2; kind++) {
      const uint32_t *list = kind ? L.list_bd : L.list_in;
      const int64_t n = kind ? L.n_bd : L.n_in;
      if (n == 0) continue;
      reads.insert(need(list, sizeof(uint32_t) * (size_t)n, kind ? "boundary tile list" : "interior tile list"));
      for (int64_t e = 0; e < n; e++) {
        const int64_t t = list[e];
        if (t >= L.tiles) fail("tile list entry %lld is tile %lld of a launch of %lld", (long long)e, (long long)t,
                               (long long)L.tiles);
        if (hit[(size_t)t]++) fail("tile %lld listed twice", (long long)t);
        const gfk::GridDesc &g = L.d[L.tm ? L.tm[t] : 0];
        const gfk::Tile2 tt = gfk::step2_tile_at(g, (int)(t - g.tile_base));
        if (gfk::step2_interior(g, tt.i0, tt.j0, tt.j1) != (kind == 0))
          fail("tile %lld in the %s list", (long long)t, kind ? "boundary" : "interior");
      }
    }
  }
  for (size_t m = 0; m < ndesc; m++)
    if (seen[m]) {
      const gfk::GridDesc &g = L.d[m];
      check_desc(g, L.span, L.th, L.count, (int64_t)base + L.n_off, L.probe != 0, reads, writes);
      // the launch must read the buffer the model's last launch wrote (any,
      // after its state was cleared)
      if ((int64_t)base + L.n_off < g.steps) {
        auto it = g_state_buf.find((uintptr_t)g.f[0][0]);
        if (it != g_state_buf.end() && it->second != rb)
          fail("step %lld reads buffer %d but the model's wavefields are in buffer %d", (long long)(base + L.n_off), rb,
               it->second);
        g_state_buf[(uintptr_t)g.f[0][0]] = rb ^ 1;
        it = g_state_buf.find((uintptr_t)g.spsi_r[0][0]);
        if (it != g_state_buf.end() && it->second != mb)
          fail("step %lld reads memory buffer %d but the model's memory variables are in buffer %d",
               (long long)(base + L.n_off), mb, it->second);
        g_state_buf[(uintptr_t)g.spsi_r[0][0]] = L.span == 2 ? mb ^ 1 : mb;
      }
    }
  enqueue(s, reads, writes, "k_step");
}

}  // namespace

struct ihipGraph {
  std::vector<Launch> launches;
};
struct hipGraphExec {
  std::vector<Launch> launches;
};

namespace {
std::map<ihipStream_t *, ihipGraph *> g_capturing;

void submit(hipStream_t s, const Launch &L) {
  auto it = g_capturing.find(s);
  if (it != g_capturing.end()) {
    it->second->launches.push_back(L);
    return;
  }
  run_launch(s, L);
}

void no_capture(hipStream_t s, const char *what) {
  if (g_capturing.count(s)) fail("%s recorded into a captured graph", what);
}
}  // namespace

// ================================================================ HIP API

extern "C" {

hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_t *p, int) {
  memset(p, 0, sizeof *p);
  strcpy(p->name, "hipstub");
  strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
  p->multiProcessorCount = 256;
  return hipSuccess;
}
const char *hipGetErrorString(hipError_t) { return "hipstub error"; }
hipError_t hipGetLastError(void) { return hipSuccess; }
hipError_t hipDeviceSynchronize(void) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  retire_all();
  return hipSuccess;
}

hipError_t hipMalloc(void **p, size_t n) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const size_t bytes = std::max<size_t>(n, 1);
  void *q = aligned_alloc(256, (bytes + 255) / 256 * 256);
  if (!q) return hipErrorOutOfMemory;
  memset(q, 0xA5, bytes);  // device memory starts as garbage
  g_live[(uintptr_t)q] = bytes;
  g_stats.mallocs++;
  *p = q;
  return hipSuccess;
}

hipError_t hipFree(void *p) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!p) return hipSuccess;
  auto it = g_live.find((uintptr_t)p);
  if (it == g_live.end()) fail("hipFree(%p): not the base of a live allocation", p);
  retire_all();  // hipFree waits for the device
  for (auto jt = g_state_buf.begin(); jt != g_state_buf.end();)
    jt = (jt->first >= it->first && jt->first < it->first + it->second) ? g_state_buf.erase(jt) : std::next(jt);
  g_live.erase(it);
  g_stats.frees++;
  free(p);
  return hipSuccess;
}

hipError_t hipHostMalloc(void **p, size_t n, unsigned int) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  *p = malloc(std::max<size_t>(n, 1));
  g_host.insert((uintptr_t)*p);
  return hipSuccess;
}
hipError_t hipHostFree(void *p) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!p) return hipSuccess;
  if (!g_host.erase((uintptr_t)p)) fail("hipHostFree(%p): not a pinned allocation", p);
  retire_all();
  free(p);
  return hipSuccess;
}

static void copy_sets(void *dst, const void *src, size_t n, hipMemcpyKind kind, std::set<uintptr_t> &r,
                      std::set<uintptr_t> &w) {
  if (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) w.insert(need(dst, n, "copy destination"));
  if (kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) r.insert(need(src, n, "copy source"));
  if (kind == hipMemcpyDefault) fail("hipMemcpyDefault is not used by the library");
}

hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind kind) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  std::set<uintptr_t> r, w;
  copy_sets(dst, src, n, kind, r, w);
  check_race(nullptr, r, w, "hipMemcpy");
  memcpy(dst, src, n);
  g_stats.copies++;
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind kind, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(s, "hipMemcpyAsync");
  std::set<uintptr_t> r, w;
  copy_sets(dst, src, n, kind, r, w);
  enqueue(s, r, w, "hipMemcpyAsync");
  memcpy(dst, src, n);
  g_stats.copies++;
  return hipSuccess;
}

hipError_t hipMemset(void *dst, int v, size_t n) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  std::set<uintptr_t> r, w{need(dst, n, "hipMemset")};
  check_race(nullptr, r, w, "hipMemset");
  memset(dst, v, n);
  return hipSuccess;
}

hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(s, "hipMemsetAsync");
  enqueue(s, {}, {need(dst, n, "hipMemsetAsync")}, "hipMemsetAsync");
  memset(dst, v, n);
  return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int flags) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  *s = new ihipStream_t();
  (*s)->nonblocking = (flags & hipStreamNonBlocking) != 0;
  g_streams.push_back(*s);
  return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (g_capturing.count(s)) fail("stream destroyed while capturing");
  g_streams.erase(std::remove(g_streams.begin(), g_streams.end(), s), g_streams.end());
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (g_capturing.count(s)) fail("hipStreamSynchronize on a capturing stream");
  if (s) s->pending.clear();
  g_stats.syncs++;
  return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new ihipEvent_t();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(s, "hipEventRecord");
  e->stream = s;
  e->seq = g_seq;
  e->recorded = true;
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!e->recorded) return hipSuccess;
  if (e->stream) {
    auto &p = e->stream->pending;
    p.erase(std::remove_if(p.begin(), p.end(), [&](const Op &o) { return o.seq <= e->seq; }), p.end());
  }
  return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  if (!a->recorded || !b->recorded) return hipErrorInvalidHandle;
  *ms = 0.001f * (float)(1 + (b->seq - a->seq));
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (g_capturing.count(s)) fail("nested capture");
  g_capturing[s] = new ihipGraph();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  auto it = g_capturing.find(s);
  if (it == g_capturing.end()) fail("end of a capture that was not begun");
  *g = it->second;
  g_capturing.erase(it);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *x = new hipGraphExec();
  (*x)->launches = g->launches;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(s, "hipGraphLaunch");
  g_stats.replays++;
  for (const Launch &L : x->launches) run_launch(s, L);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete x;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}

// Totals for the driver's report.
void hipstub_stats(long out[7]) {
  out[0] = g_stats.mallocs;
  out[1] = g_stats.frees;
  out[2] = g_stats.copies;
  out[3] = g_stats.launches;
  out[4] = g_stats.replays;
  out[5] = g_stats.syncs;
  out[6] = (long)g_live.size();
}

}  // extern "C"

// ======================================================= kernel launchers

namespace gfk {

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "fill kernel");
  enqueue(st, {}, {need(p, sizeof(uint32_t) * n, "fill_u32")}, "k_fill_u32");
  std::fill(p, p + n, v);
  return hipSuccess;
}

hipError_t launch_fill_f32(float *p, int64_t n, float v, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "fill kernel");
  // a cleared state may be read from either buffer
  for (auto it = g_state_buf.begin(); it != g_state_buf.end();)
    it = (it->first >= (uintptr_t)p && it->first < (uintptr_t)(p + n)) ? g_state_buf.erase(it) : std::next(it);
  enqueue(st, {}, {need(p, sizeof(float) * n, "fill_f32")}, "k_fill_f32");
  std::fill(p, p + n, v);
  return hipSuccess;
}

hipError_t launch_set_base(int32_t *p, int32_t v, int32_t buf, int32_t mbuf, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (buf != 0 && buf != 1) fail("set_base with buffer %d", buf);
  if (mbuf != 0 && mbuf != 1) fail("set_base with memory buffer %d", mbuf);
  Launch L;
  L.kind = 1;
  L.base_ptr = p;
  L.base_val = v;
  L.base_buf = buf;
  L.base_mbuf = mbuf;
  submit(st, L);
  return hipSuccess;
}

hipError_t launch_build_tables(const TableParams &P, const LayerParams *d_layers, float *rowc, float *colc,
                               hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "table kernel");
  std::set<uintptr_t> r{need(d_layers, sizeof(LayerParams) * P.nlayers, "layer table")};
  if (kRowN + 2 * P.nmech_table > kRowW) fail("row table of %d mechanisms", P.nmech_table);
  std::set<uintptr_t> w{need(rowc, sizeof(float) * kRowW * P.nz, "row table"),
                        need(colc, sizeof(float) * kColN * P.pitch, "column table")};
  enqueue(st, r, w, "k_build_tables");
  return hipSuccess;
}

hipError_t launch_step(const StepLaunch &S, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (S.tiles <= 0) return hipSuccess;
  if (S.waves != 1 && S.waves != 2 && S.waves != 4) fail("launch with %d tiles per workgroup", S.waves);
  if (S.span != 1 && S.span != 2) fail("launch of the %d-step kernel", S.span);
  if (S.count < 1 || S.count > S.span) fail("launch of %d steps with the %d-step kernel", S.count, S.span);
  if (S.rb_off != 0 && S.rb_off != 1) fail("launch reading buffer %d", S.rb_off);
  if (S.mb_off != 0 && S.mb_off != 1) fail("launch reading memory buffer %d", S.mb_off);
  Launch L;
  L.nmech = S.nmech;
  L.scalar = S.scalar;
  L.span = S.span;
  L.count = S.count;
  L.th = S.th;
  L.waves = S.waves;
  L.d = S.d;
  L.tm = S.tile_model;
  L.tiles = S.tiles;
  L.n_off = S.n_off;
  L.rb_off = S.rb_off;
  L.mb_off = S.mb_off;
  L.nbase = S.nbase;
  L.probe = S.probe;
  L.list_in = S.list_in;
  L.list_bd = S.list_bd;
  L.n_in = S.n_in;
  L.n_bd = S.n_bd;
  submit(st, L);
  return hipSuccess;
}

hipError_t launch_dft_f32(const float *x, int64_t ld, int32_t rows, int32_t n, double, double, const double *freqs,
                          int32_t nbins, double *out, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "dft kernel");
  std::set<uintptr_t> r{need(x, sizeof(float) * ((size_t)ld * (rows - 1) + n), "dft input"),
                        need(freqs, sizeof(double) * nbins, "dft bins")};
  enqueue(st, r, {need(out, sizeof(double) * 2 * rows * nbins, "dft output")}, "k_dft");
  return hipSuccess;
}

hipError_t launch_dft_f64(const double *x, int64_t ld, int32_t rows, int32_t n, double, double, const double *freqs,
                          int32_t nbins, double *out, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "dft kernel");
  std::set<uintptr_t> r{need(x, sizeof(double) * ((size_t)ld * (rows - 1) + n), "dft input"),
                        need(freqs, sizeof(double) * nbins, "dft bins")};
  enqueue(st, r, {need(out, sizeof(double) * 2 * rows * nbins, "dft output")}, "k_dft");
  return hipSuccess;
}

hipError_t launch_surface_velocity(const GridDesc *d, int32_t, int32_t col, float *out, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "surface velocity kernel");
  if (col < 0 || col >= d->nr) fail("surface velocity column %d of %d", col, d->nr);
  enqueue(st, {need(d, sizeof(GridDesc), "descriptor")}, {need(out, sizeof(float) * 2, "surface velocity output")},
          "k_surface_velocity");
  return hipSuccess;
}

hipError_t launch_max_speed(const GridDesc *d, int32_t, uint32_t *slot, int64_t, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  no_capture(st, "max speed kernel");
  enqueue(st, {need(d, sizeof(GridDesc), "descriptor")}, {need(slot, sizeof(uint32_t), "max speed slot")},
          "k_max_speed");
  return hipSuccess;
}

}  // namespace gfk
