// surrogate_api.cpp — the Green's-function surrogate behind a gf_sur handle.
//
// Network (all GEMMs on the BF16 MFMA kernel of surrogate.hip, FP32
// accumulation, FP32 master weights and Adam state):
//
//   encoder  x[15] -> 512 -> 512 -> 512 -> 512 -> 512 -> 512, GELU after each
//   decoder  a dense projection 512 -> 2048, GELU, read as 64 positions x
//            32 channels (one position per 8 bins), then transposed
//            convolutions whose kernel equals their stride:
//              64 -> 128 positions, 32 -> 32 channels, GELU
//              128 -> 512,          32 -> 2 (re, im),  linear
//
// A transposed convolution with kernel = stride s maps each input position
// to s output positions independently, so over rows (query, position) it is
// the GEMM [rows][C_in] x [C_in][s*C_out]; with the output columns ordered
// (tap, channel), the result [rows][s][C_out] is already the next layer's
// [rows*s][C_out] input.  No layout pass runs between layers.
//
// Weights are stored [N][K] (out x in); a transposed convolution's bias is
// one per output channel (shared by its s taps).  Inputs: 15 features padded
// to 16 (the GEMM's K granularity) with a zero column, whose weights stay zero.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "sim.h"
#include "surrogate_kernels.h"

namespace gfs {
int api_set_err(int code, const std::string &msg);
}

#define HIPCHK(expr)                                       \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return gfs::hip_fail(e_, #expr); \
  } while (0)

namespace {

constexpr int kFeatures = 15;
constexpr int kInPad = 16;
constexpr int kOut = 1024;  // 512 bins x (re, im)
constexpr int kFusedLayers = 7;  // encoder (6) + the decoder's dense layer, fused at inference

struct Layer {
  int k, n, gelu, rows, bias_mod;  // rows: GEMM rows per query
  int64_t w_off, b_off;
};

std::vector<Layer> architecture() {
  std::vector<Layer> L;
  L.push_back({kInPad, 512, 1, 1, 512, 0, 0});
  for (int i = 0; i < 5; i++) L.push_back({512, 512, 1, 1, 512, 0, 0});
  L.push_back({512, 2048, 1, 1, 2048, 0, 0});
  L.push_back({32, 64, 1, 64, 32, 0, 0});
  L.push_back({32, 8, 0, 128, 2, 0, 0});
  int64_t off = 0;
  for (auto &l : L) {
    l.w_off = off;
    off += (int64_t)l.n * l.k;
    l.b_off = off;
    off += (l.bias_mod + 7) / 8 * 8;  // keep every weight block 16-byte aligned
  }
  return L;
}

int64_t param_count(const std::vector<Layer> &L) {
  const Layer &l = L.back();
  return l.b_off + (l.bias_mod + 7) / 8 * 8;
}

// Philox-free host draws for initialisation: splitmix64 into Box-Muller.
struct InitRng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((double)(next() >> 11) + 0.5) / 9007199254740992.0; }
  double normal() { return sqrt(-2.0 * log(uniform())) * cos(6.283185307179586 * uniform()); }
};

uint16_t host_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

}  // namespace

struct gf_sur {
  std::vector<Layer> L;
  int64_t P = 0;
  int max_batch = 0;
  double band_lo = 0, band_hi = 0;
  int polar = 0;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // the backward pass's second stream and its ordering events
  hipStream_t st2 = nullptr;
  // Gradient buffer sets the backward pass rotates through: the input
  // gradient of layer l writes the set that layer l + 2's weight gradient
  // read, so the second stream has two layers' time to finish with it.
  static constexpr int kGradSets = 3;
  hipEvent_t ev_ready = nullptr, ev_free[kGradSets] = {};
  // a training step's weight transposes run on st2 beside the next forward
  // pass; the next input gradient (the only reader of pbt) waits for ev_pbt
  hipEvent_t ev_adam = nullptr, ev_pbt = nullptr;
  bool pbt_pending = false;
  // parameters and optimizer state
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pb = nullptr;
  uint16_t *pbt = nullptr;  // each layer's BF16 weights transposed, [K][N]
  // the first kFusedLayers + 1 layers' BF16 weights in MFMA fragment order, for
  // the fused inference kernel; rebuilt after the weights change
  uint16_t *wpk = nullptr;
  int64_t wpk_off[8] = {};
  bool packed = false;
  int64_t adam_t = 0;
  // activations: X[l] is the BF16 input of layer l (X[L] the last GELU output
  // is not needed); Z[l] the FP32 pre-activation (the output for the last)
  std::vector<uint16_t *> X;
  std::vector<float *> Z;
  float *dF[kGradSets] = {};
  uint16_t *dB[kGradSets] = {};
  float *split = nullptr;
  int64_t split_floats = 0;
  float *colpart = nullptr;
  // row-tile partial column sums written by the input-gradient GEMMs (the
  // next layer's bias gradient), one buffer per gradient set
  float *bpart[kGradSets] = {};
  float *feat = nullptr, *tgt = nullptr, *wgt = nullptr, *loss = nullptr;
  double *scale = nullptr, *delay = nullptr;
  float *outp = nullptr;
  gf_sur_stats_t stats{};

  ~gf_sur() {
    hipStreamSynchronize(st);
    if (st2) hipStreamSynchronize(st2);
    for (void *q : {(void *)p, (void *)g, (void *)m, (void *)v, (void *)pb, (void *)pbt, (void *)wpk, (void *)split,
                    (void *)colpart, (void *)feat, (void *)tgt, (void *)wgt, (void *)loss, (void *)scale, (void *)delay,
                    (void *)outp})
      if (q) hipFree(q);
    for (int i = 0; i < kGradSets; i++)
      for (void *q : {(void *)dF[i], (void *)dB[i], (void *)bpart[i]})
        if (q) hipFree(q);
    for (auto q : X)
      if (q) hipFree(q);
    for (auto q : Z)
      if (q) hipFree(q);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    for (hipEvent_t e : {ev_ready, ev_adam, ev_pbt})
      if (e) hipEventDestroy(e);
    for (hipEvent_t e : ev_free)
      if (e) hipEventDestroy(e);
    if (st2) hipStreamDestroy(st2);
    if (st) hipStreamDestroy(st);
  }

  int init(const gf_sur_config &c) {
    L = architecture();
    P = param_count(L);
    max_batch = c.max_batch;
    band_lo = c.band_lo_hz;
    band_hi = c.band_hi_hz;
    polar = c.polar;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_adam, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_pbt, hipEventDisableTiming));
    for (hipEvent_t &e : ev_free) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(dalloc(&p, P));
    HIPCHK(dalloc(&g, P));
    HIPCHK(dalloc(&m, P));
    HIPCHK(dalloc(&v, P));
    HIPCHK(dalloc(&pb, P));
    HIPCHK(dalloc(&pbt, P));
    int64_t wo = 0;
    for (int l = 0; l <= kFusedLayers; l++) {
      wpk_off[l] = wo;
      wo += (int64_t)L[l].n * L[l].k;
    }
    HIPCHK(dalloc(&wpk, wo));
    const int64_t B = max_batch;
    X.assign(L.size(), nullptr);
    Z.assign(L.size(), nullptr);
    int64_t widest = 0;
    for (size_t l = 0; l < L.size(); l++) {
      HIPCHK(dalloc(&X[l], B * L[l].rows * L[l].k));
      HIPCHK(dalloc(&Z[l], B * L[l].rows * L[l].n));
      widest = std::max<int64_t>(widest, (int64_t)L[l].rows * std::max(L[l].n, L[l].k));
      split_floats = std::max<int64_t>(split_floats, (int64_t)splits_for(l, B) * L[l].n * L[l].k);
    }
    for (int i = 0; i < kGradSets; i++) {
      HIPCHK(dalloc(&dF[i], B * widest));
      HIPCHK(dalloc(&dB[i], B * widest));
    }
    HIPCHK(dalloc(&split, split_floats));
    int64_t colpart_floats = 1;
    for (const Layer &y : L)
      colpart_floats = std::max<int64_t>(colpart_floats, gfsk::bias_grad_part_floats(B * y.rows, y.n));
    HIPCHK(dalloc(&colpart, colpart_floats));
    int64_t bpart_floats = 1;
    for (size_t l = 1; l < L.size(); l++)
      if (bias_from_epilogue(l)) bpart_floats = std::max<int64_t>(bpart_floats, (int64_t)row_tiles_dx(l, B) * L[l].k);
    for (float *&b : bpart) HIPCHK(dalloc(&b, bpart_floats));
    HIPCHK(dalloc(&feat, B * kFeatures));
    HIPCHK(dalloc(&tgt, B * kOut));
    HIPCHK(dalloc(&wgt, B * kOut));
    HIPCHK(dalloc(&loss, B));
    HIPCHK(dalloc(&scale, B));
    HIPCHK(dalloc(&delay, B));
    HIPCHK(dalloc(&outp, B * kOut));
    return initialise(c.seed);
  }

  // K splits of a weight-gradient GEMM over `rows` = B * L[l].rows: enough
  // workgroups to cover the CUs, at least 256 rows per split.
  int splits_for(size_t l, int64_t B) const {
    const int64_t rows = B * L[l].rows;
    const int64_t tiles = (int64_t)((L[l].n + 127) / 128) * ((L[l].k + 127) / 128);
    int64_t s = std::max<int64_t>(1, 512 / std::max<int64_t>(tiles, 1));
    s = std::min<int64_t>(s, std::max<int64_t>(1, rows / 256));
    return (int)s;
  }

  int initialise(uint64_t seed) {
    std::vector<float> h(P, 0.0f);
    InitRng r{seed ^ 0x5EED5EEDull};
    for (size_t l = 0; l < L.size(); l++) {
      const Layer &y = L[l];
      const int fan_in = l == 0 ? kFeatures : y.k;
      const double sd = (l + 1 == L.size() ? 0.5 : sqrt(2.0)) / sqrt((double)fan_in);
      for (int n = 0; n < y.n; n++)
        for (int k = 0; k < (l == 0 ? kFeatures : y.k); k++) h[y.w_off + (int64_t)n * y.k + k] = (float)(sd * r.normal());
    }
    return upload(h.data());
  }

  int upload(const float *h) {
    std::vector<uint16_t> b(P);
    for (int64_t i = 0; i < P; i++) b[i] = host_bf16(h[i]);
    HIPCHK(hipMemcpyAsync(p, h, sizeof(float) * P, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(pb, b.data(), sizeof(uint16_t) * P, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(m, 0, sizeof(float) * P, st));
    HIPCHK(hipMemsetAsync(v, 0, sizeof(float) * P, st));
    int rc = transpose_weights();
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(st));
    adam_t = 0;
    return GF_OK;
  }

  // pbt[layer] = pb[layer]^T, the B operand of the input-gradient GEMMs.
  // beside = true (after an Adam step): on st2, after the update on st,
  // leaving ev_pbt for the next backward pass (or wait_pbt) to wait on.
  int transpose_weights(bool beside = false) {
    gfsk::TransposeTable t;
    if (L.size() > (size_t)gfsk::kMaxTransposeLayers) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: more layers than the transpose table holds");
    t.layers = (int32_t)L.size();
    for (size_t l = 0; l < L.size(); l++) {
      t.off[l] = L[l].w_off;
      t.n[l] = L[l].n;
      t.k[l] = L[l].k;
      t.first_tile[l + 1] = t.first_tile[l] + ((L[l].n + 15) / 16) * ((L[l].k + 15) / 16);
    }
    if (beside) {
      HIPCHK(hipEventRecord(ev_adam, st));
      HIPCHK(hipStreamWaitEvent(st2, ev_adam, 0));
      HIPCHK(gfsk::launch_transpose_layers(pb, pbt, t, st2));
      HIPCHK(hipEventRecord(ev_pbt, st2));
      pbt_pending = true;
    } else {
      HIPCHK(wait_pbt());
      HIPCHK(gfsk::launch_transpose_layers(pb, pbt, t, st));
    }
    packed = false;
    return GF_OK;
  }

  // st waits for transposes left running on st2.
  hipError_t wait_pbt() {
    if (!pbt_pending) return hipSuccess;
    pbt_pending = false;
    return hipStreamWaitEvent(st, ev_pbt, 0);
  }

  // Inference: the whole network in one launch (gfsk::launch_encoder_fused,
  // the transposed convolutions in the dense layer's epilogue).
  // GEOFDTD_SUR_UNFUSED=1 runs every layer as its own GEMM instead.
  static bool fused_enabled() {
    const char *e = getenv("GEOFDTD_SUR_UNFUSED");
    return !(e && atoi(e) != 0);
  }

  // feat != null: the kernel reads the FP32 features [B][kFeatures] there
  // and rounds them itself (no staging launch); else the padded BF16 X[0].
  // spec != null: de-normalize into spec (spec_rows rows) in the same launch.
  int forward_fused(int B, const float *feat = nullptr, float *spec = nullptr, int spec_rows = 0, double df = 0,
                    int k_lo = 0, int k_hi = 0) {
    if (!packed) {
      for (int l = 0; l <= kFusedLayers; l++)
        HIPCHK(gfsk::launch_pack_fragments(pb + L[l].w_off, L[l].n, L[l].k, wpk + wpk_off[l], st));
      packed = true;
    }
    gfsk::FusedArgs a;
    if (feat) {
      a.xf = feat;
      a.xf_cols = kFeatures;
    } else {
      a.x0 = X[0];
    }
    a.M = B;
    for (int l = 0; l <= kFusedLayers; l++) a.wp[l] = wpk + wpk_off[l];
    for (size_t l = 0; l < L.size(); l++) a.bias[l] = p + L[l].b_off;
    a.w8 = pb + L[8].w_off;
    if (spec) {
      a.spec = spec;
      a.spec_rows = spec_rows;
      a.scale = scale;
      a.delay = delay;
      a.df = df;
      a.k_lo = k_lo;
      a.k_hi = k_hi;
      a.polar = polar;
    } else {
      a.zout = Z.back();
    }
    HIPCHK(gfsk::launch_encoder_fused(a, st));
    return GF_OK;
  }

  // Forward over B queries whose padded BF16 features are in X[0].  keep:
  // store pre-activations for a backward pass.  Output: Z.back() [B][1024].
  int forward(int B, bool keep) {
    if (!keep && fused_enabled()) return forward_fused(B);
    return forward_from(0, B, keep);
  }

  // Layers first.. as one GEMM each.
  int forward_from(size_t first, int B, bool keep) {
    for (size_t l = first; l < L.size(); l++) {
      const Layer &y = L[l];
      gfsk::GemmArgs a;
      a.M = B * y.rows;
      a.N = y.n;
      a.K = y.k;
      a.A = X[l];
      a.lda = y.k;
      a.B = pb + y.w_off;
      a.ldb = y.k;
      a.bias = p + y.b_off;
      a.bias_mod = y.bias_mod;
      int epi;
      if (y.gelu) {
        epi = gfsk::kEpiBiasGelu;
        a.z = keep ? Z[l] : nullptr;
        a.ldz = y.n;
        a.cb = X[l + 1];
        a.ldcb = y.n;
      } else {
        epi = gfsk::kEpiBias;
        a.cf = Z[l];
        a.ldc = y.n;
      }
      HIPCHK(gfsk::launch_gemm(a, false, true, epi, 1, st));
    }
    return GF_OK;
  }

  // Layer l's input gradient (into layer l - 1) is formed in layer l's row
  // view, k = L[l].k columns.  When those columns are layer l - 1's bias
  // channels, the GEMM's epilogue also writes the bias gradient's column sums
  // by row tile, and no FP32 copy of the gradient is needed.
  bool bias_from_epilogue(size_t l) const { return L[l].k == L[l - 1].bias_mod; }

  gfsk::GemmArgs dx_args(size_t l, int64_t B) const {
    gfsk::GemmArgs x;
    x.M = (int32_t)(B * L[l].rows);
    x.N = L[l].k;
    x.K = L[l].n;
    return x;
  }

  int32_t row_tiles_dx(size_t l, int64_t B) const {
    return gfsk::gemm_row_tiles(dx_args(l, B), false, true, gfsk::kEpiGeluBwd);
  }

  // Backward from the output gradient in dF[0]/dB[0] ([B][1024]); weight and
  // bias gradients into g.  Layer l's weight and bias gradients run on st2
  // beside its input gradient on st: both read the gradient set `cur`, and
  // the input gradient writes the next set, which layer l + 2's gradients on
  // st2 (event ev_free[nxt]) must have finished reading.  On return st has
  // waited for all of st2's work.
  int backward(int B) {
    // on an error return, st2's work must not outlive the call (a later call
    // would reuse the buffers it reads)
    struct Drain {
      hipStream_t s;
      bool done = false;
      ~Drain() {
        if (!done) (void)hipStreamSynchronize(s);
      }
    } drain{st2};
    HIPCHK(wait_pbt());
    int cur = 0;
    bool pending[kGradSets] = {};
    int32_t part_rows[kGradSets] = {};  // row tiles of bias partials in bpart[i]; 0: use dF[i]
    HIPCHK(hipEventRecord(ev_ready, st));
    for (int l = (int)L.size() - 1; l >= 0; l--) {
      const Layer &y = L[l];
      const int64_t rows = (int64_t)B * y.rows;
      // st2 waits for the gradient set cur (ev_ready as recorded now, after
      // the input gradient that wrote it), before ev_ready is recorded again
      HIPCHK(hipStreamWaitEvent(st2, ev_ready, 0));
      const int nxt = (cur + 1) % kGradSets;
      if (l > 0) {
        // the input gradient first: it is the critical path, and enqueued
        // ahead of the weight-gradient launches it starts as soon as it can
        if (pending[nxt]) HIPCHK(hipStreamWaitEvent(st, ev_free[nxt], 0));
        // dX = dZ W, times GELU'(Z[l-1]) of the layer below
        gfsk::GemmArgs x = dx_args(l, B);
        x.A = dB[cur];
        x.lda = y.n;
        x.B = pbt + y.w_off;  // B(k=n_out, n=k_in) = W[n_out][k_in] = W^T[k_in][n_out]: k contiguous
        x.ldb = y.n;
        x.z = Z[l - 1];
        x.ldz = y.k;
        x.ldc = y.k;
        x.cb = dB[nxt];
        x.ldcb = y.k;
        if (bias_from_epilogue(l)) {
          x.colpart = bpart[nxt];
          part_rows[nxt] = gfsk::gemm_row_tiles(x, false, true, gfsk::kEpiGeluBwd);
        } else {
          x.cf = dF[nxt];
          part_rows[nxt] = 0;
        }
        HIPCHK(gfsk::launch_gemm(x, false, true, gfsk::kEpiGeluBwd, 1, st));
        HIPCHK(hipEventRecord(ev_ready, st));
      }
      // dW[n][k] = sum_rows dZ[row][n] X[row][k], split over rows
      const int s = splits_for(l, B);
      gfsk::GemmArgs w;
      w.M = y.n;
      w.N = y.k;
      w.K = (int32_t)rows;
      w.A = dB[cur];
      w.lda = y.n;
      w.B = X[l];
      w.ldb = y.k;
      w.cf = split;
      w.ldc = y.k;
      w.k_split = (int32_t)(((rows + s - 1) / s + 31) / 32 * 32);
      const int splits = (int)((rows + w.k_split - 1) / w.k_split);
      w.split_stride = (int64_t)y.n * y.k;
      HIPCHK(gfsk::launch_gemm(w, true, false, gfsk::kEpiF32, splits, st2));
      if (part_rows[cur] > 0) {
        // the weight gradient's split sums and the bias gradient's partial
        // sums, in one launch when both take one pass
        HIPCHK(gfsk::launch_sum_splits_pair(split, splits, (int64_t)y.n * y.k, g + y.w_off, bpart[cur],
                                            part_rows[cur], y.bias_mod, g + y.b_off, st2));
      } else {
        HIPCHK(gfsk::launch_sum_splits(split, splits, (int64_t)y.n * y.k, g + y.w_off, st2));
        HIPCHK(gfsk::launch_bias_grad(dF[cur], rows, y.n, y.bias_mod, colpart, g + y.b_off, st2));
      }
      HIPCHK(hipEventRecord(ev_free[cur], st2));
      pending[cur] = true;
      if (l == 0) break;
      cur = nxt;
    }
    HIPCHK(hipStreamWaitEvent(st, ev_free[cur], 0));
    drain.done = true;
    return GF_OK;
  }

  int stage_features(const float *dfeat, int B) {
    return gfsk::launch_pad_rows_bf16(dfeat, kFeatures, X[0], kInPad, B, st) == hipSuccess
               ? GF_OK
               : gfs::hip_fail(hipGetLastError(), "pad features");
  }
};

namespace {

int check_batch(const gf_sur *h, int64_t n) {
  if (!h) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null surrogate handle");
  if (n < 0) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: negative query count");
  return GF_OK;
}

double flops_per_query(const std::vector<Layer> &L) {
  double f = 0;
  for (size_t l = 0; l < L.size(); l++) f += 2.0 * L[l].rows * (l == 0 ? kFeatures : L[l].k) * L[l].n;
  return f;
}

}  // namespace

extern "C" {

int gf_sur_create(const gf_sur_config *c, gf_sur **out) {
  if (!c || !out) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = nullptr;
  // the decoder's last GEMM has 128 rows per query and one 128-row tile per
  // grid row: 16384 queries keep the grid inside its 65535 rows
  if (c->max_batch < 8 || c->max_batch % 8 != 0 || c->max_batch > (1 << 14))
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: surrogate max_batch must be a multiple of 8 in [8, 16384]");
  std::string why;
  int rc = gfs::check_device(why);
  if (rc) return gfs::api_set_err(rc, why);
  gf_sur *h = new gf_sur();
  if ((rc = h->init(*c))) {
    delete h;
    return rc;
  }
  h->stats.flops_per_query = flops_per_query(h->L);
  *out = h;
  return gfs::api_set_err(GF_OK, "");
}

void gf_sur_destroy(gf_sur *h) { delete h; }

int gf_sur_layers(const gf_sur *h, int32_t *n) {
  if (!h || !n) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *n = (int32_t)h->L.size();
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_layer(const gf_sur *h, int32_t l, gf_sur_layer_t *out) {
  if (!h || !out || l < 0 || l >= (int32_t)h->L.size())
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad layer index");
  const Layer &y = h->L[l];
  *out = gf_sur_layer_t{y.k, y.n, y.gelu, y.rows, y.bias_mod, y.w_off, y.b_off};
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_param_count(const gf_sur *h, int64_t *n) {
  if (!h || !n) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *n = h->P;
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_get_params(gf_sur *h, float *out) {
  if (!h || !out) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  HIPCHK(hipMemcpyAsync(out, h->p, sizeof(float) * h->P, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_set_params(gf_sur *h, const float *in) {
  if (!h || !in) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  int rc = h->upload(in);
  return rc ? rc : gfs::api_set_err(GF_OK, "");
}

int gf_sur_forward(gf_sur *h, const float *x, int32_t n, float *out) {
  int rc = check_batch(h, n);
  if (rc) return rc;
  if (n > 0 && (!x || !out)) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  double ms_total = 0;
  for (int32_t q0 = 0; q0 < n; q0 += h->max_batch) {
    const int B = std::min(h->max_batch, n - q0), Bp = (B + 7) / 8 * 8;
    HIPCHK(hipMemsetAsync(h->feat, 0, sizeof(float) * Bp * kFeatures, h->st));
    HIPCHK(hipMemcpyAsync(h->feat, x + (int64_t)q0 * kFeatures, sizeof(float) * B * kFeatures,
                          hipMemcpyHostToDevice, h->st));
    HIPCHK(hipEventRecord(h->e0, h->st));
    if (gf_sur::fused_enabled()) {
      if ((rc = h->forward_fused(Bp, h->feat))) return rc;
    } else if ((rc = h->stage_features(h->feat, Bp)) || (rc = h->forward(Bp, false))) {
      return rc;
    }
    HIPCHK(hipEventRecord(h->e1, h->st));
    HIPCHK(hipMemcpyAsync(out + (int64_t)q0 * kOut, h->Z.back(), sizeof(float) * B * kOut, hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
    ms_total += ms;
  }
  h->stats.forward_ms = ms_total;
  h->stats.forward_queries = n;
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_predict(gf_sur *h, const float *x, const double *scale, const double *delay, double bin_hz, int32_t n,
                   float *out) {
  int rc = check_batch(h, n);
  if (rc) return rc;
  if (n > 0 && (!x || !out || !scale || !delay)) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  double ms_total = 0;
  for (int32_t q0 = 0; q0 < n; q0 += h->max_batch) {
    const int B = std::min(h->max_batch, n - q0), Bp = (B + 7) / 8 * 8;
    HIPCHK(hipMemsetAsync(h->feat, 0, sizeof(float) * Bp * kFeatures, h->st));
    HIPCHK(hipMemcpyAsync(h->feat, x + (int64_t)q0 * kFeatures, sizeof(float) * B * kFeatures,
                          hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->scale, scale + q0, sizeof(double) * B, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->delay, delay + q0, sizeof(double) * B, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipEventRecord(h->e0, h->st));
    int k_lo = 0, k_hi = kOut / 2 - 1;
    if (h->band_hi > 0 && bin_hz > 0) {
      k_lo = (int)ceil(h->band_lo / bin_hz - 1e-9);
      k_hi = (int)floor(h->band_hi / bin_hz + 1e-9);
    }
    if (gf_sur::fused_enabled()) {
      // the features' rounding, the whole network and the de-normalization
      // in one launch
      if ((rc = h->forward_fused(Bp, h->feat, h->outp, B, bin_hz, k_lo, k_hi))) return rc;
    } else {
      if ((rc = h->stage_features(h->feat, Bp)) || (rc = h->forward(Bp, false))) return rc;
      HIPCHK(gfsk::launch_denormalize(h->Z.back(), kOut / 2, h->scale, h->delay, bin_hz, k_lo, k_hi, h->polar,
                                      h->outp, B, h->st));
    }
    HIPCHK(hipEventRecord(h->e1, h->st));
    HIPCHK(hipMemcpyAsync(out + (int64_t)q0 * kOut, h->outp, sizeof(float) * B * kOut, hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->e0, h->e1));
    ms_total += ms;
  }
  h->stats.forward_ms = ms_total;
  h->stats.forward_queries = n;
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_loss_grad(gf_sur *h, const float *x, const float *t, const float *w, int32_t n, double *loss, float *grad) {
  int rc = check_batch(h, n);
  if (rc) return rc;
  if (n <= 0 || n > h->max_batch || n % 8 != 0)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: loss_grad needs a multiple of 8 queries, at most max_batch");
  if (!x || !t || !loss) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  HIPCHK(hipMemcpyAsync(h->feat, x, sizeof(float) * n * kFeatures, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->tgt, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice, h->st));
  if (w) HIPCHK(hipMemcpyAsync(h->wgt, w, sizeof(float) * n * kOut, hipMemcpyHostToDevice, h->st));
  if ((rc = h->stage_features(h->feat, n)) || (rc = h->forward(n, true))) return rc;
  HIPCHK(gfsk::launch_rel_l2(h->Z.back(), h->tgt, w ? h->wgt : nullptr, n, kOut, 1.0f / n, h->loss, h->dF[0],
                             h->dB[0], h->st));
  if ((rc = h->backward(n))) return rc;
  std::vector<float> lb(n);
  HIPCHK(hipMemcpyAsync(lb.data(), h->loss, sizeof(float) * n, hipMemcpyDeviceToHost, h->st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, h->g, sizeof(float) * h->P, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  double s = 0;
  for (float v : lb) s += v;
  *loss = s / n;
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_train(gf_sur *h, const float *x, const float *t, const float *w, int64_t n, const gf_sur_train_opts *o,
                 float *losses) {
  int rc = check_batch(h, n);
  if (rc) return rc;
  if (!o || !x || !t) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  const int B = o->batch;
  if (B < 8 || B % 8 != 0 || B > h->max_batch || n < B || o->steps < 0)
    return gfs::api_set_err(GF_ERR_ARGUMENT,
                            "geofdtd: train batch must be a multiple of 8, at most max_batch and the sample count");
  float *dx = nullptr, *dt = nullptr, *dw = nullptr, *dls = nullptr;
  // each epoch's permutation, uploaded once into one of two buffers; the
  // event marks the last gather that read a buffer, before it is refilled two
  // epochs later
  int32_t *dperm[2] = {}, *hperm[2] = {};
  hipEvent_t ev_perm[2] = {};
  auto cleanup = [&]() {
    hipStreamSynchronize(h->st);
    // transposes a failed step left on st2 read pb, which a later call may
    // overwrite
    hipStreamSynchronize(h->st2);
    h->pbt_pending = false;
    if (dx) hipFree(dx);
    if (dt) hipFree(dt);
    if (dw) hipFree(dw);
    if (dls) hipFree(dls);
    for (int i = 0; i < 2; i++) {
      if (dperm[i]) hipFree(dperm[i]);
      if (hperm[i]) hipHostFree(hperm[i]);
      if (ev_perm[i]) hipEventDestroy(ev_perm[i]);
    }
  };
  auto fail = [&](hipError_t e, const char *w) {
    cleanup();
    return gfs::hip_fail(e, w);
  };
  hipError_t e;
  if ((e = dalloc(&dx, (size_t)n * kFeatures)) || (e = dalloc(&dt, (size_t)n * kOut)) ||
      (w && (e = dalloc(&dw, (size_t)n * kOut))) ||
      (e = dalloc(&dls, (size_t)std::max(o->steps, 1) * B)))
    return fail(e, "train allocation");
  for (int i = 0; i < 2; i++)
    if ((e = dalloc(&dperm[i], (size_t)n)) ||
        (e = hipHostMalloc((void **)&hperm[i], sizeof(int32_t) * n, hipHostMallocDefault)) ||
        (e = hipEventCreateWithFlags(&ev_perm[i], hipEventDisableTiming)))
      return fail(e, "train allocation");
  if ((e = hipMemcpyAsync(dx, x, sizeof(float) * n * kFeatures, hipMemcpyHostToDevice, h->st)) ||
      (e = hipMemcpyAsync(dt, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice, h->st)) ||
      (w && (e = hipMemcpyAsync(dw, w, sizeof(float) * n * kOut, hipMemcpyHostToDevice, h->st))))
    return fail(e, "train upload");
  // epochs of a seeded permutation; a batch never straddles two epochs
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  InitRng r{o->seed ^ 0xBA7C4ull};
  int64_t pos = n;
  float ms_total = 0;
  const double b1 = o->beta1 > 0 ? o->beta1 : 0.9, b2 = o->beta2 > 0 ? o->beta2 : 0.999;
  const double eps = o->eps > 0 ? o->eps : 1e-8;
  // The steps are enqueued without waiting for one another (the stream
  // orders them), so the host's launches run ahead of the device; the time
  // is taken over the whole run.
  int epoch = -1;
  if ((e = hipEventRecord(h->e0, h->st))) return fail(e, "event");
  for (int32_t s = 0; s < o->steps; s++) {
    if (pos + B > n) {
      for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int64_t)(r.uniform() * (double)(i + 1))]);
      pos = 0;
      epoch++;
      const int b = epoch & 1;
      if (epoch >= 2 && (e = hipEventSynchronize(ev_perm[b]))) return fail(e, "train sync");
      memcpy(hperm[b], perm.data(), sizeof(int32_t) * n);
      if ((e = hipMemcpyAsync(dperm[b], hperm[b], sizeof(int32_t) * n, hipMemcpyHostToDevice, h->st)))
        return fail(e, "train indices");
    }
    const int32_t *didx = dperm[epoch & 1] + pos;
    pos += B;
    if ((e = gfsk::launch_gather_rows(dx, didx, kFeatures, h->feat, B, h->st)) ||
        (e = gfsk::launch_gather_rows(dt, didx, kOut, h->tgt, B, h->st)) ||
        (w && (e = gfsk::launch_gather_rows(dw, didx, kOut, h->wgt, B, h->st))))
      return fail(e, "gather");
    if (pos + B > n && (e = hipEventRecord(ev_perm[epoch & 1], h->st))) return fail(e, "event");
    if ((rc = h->stage_features(h->feat, B)) || (rc = h->forward(B, true))) {
      cleanup();
      return rc;
    }
    if ((e = gfsk::launch_rel_l2(h->Z.back(), h->tgt, w ? h->wgt : nullptr, B, kOut, 1.0f / B, dls + (int64_t)s * B,
                                 h->dF[0], h->dB[0], h->st)))
      return fail(e, "loss");
    if ((rc = h->backward(B))) {
      cleanup();
      return rc;
    }
    // cosine decay from lr to lr_min over the run
    const double frac = o->steps > 1 ? (double)s / (o->steps - 1) : 0.0;
    const double lr = o->lr_min + 0.5 * (o->lr - o->lr_min) * (1 + cos(3.141592653589793 * frac));
    h->adam_t++;
    const double c1 = 1 - pow(b1, (double)h->adam_t), c2 = 1 - pow(b2, (double)h->adam_t);
    if ((e = gfsk::launch_adam(h->p, h->g, h->m, h->v, h->pb, h->P, (float)lr, (float)b1, (float)b2, (float)eps,
                               (float)c1, (float)c2, h->st)))
      return fail(e, "adam");
    if ((rc = h->transpose_weights(true))) {
      cleanup();
      return rc;
    }
  }
  if ((e = h->wait_pbt())) return fail(e, "transpose");
  if ((e = hipEventRecord(h->e1, h->st))) return fail(e, "event");
  if ((e = hipEventSynchronize(h->e1))) return fail(e, "train run");
  hipEventElapsedTime(&ms_total, h->e0, h->e1);
  if (losses && o->steps > 0) {
    std::vector<float> all((size_t)o->steps * B);
    if ((e = hipMemcpy(all.data(), dls, sizeof(float) * all.size(), hipMemcpyDeviceToHost))) return fail(e, "losses");
    for (int32_t s = 0; s < o->steps; s++) {
      double acc = 0;
      for (int b = 0; b < B; b++) acc += all[(size_t)s * B + b];
      losses[s] = (float)(acc / B);
    }
  }
  cleanup();
  h->stats.train_ms = ms_total;
  h->stats.train_steps = o->steps;
  h->stats.train_queries = (int64_t)o->steps * B;
  return gfs::api_set_err(GF_OK, "");
}

int gf_sur_stats(const gf_sur *h, gf_sur_stats_t *out) {
  if (!h || !out) return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: null argument");
  *out = h->stats;
  return gfs::api_set_err(GF_OK, "");
}

int gf_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t *A, int64_t lda, int32_t a_trans, const uint16_t *B,
                 int64_t ldb, int32_t b_nk, int32_t splits, float *C) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || splits < 1)
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: bad gemm arguments");
  if (K % 8 || lda % 8 || ldb % 8 || (a_trans && M % 8) || (!b_nk && N % 8))
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: gemm sizes break the 8-element vector rule");
  if ((!a_trans && lda < K) || (a_trans && lda < M) || (b_nk && ldb < K) || (!b_nk && ldb < N))
    return gfs::api_set_err(GF_ERR_ARGUMENT, "geofdtd: gemm leading dimension too small");
  std::string why;
  int rc = gfs::check_device(why);
  if (rc) return gfs::api_set_err(rc, why);
  const size_t na = a_trans ? (size_t)K * lda : (size_t)M * lda;
  const size_t nb = b_nk ? (size_t)N * ldb : (size_t)K * ldb;
  gfsk::GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = N;
  g.k_split = (((K + splits - 1) / splits) + 31) / 32 * 32;
  const int s = (K + g.k_split - 1) / g.k_split;
  g.split_stride = (int64_t)M * N;
  uint16_t *dA = nullptr, *dBm = nullptr;
  float *dP = nullptr, *dC = nullptr;
  hipError_t e;
  if (!(e = dalloc(&dA, na)) && !(e = dalloc(&dBm, nb)) && !(e = dalloc(&dP, (size_t)s * M * N)) &&
      !(e = dalloc(&dC, (size_t)M * N)) && !(e = hipMemcpy(dA, A, na * 2, hipMemcpyHostToDevice)) &&
      !(e = hipMemcpy(dBm, B, nb * 2, hipMemcpyHostToDevice))) {
    g.A = dA;
    g.B = dBm;
    g.cf = dP;
    if (!(e = gfsk::launch_gemm(g, a_trans != 0, b_nk != 0, gfsk::kEpiF32, s, nullptr)) &&
        !(e = gfsk::launch_sum_splits(dP, s, (int64_t)M * N, dC, nullptr)) && !(e = hipDeviceSynchronize()))
      e = hipMemcpy(C, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost);
  }
  hipFree(dA);
  hipFree(dBm);
  hipFree(dP);
  hipFree(dC);
  if (e) return gfs::hip_fail(e, "gf_gemm_bf16");
  return gfs::api_set_err(GF_OK, "");
}

}  // extern "C"
