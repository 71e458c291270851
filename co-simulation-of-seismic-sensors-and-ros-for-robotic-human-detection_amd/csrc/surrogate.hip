// surrogate.hip — the Green's-function surrogate's kernels for gfx950.
//
// The network is a chain of GEMMs (an MLP encoder and a transposed-convolution
// decoder whose kernel equals its stride, which is a GEMM over (query,
// position) rows).  Every GEMM, forward and backward, is one kernel:
//
//   C[M][N] = sum_k A(m,k) B(k,n),   A, B BF16,  accumulation FP32 on MFMA
//
// with A stored either [M][K] or [K][M] and B either [N][K] or [K][N], so the
// forward product X W^T, the input gradient dZ W and the weight gradient
// dZ^T X all run on the same tiles.  Tiles are staged in LDS k-contiguous (the
// MFMA operand layout), so a lane fetches its 8 BF16 of one row with one
// 16-byte LDS read.  The 32x32x16 BF16 MFMA gives lane l of a 32x32 output
// tile column l%32 and rows 8(i/4) + 4(l/32) + i%4 for its 16 accumulators.
//
// Epilogues are fused: bias + GELU (keeping the FP32 pre-activation for the
// backward pass), bias only (the output layer), GELU' times the incoming
// gradient (the backward pass of the layer below), or plain FP32 (weight
// gradients, one slice per split of K).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>

#include "sur_math.h"
#include "surrogate_kernels.h"

namespace gfsk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// K per LDS stage is a template parameter BK (a multiple of 16: BK/16 MFMA
// steps per stage); LDS rows are padded to BK + 8 elements.


// One ROWS x kBK tile, moved in two halves so that the global loads of the
// next stage are in flight while the MFMAs of this one run: fetch_tile reads
// it into registers (16-byte vectors), store_tile writes them to LDS as
// dst[row][k].  TRANS = false: element (r, k) at src[r*ld + k] (k
// contiguous; vectors copied as they are).  TRANS = true: element (r, k) at
// src[k*ld + r] (r contiguous; vectors along r, scattered into 8 LDS rows).
// Rows at or beyond R and k at or beyond kend read as zero; the host
// guarantees that a vector is wholly in or out (K, and R for TRANS, multiples
// of 8, and 16-byte aligned rows).
template <int ROWS, int BK>
struct TileRegs {
  static constexpr int kVecs = ROWS * BK / 8;
  static constexpr int kPer = (kVecs + 255) / 256;
  uint4 v[kPer];
};

template <int ROWS, bool TRANS, int BK>
__device__ __forceinline__ void fetch_tile(TileRegs<ROWS, BK> &t, const uint16_t *__restrict__ src, int64_t ld,
                                           int r0, int k0, int R, int kend) {
  constexpr int kVecs = TileRegs<ROWS, BK>::kVecs;
#pragma unroll
  for (int i = 0; i < TileRegs<ROWS, BK>::kPer; i++) {
    const int v = threadIdx.x + 256 * i;
    uint4 val = make_uint4(0, 0, 0, 0);
    if (v < kVecs) {
      if (!TRANS) {
        const int row = v / (BK / 8), kv = v % (BK / 8);
        const int gr = r0 + row, gk = k0 + kv * 8;
        if (gr < R && gk < kend) val = *(const uint4 *)(src + (int64_t)gr * ld + gk);
      } else {
        const int k = v / (ROWS / 8), rv = v % (ROWS / 8);
        const int gk = k0 + k, gr = r0 + rv * 8;
        if (gk < kend && gr < R) val = *(const uint4 *)(src + (int64_t)gk * ld + gr);
      }
    }
    t.v[i] = val;
  }
}

template <int ROWS, bool TRANS, int BK>
__device__ __forceinline__ void store_tile(const TileRegs<ROWS, BK> &t, uint16_t *dst) {
  constexpr int kVecs = TileRegs<ROWS, BK>::kVecs, LDK = BK + 8;
#pragma unroll
  for (int i = 0; i < TileRegs<ROWS, BK>::kPer; i++) {
    const int v = threadIdx.x + 256 * i;
    if (v >= kVecs) break;
    if (!TRANS) {
      const int row = v / (BK / 8), kv = v % (BK / 8);
      *(uint4 *)(dst + row * LDK + kv * 8) = t.v[i];
    } else {
      const int k = v / (ROWS / 8), rv = v % (ROWS / 8);
      const uint16_t *e = (const uint16_t *)&t.v[i];
#pragma unroll
      for (int j = 0; j < 8; j++) dst[(rv * 8 + j) * LDK + k] = e[j];
    }
  }
}

// Four wavefronts: WM x WN of them tile the block's output, and the
// KW = 4 / (WM WN) groups of those take turns over a stage's 16-deep k-steps
// (k-step s goes to group s % KW); the groups' accumulators are added in LDS
// before the epilogue, in group order.  KW > 1 is for short, skinny outputs
// with a long K (weight gradients of the small layers).
template <int BM, int BN, int BK, int WM, int WN, bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  constexpr int TM = BM / WM / 32, TN = BN / WN / 32, kLDK = BK + 8;
  constexpr int KW = 4 / (WM * WN);
  static_assert(4 % (WM * WN) == 0 && TM >= 1 && TN >= 1, "4 wavefronts, 32x32 MFMA tiles");
  static_assert(BK % (16 * KW) == 0, "whole MFMA steps per stage and k group");
  static_assert(KW == 1 || (KW - 1) * WM * WN * TM * TN * 16 * 64 * 4 <= BM * kLDK * 2, "k-group sums fit in As");
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * kLDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * kLDK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wmn = wave % (WM * WN), wk = wave / (WM * WN);
  const int wm = wmn / WN, wn = wmn % WN;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * g.k_split;
  const int kend = min(g.K, kbeg + g.k_split);

  f32x16 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++)
#pragma unroll
      for (int i = 0; i < 16; i++) acc[a][b][i] = 0.0f;

  // software pipeline: the next stage's global loads are issued before this
  // stage's MFMAs, and land in LDS after them
  TileRegs<BM, BK> ra;
  TileRegs<BN, BK> rb;
  if (kbeg < kend) {
    fetch_tile<BM, AT, BK>(ra, g.A, g.lda, m0, kbeg, g.M, kend);
    fetch_tile<BN, !BT, BK>(rb, g.B, g.ldb, n0, kbeg, g.N, kend);
  }
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    store_tile<BM, AT, BK>(ra, As);
    store_tile<BN, !BT, BK>(rb, Bs);
    __syncthreads();
    if (k0 + BK < kend) {
      fetch_tile<BM, AT, BK>(ra, g.A, g.lda, m0, k0 + BK, g.M, kend);
      fetch_tile<BN, !BT, BK>(rb, g.B, g.ldb, n0, k0 + BK, g.N, kend);
    }
#pragma unroll
    for (int s = 0; s < BK / 16; s++) {
      if (KW > 1 && s % KW != wk) continue;
      bf16x8 fa[TM], fb[TN];
      const int kk = s * 16 + (lane >> 5) * 8;
#pragma unroll
      for (int a = 0; a < TM; a++)
        fa[a] = *(const bf16x8 *)(As + (wm * TM * 32 + a * 32 + (lane & 31)) * kLDK + kk);
#pragma unroll
      for (int b = 0; b < TN; b++)
        fb[b] = *(const bf16x8 *)(Bs + (wn * TN * 32 + b * 32 + (lane & 31)) * kLDK + kk);
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }

  if (KW > 1) {
    // the loop ended on a barrier, so As is free: groups 1.. park their sums,
    // one slot per (group, output position wmn, tile, element, lane)
    float *red = (float *)As;
    if (wk > 0)
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++)
#pragma unroll
          for (int i = 0; i < 16; i++)
            red[((((wk - 1) * WM * WN + wmn) * TM * TN + a * TN + b) * 16 + i) * 64 + lane] = acc[a][b][i];
    __syncthreads();
    if (wk > 0) return;
#pragma unroll
    for (int w = 0; w < KW - 1; w++)
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++)
#pragma unroll
          for (int i = 0; i < 16; i++)
            acc[a][b][i] += red[(((w * WM * WN + wmn) * TM * TN + a * TN + b) * 16 + i) * 64 + lane];
  }

  float csum[TN];
#pragma unroll
  for (int b = 0; b < TN; b++) csum[b] = 0.0f;
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) {
      const int col = n0 + wn * TN * 32 + b * 32 + (lane & 31);
      if (col >= g.N) continue;
      const float bias = (EPI == kEpiBiasGelu || EPI == kEpiBias) ? g.bias[col % g.bias_mod] : 0.0f;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = m0 + wm * TM * 32 + a * 32 + (i / 4) * 8 + (lane >> 5) * 4 + (i % 4);
        if (row >= g.M) continue;
        const float v = acc[a][b][i];
        if (EPI == kEpiF32) {
          g.cf[(int64_t)blockIdx.z * g.split_stride + (int64_t)row * g.ldc + col] = v;
        } else if (EPI == kEpiBias) {
          g.cf[(int64_t)row * g.ldc + col] = v + bias;
        } else if (EPI == kEpiBiasGelu) {
          const float z = v + bias;
          if (g.z) g.z[(int64_t)row * g.ldz + col] = z;
          g.cb[(int64_t)row * g.ldcb + col] = to_bf16(gelu(z));
        } else {  // kEpiGeluBwd
          const float d = v * gelu_grad(g.z[(int64_t)row * g.ldz + col]);
          if (g.cf) g.cf[(int64_t)row * g.ldc + col] = d;
          g.cb[(int64_t)row * g.ldcb + col] = to_bf16(d);
          csum[b] += d;
        }
      }
    }

  if constexpr (EPI == kEpiGeluBwd) {
    // column sums of d over the block's rows (the next layer's bias
    // gradient, in parts): lanes l and l + 32 hold the same column, then the
    // WM wavefronts sharing a column add in LDS in wavefront order
    static_assert(KW == 1 && WM * BN * 4 <= BM * kLDK * 2, "column sums fit in As");
    if (g.colpart) {
      float *red = (float *)As;  // free: the main loop ended on a barrier
#pragma unroll
      for (int b = 0; b < TN; b++) {
        const float t = csum[b] + __shfl_xor(csum[b], 32);
        if (lane < 32) red[wm * BN + wn * TN * 32 + b * 32 + lane] = t;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < BN; c += 256) {
        if (n0 + c >= g.N) continue;
        float t = 0.0f;
        for (int w = 0; w < WM; w++) t += red[w * BN + c];
        g.colpart[(int64_t)blockIdx.y * g.N + n0 + c] = t;
      }
    }
  }
}

// Tile shapes (M x N x K-stage, wavefronts along M x N).  The default picks by
// N and by how many workgroups a shape gives; GEOFDTD_GEMM_SHAPE forces one of
// the shapes for N > 32 (a tuning aid: every shape is correct for every size).
template <int BM, int BN, int BK, int WM, int WN, bool AT, bool BT, int EPI>
hipError_t launch_tiles(const GemmArgs &g, int splits, hipStream_t st) {
  const dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, splits);
  hipLaunchKernelGGL((k_gemm<BM, BN, BK, WM, WN, AT, BT, EPI>), grid, dim3(256), 0, st, g);
  return hipGetLastError();
}

inline int forced_shape() {
  static const int v = [] {
    const char *e = getenv("GEOFDTD_GEMM_SHAPE");
    return e ? atoi(e) : -1;
  }();
  return v;
}

// The tile shape for a product: one id, used by the launcher and by
// gemm_row_tiles so that both always agree.
enum Shape { kSkinny32, kSkinny64, kNarrow, kS128x128, kS128x64, kS128x128k128, kS128x64k128, kS64x64, kS64x64k128 };

static Shape pick_shape(const GemmArgs &g, bool at, bool bt, int epi) {
  // weight gradients of the small layers: an output of at most 64 x 32 over
  // a long K, so the wavefronts share the K range in 256-deep stages
  if (at && !bt && epi == kEpiF32 && g.N <= 32 && g.M <= 32) return kSkinny32;
  if (at && !bt && epi == kEpiF32 && g.N <= 32 && g.M <= 64) return kSkinny64;
  if (g.N <= 32) return kNarrow;
  switch (forced_shape()) {
    case 0: return kS128x128;
    case 1: return kS128x64;
    case 2: return kS128x128k128;
    case 3: return kS128x64k128;
    case 4: return kS64x64;
    case 5: return kS64x64k128;
    default: break;
  }
  // 64 x 64 tiles, four wavefronts of 32 x 32: the network's GEMMs are short
  // in K (16-2048) and latency-bound, and small tiles let a CU interleave up
  // to eight workgroups (measured best of the shapes above for inference at
  // batch 8192 and for training at 4096)
  return kS64x64;
}

static int shape_rows(Shape s) {
  switch (s) {
    case kSkinny32: return 32;
    case kSkinny64: case kS64x64: case kS64x64k128: return 64;
    default: return 128;
  }
}

template <bool AT, bool BT, int EPI>
hipError_t launch_shape(const GemmArgs &g, int splits, hipStream_t st) {
  switch (pick_shape(g, AT, BT, EPI)) {
    case kSkinny32:
    case kSkinny64:
      if constexpr (AT && !BT && EPI == kEpiF32) {
        if (g.M <= 32) return launch_tiles<32, 32, 256, 1, 1, AT, BT, EPI>(g, splits, st);
        return launch_tiles<64, 32, 256, 2, 1, AT, BT, EPI>(g, splits, st);
      }
      return hipErrorInvalidValue;
    case kNarrow: return launch_tiles<128, 32, 64, 4, 1, AT, BT, EPI>(g, splits, st);
    case kS128x128: return launch_tiles<128, 128, 64, 2, 2, AT, BT, EPI>(g, splits, st);
    case kS128x64: return launch_tiles<128, 64, 64, 4, 1, AT, BT, EPI>(g, splits, st);
    case kS128x128k128: return launch_tiles<128, 128, 128, 2, 2, AT, BT, EPI>(g, splits, st);
    case kS128x64k128: return launch_tiles<128, 64, 128, 4, 1, AT, BT, EPI>(g, splits, st);
    case kS64x64: return launch_tiles<64, 64, 64, 2, 2, AT, BT, EPI>(g, splits, st);
    case kS64x64k128: return launch_tiles<64, 64, 128, 2, 2, AT, BT, EPI>(g, splits, st);
  }
  return hipErrorInvalidValue;
}

template <int EPI>
hipError_t launch_layouts(const GemmArgs &g, bool at, bool bt, int splits, hipStream_t st) {
  if (!at && bt) return launch_shape<false, true, EPI>(g, splits, st);
  if (!at && !bt) return launch_shape<false, false, EPI>(g, splits, st);
  if (at && !bt) return launch_shape<true, false, EPI>(g, splits, st);
  return launch_shape<true, true, EPI>(g, splits, st);
}

int32_t gemm_row_tiles(const GemmArgs &g, bool at, bool bt, int epi) {
  const int r = shape_rows(pick_shape(g, at, bt, epi));
  return (g.M + r - 1) / r;
}

hipError_t launch_gemm(const GemmArgs &g, bool at, bool bt, int epi, int splits, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if ((g.M + 127) / 128 > 65535 || splits < 1 || splits > 65535) return hipErrorInvalidConfiguration;
  switch (epi) {
    case kEpiF32: return launch_layouts<kEpiF32>(g, at, bt, splits, st);
    case kEpiBias: return launch_layouts<kEpiBias>(g, at, bt, splits, st);
    case kEpiBiasGelu: return launch_layouts<kEpiBiasGelu>(g, at, bt, splits, st);
    case kEpiGeluBwd: return launch_layouts<kEpiGeluBwd>(g, at, bt, splits, st);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------ elementwise

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// Rows of `in` ([rows][cols_in] FP32) into the first cols_in columns of
// `out` ([rows][cols_out] BF16), zero in the rest.
__global__ void k_pad_rows_bf16(const float *__restrict__ in, int32_t cols_in, uint16_t *__restrict__ out,
                                int32_t cols_out, int64_t rows) {
  const int64_t n = rows * cols_out;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols_out;
    const int c = (int)(i % cols_out);
    out[i] = c < cols_in ? to_bf16(in[r * cols_in + c]) : (uint16_t)0;
  }
}

// Rows `idx[b]` of src ([*][cols] FP32) into dst row b.
__global__ void k_gather_rows(const float *__restrict__ src, const int32_t *__restrict__ idx, int32_t cols,
                              float *__restrict__ dst, int32_t rows) {
  const int64_t n = (int64_t)rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols;
    const int c = (int)(i % cols);
    dst[i] = src[(int64_t)idx[r] * cols + c];
  }
}

// The loss, one workgroup per sample (scale = 1/B gives the gradient of the
// batch mean; dF and dB receive the gradient in FP32 and BF16):
//   w == null: relative spectral L2,  loss_b = |o_b - t_b|^2 / |t_b|^2,
//              d o_b = scale * 2 (o_b - t_b) / |t_b|^2
//   w != null: weighted squares,      loss_b = sum_i w_i (o_i - t_i)^2,
//              d o_i = scale * 2 w_i (o_i - t_i)
// (with outputs (log-amplitude, phase) and w = |T|^2 / sum |T|^2 per bin, the
// weighted form is the relative L2 to first order in the error).
__global__ __launch_bounds__(256) void k_rel_l2(const float *__restrict__ o, const float *__restrict__ t,
                                                const float *__restrict__ w, int32_t len, float scale,
                                                float *__restrict__ loss, float *__restrict__ dF,
                                                uint16_t *__restrict__ dB) {
  const int b = blockIdx.x;
  const float *ob = o + (int64_t)b * len, *tb = t + (int64_t)b * len;
  const float *wb = w ? w + (int64_t)b * len : nullptr;
  float num = 0, den = 0;
  for (int i = threadIdx.x; i < len; i += 256) {
    const float d = ob[i] - tb[i];
    num += wb ? wb[i] * d * d : d * d;
    den += wb ? 0.0f : tb[i] * tb[i];
  }
  __shared__ float sn[256], sd[256];
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sn[threadIdx.x] += sn[threadIdx.x + off];
      sd[threadIdx.x] += sd[threadIdx.x + off];
    }
    __syncthreads();
  }
  const float inv = wb ? 1.0f : (sd[0] > 0 ? 1.0f / sd[0] : 0.0f);
  if (threadIdx.x == 0) loss[b] = sn[0] * inv;
  if (dF) {
    for (int i = threadIdx.x; i < len; i += 256) {
      const float g = scale * 2.0f * (ob[i] - tb[i]) * (wb ? wb[i] : inv);
      dF[(int64_t)b * len + i] = g;
      dB[(int64_t)b * len + i] = to_bf16(g);
    }
  }
}

// Column sums of x[rows][cols] over chunks of rows: part[chunk][col].  One
// workgroup per (256-column stripe, chunk): its threads read 16-byte vectors,
// tpr = min(cols, 256) / 4 of them across a row and 256 / tpr rows at once,
// so every load is a coalesced full row segment; the row phases are then
// added in an LDS tree.  cols is a multiple of 4 that divides 256, or a
// multiple of 256.  The chunk length keeps about 1024 workgroups in flight.
__host__ __device__ inline int64_t colsum_chunk_rows(int64_t rows, int32_t cols) {
  const int64_t stripes = cols > 256 ? cols / 256 : 1;
  int64_t c = 8;
  while (c < 4096 && (rows + c - 1) / c * stripes > 1024) c *= 2;
  return c;
}

__global__ __launch_bounds__(256) void k_colsum_part(const float *__restrict__ x, int64_t rows, int32_t cols,
                                                     int64_t chunk, float *__restrict__ part) {
  const int cw = cols < 256 ? cols : 256, tpr = cw / 4, phases = 256 / tpr;
  const int t = threadIdx.x, c4 = t % tpr, ph = t / tpr;
  const int col = blockIdx.x * 256 + c4 * 4;
  const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t r = r0 + ph; r < r1; r += phases) {
    const float4 v = *(const float4 *)(x + r * cols + col);
    s.x += v.x;
    s.y += v.y;
    s.z += v.z;
    s.w += v.w;
  }
  __shared__ float4 red[256];
  red[t] = s;
  __syncthreads();
  for (int h = phases / 2; h > 0; h /= 2) {
    if (ph < h) {
      const float4 o = red[t + h * tpr];
      s.x += o.x;
      s.y += o.y;
      s.z += o.z;
      s.w += o.w;
      red[t] = s;
    }
    __syncthreads();
  }
  if (ph == 0) *(float4 *)(part + (int64_t)blockIdx.y * cols + col) = s;
}

// out[j] = sum over chunks and over columns c with c % mod == j of part: one
// workgroup per channel j, its lanes striding over the (chunk, tap) entries,
// then an LDS tree.
__global__ __launch_bounds__(256) void k_colsum_fold(const float *__restrict__ part, int32_t chunks, int32_t cols,
                                                     int32_t mod, float *__restrict__ out) {
  const int j = blockIdx.x;
  const int taps = cols / mod;
  float s = 0;
  for (int e = threadIdx.x; e < chunks * taps; e += 256) s += part[(int64_t)(e / taps) * cols + j + (e % taps) * mod];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// The K splits of a weight gradient, out[i] = sum_s part[s][i], in two
// passes.  Pass 1: the splits fall into groups of gs consecutive slices;
// workgroup (x, y) adds group y for outputs 64 x .. 64 x + 63 (its four
// wavefronts take every fourth slice, then add in LDS) and writes the group's
// sum over the group's first slice, which no other workgroup reads -- or to
// out when there is one group.  Pass 2 adds the groups' sums.
// The sum of slices s0 .. s1 - 1 for outputs 64 bx .. 64 bx + 63, returned
// by wavefront 0 (the other wavefronts' lanes return 0).
__device__ __forceinline__ float sum_slices(const float *__restrict__ part, int s0, int s1, int64_t n, int64_t bx,
                                            int lane, int wave, bool &valid) {
  const int64_t i = bx * 64 + lane;
  // four running sums per lane (slices k, k + 4, k + 8, k + 12 of the
  // wavefront's stride-4 sequence), so that four loads are in flight
  float a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  if (i < n) {
    int k = s0 + wave;
    for (; k + 12 < s1; k += 16) {
      a0 += part[(int64_t)k * n + i];
      a1 += part[(int64_t)(k + 4) * n + i];
      a2 += part[(int64_t)(k + 8) * n + i];
      a3 += part[(int64_t)(k + 12) * n + i];
    }
    for (; k < s1; k += 4) a0 += part[(int64_t)k * n + i];
  }
  const float s = (a0 + a1) + (a2 + a3);
  __shared__ float red[4][64];
  red[wave][lane] = s;
  __syncthreads();
  valid = wave == 0 && i < n;
  return valid ? (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]) : 0.0f;
}

__global__ __launch_bounds__(256) void k_sum_groups(float *__restrict__ part, int32_t splits, int32_t gs, int64_t n,
                                                    float *__restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s0 = blockIdx.y * gs, s1 = min(splits, s0 + gs);
  bool valid;
  const float v = sum_slices(part, s0, s1, n, blockIdx.x, lane, wave, valid);
  if (valid) {
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    if (gridDim.y == 1)
      out[i] = v;
    else
      part[(int64_t)s0 * n + i] = v;
  }
}

// Two one-pass sums in one launch (the weight gradient's split-K slices and
// the bias gradient's row-tile partials): workgroups below `blocks_a` do the
// first, the rest the second, each as k_sum_groups with one group does it.
__global__ __launch_bounds__(256) void k_sum_pair(const float *__restrict__ pa, int32_t sa, int64_t na,
                                                  float *__restrict__ oa, int64_t blocks_a,
                                                  const float *__restrict__ pb, int32_t sb, int64_t nb,
                                                  float *__restrict__ ob) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool first = (int64_t)blockIdx.x < blocks_a;
  const int64_t bx = first ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - blocks_a;
  bool valid;
  const float v = first ? sum_slices(pa, 0, sa, na, bx, lane, wave, valid) : sum_slices(pb, 0, sb, nb, bx, lane, wave, valid);
  if (valid) (first ? oa : ob)[bx * 64 + lane] = v;
}

__global__ void k_sum_strided(const float *__restrict__ part, int32_t terms, int64_t stride, int64_t n,
                              float *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float s = 0;
    for (int k = 0; k < terms; k++) s += part[(int64_t)k * stride + i];
    out[i] = s;
  }
}

// Adam (Kingma & Ba 2015) on FP32 master parameters, with the BF16 copy the
// GEMMs read refreshed in the same pass.  c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ pb, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    pb[i] = to_bf16(pi);
  }
}

// Normalized network output to the physical spectrum, per query b and bin k:
//   G = s_b T e^{-i 2 pi f_k w_b},  f_k = k * df,
// T = o_0 + i o_1 (cartesian) or e^{o_0 + i o_1} (polar: log-amplitude and
// phase), s_b the amplitude scale and w_b the delay the targets were divided
// by; zero outside bins [k_lo, k_hi].
// One bin's de-normalization; products are not contracted, so every kernel
// that calls it gives the same bits.  The phase -2 pi k df delay (+ the
// network's phase, polar) is formed and reduced to [-pi, pi] in FP64 (it
// reaches ~10^3 rad); the sine, cosine, exponential and products are FP32
// (about 1e-7 relative).
__device__ __forceinline__ void denorm_bin(float o0, float o1, double scale, double delay, double df, int k, int k_lo,
                                           int k_hi, int polar, float &re_out, float &im_out) {
#pragma clang fp contract(off)
  if (k < k_lo || k > k_hi) {
    re_out = 0.0f;
    im_out = 0.0f;
    return;
  }
  constexpr double kTwoPi = 6.283185307179586;
  double ph = -kTwoPi * (double)k * df * delay;
  float re = o0, im = o1;
  if (polar) {
    ph += (double)o1;
    re = expf(o0);
    im = 0.0f;
  }
  const float phr = (float)(ph - kTwoPi * rint(ph * (1.0 / kTwoPi)));
  float sn, cs;
  sincosf(phr, &sn, &cs);
  const float sc = (float)scale;
  re_out = sc * (re * cs - im * sn);
  im_out = sc * (re * sn + im * cs);
}

__global__ void k_denormalize(const float *__restrict__ o, int32_t bins, const double *__restrict__ scale,
                              const double *__restrict__ delay, double df, int32_t k_lo, int32_t k_hi,
                              int32_t polar, float *__restrict__ out, int32_t rows) {
  const int64_t n = (int64_t)rows * bins;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / bins;
    const int k = (int)(i % bins);
    denorm_bin(o[2 * i], o[2 * i + 1], scale[b], delay[b], df, k, k_lo, k_hi, polar, out[2 * i], out[2 * i + 1]);
  }
}

// out[k][n] = in[n][k] for one layer's BF16 weights (16x16 tiles via LDS).
// Every layer's [n][k] weight block transposed to [k][n] at the same offset,
// in one launch: 16 x 16 tiles numbered layer by layer.
__global__ __launch_bounds__(256) void k_transpose_layers(const uint16_t *__restrict__ in, uint16_t *__restrict__ out,
                                                          TransposeTable tt) {
  __shared__ uint16_t t[16][17];
  int l = 0;
  while (l + 1 < tt.layers && (int)blockIdx.x >= tt.first_tile[l + 1]) l++;
  const int tile = blockIdx.x - tt.first_tile[l], tiles_k = (tt.k[l] + 15) / 16;
  const int n_rows = tt.n[l], k_cols = tt.k[l];
  const uint16_t *src = in + tt.off[l];
  uint16_t *dst = out + tt.off[l];
  const int n0 = tile / tiles_k * 16, k0 = tile % tiles_k * 16;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  if (n0 + ty < n_rows && k0 + tx < k_cols) t[ty][tx] = src[(int64_t)(n0 + ty) * k_cols + k0 + tx];
  __syncthreads();
  if (k0 + ty < k_cols && n0 + tx < n_rows) dst[(int64_t)(k0 + ty) * n_rows + n0 + tx] = t[tx][ty];
}

namespace {
inline dim3 grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return dim3((unsigned)(b < 4096 ? b : 4096));
}
}  // namespace

hipError_t launch_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, grid_for(n), dim3(256), 0, st, x, y, n);
  return hipGetLastError();
}

hipError_t launch_pad_rows_bf16(const float *in, int32_t cols_in, uint16_t *out, int32_t cols_out, int64_t rows,
                                hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pad_rows_bf16, grid_for(rows * cols_out), dim3(256), 0, st, in, cols_in, out, cols_out, rows);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const float *src, const int32_t *idx, int32_t cols, float *dst, int32_t rows,
                              hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows, grid_for((int64_t)rows * cols), dim3(256), 0, st, src, idx, cols, dst, rows);
  return hipGetLastError();
}

hipError_t launch_rel_l2(const float *o, const float *t, const float *w, int32_t rows, int32_t len, float scale,
                         float *loss, float *dF, uint16_t *dB, hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rel_l2, dim3(rows), dim3(256), 0, st, o, t, w, len, scale, loss, dF, dB);
  return hipGetLastError();
}

int64_t bias_grad_part_floats(int64_t max_rows, int32_t cols) {
  // at most 1024 chunks while the chunk length is below its cap, and
  // ceil(rows / cap) beyond it
  return std::max<int64_t>(1024, (max_rows + 4095) / 4096) * cols;
}

hipError_t launch_bias_grad(const float *x, int64_t rows, int32_t cols, int32_t mod, float *part, float *out,
                            hipStream_t st) {
  if (cols % 4 || (cols <= 256 && 256 % cols) || (cols > 256 && cols % 256) || mod < 1 || cols % mod)
    return hipErrorInvalidValue;
  if (rows <= 0) return hipSuccess;
  const int64_t c = colsum_chunk_rows(rows, cols);
  const int64_t chunks = (rows + c - 1) / c;
  if (chunks > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_colsum_part, dim3(cols > 256 ? cols / 256 : 1, (unsigned)chunks), dim3(256), 0, st, x, rows,
                     cols, c, part);
  hipLaunchKernelGGL(k_colsum_fold, dim3(mod), dim3(256), 0, st, part, (int32_t)chunks, cols, mod, out);
  return hipGetLastError();
}

hipError_t launch_sum_splits(float *part, int32_t splits, int64_t n, float *out, hipStream_t st) {
  if (n <= 0 || splits < 1) return hipSuccess;
  const int64_t colblocks = (n + 63) / 64;
  if (colblocks > 0x7fffffff) return hipErrorInvalidConfiguration;
  // one pass for up to 64 slices; else about 1024 workgroups in pass 1, at
  // most 16 groups (pass 2's terms), and at least 8 slices a group
  int64_t groups = std::max<int64_t>(1, 1024 / colblocks);
  groups = std::min<int64_t>({groups, 16, std::max<int64_t>(1, splits / 8)});
  if (splits <= kSumOnePass) groups = 1;
  const int32_t gs = (int32_t)((splits + groups - 1) / groups);
  groups = (splits + gs - 1) / gs;
  hipLaunchKernelGGL(k_sum_groups, dim3((unsigned)colblocks, (unsigned)groups), dim3(256), 0, st, part, splits, gs, n,
                     out);
  if (groups > 1)
    hipLaunchKernelGGL(k_sum_strided, grid_for(n), dim3(256), 0, st, part, (int32_t)groups, (int64_t)gs * n, n, out);
  return hipGetLastError();
}

hipError_t launch_sum_splits_pair(float *pa, int32_t sa, int64_t na, float *oa, float *pb, int32_t sb, int64_t nb,
                                  float *ob, hipStream_t st) {
  if (sa < 1 || sb < 1 || sa > kSumOnePass || sb > kSumOnePass || na <= 0 || nb <= 0) {
    hipError_t e = launch_sum_splits(pa, sa, na, oa, st);
    return e != hipSuccess ? e : launch_sum_splits(pb, sb, nb, ob, st);
  }
  const int64_t ba = (na + 63) / 64, bb = (nb + 63) / 64;
  if (ba + bb > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_sum_pair, dim3((unsigned)(ba + bb)), dim3(256), 0, st, pa, sa, na, oa, ba, pb, sb, nb, ob);
  return hipGetLastError();
}

hipError_t launch_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(k_adam, grid_for(n), dim3(256), 0, st, p, g, m, v, pb, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_transpose_layers(const uint16_t *in, uint16_t *out, const TransposeTable &t, hipStream_t st) {
  if (t.layers < 1 || t.layers > kMaxTransposeLayers) return hipErrorInvalidValue;
  const int tiles = t.first_tile[t.layers];
  if (tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_transpose_layers, dim3(tiles), dim3(256), 0, st, in, out, t);
  return hipGetLastError();
}

hipError_t launch_denormalize(const float *o, int32_t bins, const double *scale, const double *delay, double df,
                              int32_t k_lo, int32_t k_hi, int32_t polar, float *out, int32_t rows, hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_denormalize, grid_for((int64_t)rows * bins), dim3(256), 0, st, o, bins, scale, delay, df, k_lo,
                     k_hi, polar, out, rows);
  return hipGetLastError();
}

// ------------------------------------------------------- fused inference

__global__ void k_pack_fragments(const uint16_t *__restrict__ w, int32_t n, int32_t k, uint4 *__restrict__ out) {
  const int ks_n = k / 16;
  const int64_t total = (int64_t)(n / 32) * ks_n * 64;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(i & 63);
    const int64_t t = i >> 6;
    const int ks = (int)(t % ks_n), nt = (int)(t / ks_n);
    const int col = nt * 32 + (lane & 31), kk = ks * 16 + (lane >> 5) * 8;
    out[i] = *(const uint4 *)(w + (int64_t)col * k + kk);
  }
}

hipError_t launch_pack_fragments(const uint16_t *w, int32_t n, int32_t k, uint16_t *out, hipStream_t st) {
  if (n % 32 || k % 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pack_fragments, grid_for((int64_t)n / 32 * (k / 16) * 64), dim3(256), 0, st, w, n, k,
                     (uint4 *)out);
  return hipGetLastError();
}

constexpr int kFusedRows = 32;        // queries per workgroup: one MFMA row tile
constexpr int kFusedWaves = 16;       // wavefronts per workgroup, splitting the output columns
constexpr int kFusedTN = 1;           // 32-column tiles a wavefront of layers 1-6 carries at once
static_assert(kFusedWaves * kFusedTN == 16, "one tile group per wavefront for the 512-wide layers");
constexpr int kFusedPitch = 512 + 8;  // LDS row pitch of the activations (BF16 elements)

// One layer of the fused chain: out = GELU(in W^T + bias) for the block's 32
// rows.  Each wavefront owns N / 32 / 8 consecutive 32-column tiles, taken
// two at a time; per k-step it reads its A fragment (row lane%32, k-slice
// 8*(lane/32)) from LDS and each tile's packed B fragment from global memory.
// The B fragments come in chunks of 8 k-steps, double-buffered in registers:
// the next chunk's loads are in flight while this chunk's MFMAs run.
template <int KS, int KC, int TN>
__device__ __forceinline__ void fetch_chunk(uint4 (&buf)[TN][KC], const uint4 *__restrict__ wp, int nt0, int c,
                                            int lane) {
#pragma unroll
  for (int t = 0; t < TN; t++)
#pragma unroll
    for (int s = 0; s < KC; s++) buf[t][s] = wp[((int64_t)(nt0 + t) * KS + c * KC + s) * 64 + lane];
}

template <int KC, int TN>
__device__ __forceinline__ void mfma_chunk(f32x16 (&acc)[TN], const uint4 (&buf)[TN][KC], const uint16_t *act_in,
                                           int c, int lane) {
#pragma unroll
  for (int s = 0; s < KC; s++) {
    const bf16x8 fa = *(const bf16x8 *)(act_in + (lane & 31) * kFusedPitch + (c * KC + s) * 16 + (lane >> 5) * 8);
#pragma unroll
    for (int t = 0; t < TN; t++) {
      bf16x8 fb;
      __builtin_memcpy(&fb, &buf[t][s], 16);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc[t], 0, 0, 0);
    }
  }
}

template <int K, int N, bool GLOBAL_OUT>
__device__ __forceinline__ void fused_layer(const uint16_t *act_in, uint16_t *act_out, const uint4 *__restrict__ wp,
                                            const float *__restrict__ bias, uint16_t *__restrict__ gout, int m0, int M,
                                            int lane, int wave) {
  constexpr int KS = K / 16, NT = N / 32, TPW = NT / kFusedWaves, TN = TPW >= 2 ? 2 : 1;
  constexpr int KC = KS < 8 ? KS : 8, NC = KS / KC;
  static_assert(TPW % TN == 0 && KS % KC == 0 && (NC == 1 || NC % 2 == 0), "whole tile pairs and chunk pairs");
#pragma unroll 1
  for (int g = 0; g < TPW; g += TN) {
    const int nt0 = wave * TPW + g;
    f32x16 acc[TN];
#pragma unroll
    for (int t = 0; t < TN; t++)
#pragma unroll
      for (int i = 0; i < 16; i++) acc[t][i] = 0.0f;
    uint4 b0[TN][KC], b1[TN][KC];
    fetch_chunk<KS, KC, TN>(b0, wp, nt0, 0, lane);
    // Fully unrolled, so the waits before each chunk's MFMAs count only the
    // loads of that chunk; the scheduling barriers keep every load of the
    // next chunk issued before this chunk's MFMAs (the scheduler would
    // otherwise sink them next to their uses, leaving few loads in flight).
#pragma unroll
    for (int c = 0; c < NC; c += 2) {
      if (c + 1 < NC) fetch_chunk<KS, KC, TN>(b1, wp, nt0, c + 1, lane);
      __builtin_amdgcn_sched_barrier(0);
      mfma_chunk<KC, TN>(acc, b0, act_in, c, lane);
      if (c + 1 < NC) {
        if (c + 2 < NC) fetch_chunk<KS, KC, TN>(b0, wp, nt0, c + 2, lane);
        __builtin_amdgcn_sched_barrier(0);
        mfma_chunk<KC, TN>(acc, b1, act_in, c + 1, lane);
      }
    }
#pragma unroll
    for (int t = 0; t < TN; t++) {
      const int col = (nt0 + t) * 32 + (lane & 31);
      const float bv = bias[col];
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = (i / 4) * 8 + (lane >> 5) * 4 + (i % 4);
        const uint16_t y = to_bf16(gelu(acc[t][i] + bv));
        if (GLOBAL_OUT) {
          if (m0 + row < M) gout[(int64_t)(m0 + row) * N + col] = y;
        } else {
          act_out[row * kFusedPitch + col] = y;
        }
      }
    }
  }
}

// Layers 1-6 as one software pipeline over "chunks": kChunkK k-steps of a
// wavefront's group of kFusedTN 32-column tiles.  Layers 1-5 (512 x 512) are
// kChunks chunks per wavefront (one tile group, K = 512); layer 6 (512 ->
// 2048) has four tile groups.  The weights do not depend on the activations,
// so the B fragments of chunk C+kAhead are loaded while chunk C's MFMAs run even
// across a tile group's epilogue or a layer's barrier: the load stream never
// drains until the last chunk.
constexpr int kChunkK = 4;                 // k-steps (16 deep) per chunk
constexpr int kChunks = 32 / kChunkK;      // chunks per tile group (K = 512)
constexpr int kPipeChunks = 5 * kChunks + 4 * kChunks;
// Chunks in flight ahead of the one whose MFMAs run: chunk C + kAhead is
// issued before chunk C's MFMAs, into the register buffer (C + kAhead) %
// kBufs.  Each buffer is 16 VGPRs.
constexpr int kAhead = 2;
constexpr int kBufs = kAhead + 1;
using ChunkBuf = uint4[kFusedTN][kChunkK];

template <int C>
struct PipeChunk {
  static constexpr int L = C < 5 * kChunks ? 1 + C / kChunks : 6;            // layer
  static constexpr int G = C < 5 * kChunks ? 0 : (C - 5 * kChunks) / kChunks;  // tile group within the wavefront's
  static constexpr int KC = C % kChunks;                                        // chunk within K = 512
  static constexpr int N = L == 6 ? 2048 : 512;
  static constexpr int TPW = N / 32 / kFusedWaves;        // tiles per wavefront
  static constexpr bool kLastOfPair = KC == kChunks - 1;
  static constexpr bool kLastOfLayer = kLastOfPair && (L < 6 || G == TPW / kFusedTN - 1);
};

template <int C>
__device__ __forceinline__ void pipe_fetch(uint4 (&buf)[kFusedTN][kChunkK], const FusedArgs &a, int wave, int lane) {
  using P = PipeChunk<C>;
  const uint4 *__restrict__ wp = (const uint4 *)a.wp[P::L];
  const int nt0 = wave * P::TPW + P::G * kFusedTN;
#pragma unroll
  for (int t = 0; t < kFusedTN; t++)
#pragma unroll
    for (int s = 0; s < kChunkK; s++) buf[t][s] = wp[((int64_t)(nt0 + t) * 32 + P::KC * kChunkK + s) * 64 + lane];
}

// LDS row pitch (BF16) of the decoder tail's 32 x 32 tiles.
constexpr int kTailPitch = 32 + 8;
constexpr int kTailTile = kFusedRows * kTailPitch;
// Each wavefront's tail tiles: the dense layer's GELU output shares its tile
// with the first convolution's second output tile, written after the loads
// that read it (a wavefront's LDS operations run in order).
constexpr int kTailTilesPerWave = 2;
// The tail's constants, staged in LDS once per workgroup (every position uses
// them): layer 7's weight fragments, layer 8's weights [8][32], layer 7's 32
// biases and layer 8's two (offsets in BF16 elements).
constexpr int kTwW7 = 0, kTwW8 = 4 * 64 * 8, kTwB7 = kTwW8 + 8 * 32, kTwB8 = kTwB7 + 64, kTwSize = kTwB8 + 8;

// Orders a wavefront's LDS writes before its following LDS reads (the tail's
// tiles are private to the wavefront, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The decoder's transposed convolutions for dense-layer position p of the
// block's 32 queries, from that position's accumulators (acc, bias bv): the
// dense layer's GELU output, then 32 -> 2 x 32 channels (bias, GELU; rows
// (query, position 2p + j) of the next input), then 32 -> 4 taps x (re, im)
// (bias), written to zout as the per-layer GEMMs would write it: row
// (query, 2p + j), column tap * 2 + component -- or, de-normalized, to spec.
// The last product is formed transposed (W8 as the A operand), so each lane
// holds four consecutive outputs of one query and stores them as one vector.
// s1 is the first of the wavefront's own two LDS tiles of 32 x kTailPitch.
__device__ __forceinline__ void decoder_tail(const FusedArgs &a, const f32x16 &acc, float bv, int p, int m0,
                                             int lane, uint16_t *s1, const uint16_t *tw) {
  const int c = lane & 31, hi = lane >> 5;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int row = (i / 4) * 8 + hi * 4 + (i % 4);
    s1[row * kTailPitch + c] = to_bf16(gelu(acc[i] + bv));
  }
  wave_lds_sync();
  // first convolution: [32 rows][32 channels] x W7^T, one 32-column output
  // tile (tap j) at a time.  Tap 0 goes to the wavefront's second tile; tap 1
  // overwrites the input tile s1, after its own operand loads from it.
  const uint4 *w7 = (const uint4 *)(tw + kTwW7);
  const float b1 = ((const float *)(tw + kTwB7))[c];
  uint16_t *const out1[2] = {s1 + kTailTile, s1};
#pragma unroll
  for (int j = 0; j < 2; j++) {
    f32x16 c1;
#pragma unroll
    for (int i = 0; i < 16; i++) c1[i] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      const bf16x8 fa = *(const bf16x8 *)(s1 + c * kTailPitch + ks * 16 + hi * 8);
      const uint4 v = w7[(j * 2 + ks) * 64 + lane];
      bf16x8 fb;
      __builtin_memcpy(&fb, &v, 16);
      c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, c1, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int row = (i / 4) * 8 + hi * 4 + (i % 4);
      out1[j][row * kTailPitch + c] = to_bf16(gelu(c1[i] + b1));
    }
  }
  wave_lds_sync();
  // second convolution, transposed: W8 [8 outputs (rows 8-31 zero, as the
  // per-layer GEMM pads them)][32 channels] x the tile^T.  Lane (query c,
  // half hi) gets outputs 4 hi .. 4 hi + 3 of that query in c2[0..3].
  const int r = m0 + c;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    f32x16 c2;
#pragma unroll
    for (int i = 0; i < 16; i++) c2[i] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (c < 8) v = *(const uint4 *)(tw + kTwW8 + c * 32 + ks * 16 + hi * 8);
      bf16x8 fw;
      __builtin_memcpy(&fw, &v, 16);
      const bf16x8 fx = *(const bf16x8 *)(out1[j] + c * kTailPitch + ks * 16 + hi * 8);
      c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fw, fx, c2, 0, 0, 0);
    }
    const int q = 2 * p + j, col = q * 8 + hi * 4;
    const float *b8 = (const float *)(tw + kTwB8);
    float4 o;
    o.x = c2[0] + b8[0];
    o.y = c2[1] + b8[1];
    o.z = c2[2] + b8[0];
    o.w = c2[3] + b8[1];
    if (a.spec) {
      if (r < a.spec_rows) {
        const double sc = a.scale[r], dl = a.delay[r];
        const int k = 4 * q + 2 * hi;  // bins k and k + 1 (taps 2 hi, 2 hi + 1)
        float4 y;
        denorm_bin(o.x, o.y, sc, dl, a.df, k, a.k_lo, a.k_hi, a.polar, y.x, y.y);
        denorm_bin(o.z, o.w, sc, dl, a.df, k + 1, a.k_lo, a.k_hi, a.polar, y.z, y.w);
        *(float4 *)(a.spec + (int64_t)r * 1024 + col) = y;
      }
    } else if (r < a.M) {
      *(float4 *)(a.zout + (int64_t)r * 1024 + col) = o;
    }
  }
  wave_lds_sync();  // the next position reuses the wavefront's tiles
}

template <int C>
__device__ __forceinline__ void pipe_run(ChunkBuf (&buf)[kBufs], f32x16 (&acc)[kFusedTN], float (&bias)[kFusedTN],
                                         uint16_t (*act)[kFusedRows * kFusedPitch], uint16_t *tail,
                                         const FusedArgs &a, int m0, int wave, int lane) {
  using P = PipeChunk<C>;
  if constexpr (C + kAhead < kPipeChunks) pipe_fetch<C + kAhead>(buf[(C + kAhead) % kBufs], a, wave, lane);
  if constexpr (P::KC == 0) {
    // the pair's biases, loaded now so that the epilogue's wait on them does
    // not also wait for the chunks prefetched after them
    const int nt0 = wave * P::TPW + P::G * kFusedTN;
#pragma unroll
    for (int t = 0; t < kFusedTN; t++) bias[t] = a.bias[P::L][(nt0 + t) * 32 + (lane & 31)];
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (P::KC == 0) {
#pragma unroll
    for (int t = 0; t < kFusedTN; t++)
#pragma unroll
      for (int i = 0; i < 16; i++) acc[t][i] = 0.0f;
  }
  mfma_chunk<kChunkK, kFusedTN>(acc, buf[C % kBufs], act[P::L & 1], P::KC, lane);
  if constexpr (P::kLastOfPair) {
    const int nt0 = wave * P::TPW + P::G * kFusedTN;
    if (P::L == 6 && (a.zout || a.spec)) {
      // each 32-column tile of the dense layer is one position (32 channels)
      uint16_t *s1 = tail + wave * kTailTilesPerWave * kTailTile;
#pragma unroll
      for (int t = 0; t < kFusedTN; t++)
        decoder_tail(a, acc[t], bias[t], nt0 + t, m0, lane, s1, tail + kFusedWaves * kTailTilesPerWave * kTailTile);
    } else {
#pragma unroll
      for (int t = 0; t < kFusedTN; t++) {
        const int col = (nt0 + t) * 32 + (lane & 31);
        const float bv = bias[t];
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const int row = (i / 4) * 8 + (lane >> 5) * 4 + (i % 4);
          const uint16_t y = to_bf16(gelu(acc[t][i] + bv));
          if constexpr (P::L == 6) {
            if (m0 + row < a.M) a.out[(int64_t)(m0 + row) * P::N + col] = y;
          } else {
            act[(P::L + 1) & 1][row * kFusedPitch + col] = y;
          }
        }
      }
    }
    if constexpr (P::kLastOfLayer && P::L < 6) __syncthreads();
  }
  if constexpr (C + 1 < kPipeChunks) pipe_run<C + 1>(buf, acc, bias, act, tail, a, m0, wave, lane);
}

__global__ __launch_bounds__(kFusedWaves * 64) void k_encoder_fused(FusedArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t act[2][kFusedRows * kFusedPitch];
  __shared__ __attribute__((aligned(16))) uint16_t tail[kFusedWaves * kTailTilesPerWave * kTailTile + kTwSize];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * kFusedRows;
  if (a.zout || a.spec) {
    // the tail's constants (read after the barrier below)
    uint16_t *tw = tail + kFusedWaves * kTailTilesPerWave * kTailTile;
    for (int v = threadIdx.x; v < 4 * 64; v += blockDim.x)
      ((uint4 *)(tw + kTwW7))[v] = ((const uint4 *)a.wp[7])[v];
    for (int v = threadIdx.x; v < 8 * 32 / 8; v += blockDim.x)
      ((uint4 *)(tw + kTwW8))[v] = ((const uint4 *)a.w8)[v];
    for (int v = threadIdx.x; v < 32; v += blockDim.x) ((float *)(tw + kTwB7))[v] = a.bias[7][v];
    for (int v = threadIdx.x; v < 2; v += blockDim.x) ((float *)(tw + kTwB8))[v] = a.bias[8][v];
  }
  // the first chunks of layer 1 are in flight while layer 0 runs
  ChunkBuf buf[kBufs];
  pipe_fetch<0>(buf[0], a, wave, lane);
  if constexpr (kAhead > 1) pipe_fetch<1>(buf[1], a, wave, lane);
  // the block's features (16 BF16 per row; rows past M read as zero)
  if (a.xf) {
    // FP32 rows of xf_cols, rounded here as k_pad_rows_bf16 rounds them
    for (int v = threadIdx.x; v < kFusedRows * 16; v += blockDim.x) {
      const int row = v >> 4, col = v & 15;
      uint16_t y = 0;
      if (m0 + row < a.M && col < a.xf_cols) y = to_bf16(a.xf[(int64_t)(m0 + row) * a.xf_cols + col]);
      act[0][row * kFusedPitch + col] = y;
    }
  } else {
    for (int v = threadIdx.x; v < kFusedRows * 2; v += blockDim.x) {
      const int row = v >> 1, half = v & 1;
      uint4 val = make_uint4(0, 0, 0, 0);
      if (m0 + row < a.M) val = *(const uint4 *)(a.x0 + (int64_t)(m0 + row) * 16 + half * 8);
      *(uint4 *)(&act[0][row * kFusedPitch + half * 8]) = val;
    }
  }
  __syncthreads();
  fused_layer<16, 512, false>(act[0], act[1], (const uint4 *)a.wp[0], a.bias[0], nullptr, m0, a.M, lane, wave);
  __syncthreads();
  f32x16 acc[kFusedTN];
  float bias[kFusedTN];
  pipe_run<0>(buf, acc, bias, act, tail, a, m0, wave, lane);
}

hipError_t launch_encoder_fused(const FusedArgs &a, hipStream_t st) {
  if (a.M <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_encoder_fused, dim3((a.M + kFusedRows - 1) / kFusedRows), dim3(kFusedWaves * 64), 0, st, a);
  return hipGetLastError();
}

}  // namespace gfsk
