// surrogate_kernels.h — launchers of the surrogate's kernels (surrogate.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

namespace gfsk {

enum Epilogue {
  kEpiF32 = 0,       // cf[split][m][n] = acc                       (weight gradients)
  kEpiBias = 1,      // cf[m][n] = acc + bias[n % bias_mod]         (output layer)
  kEpiBiasGelu = 2,  // z[m][n] = acc + bias (if z); cb = bf16(gelu(z))   (hidden layers)
  kEpiGeluBwd = 3,   // d = acc * gelu'(z[m][n]); cf = d (if cf); cb = bf16(d);
                     // colpart[row tile][n] = sum of d over the tile's rows (if colpart)
};

// C[M][N] = sum_k A(m,k) B(k,n) over k in this split's range.  A(m,k) is
// A[m*lda + k] (at = false) or A[k*lda + m] (at = true); B(k,n) is
// B[n*ldb + k] (bt = true) or B[k*ldb + n] (bt = false).  BF16 operands as
// raw uint16 bits.  Requirements (checked by the host code): K a multiple of
// 8; M a multiple of 8 when at; N a multiple of 8 when !bt; lda, ldb multiples
// of 8 and both operands 16-byte aligned; k_split a multiple of 32.
struct GemmArgs {
  int32_t M = 0, N = 0, K = 0;
  const uint16_t *A = nullptr;
  int64_t lda = 0;
  const uint16_t *B = nullptr;
  int64_t ldb = 0;
  float *cf = nullptr;
  int64_t ldc = 0;
  uint16_t *cb = nullptr;
  int64_t ldcb = 0;
  const float *bias = nullptr;
  int32_t bias_mod = 1;
  float *z = nullptr;
  int64_t ldz = 0;
  int32_t k_split = 1 << 30;
  int64_t split_stride = 0;
  float *colpart = nullptr;  // [gemm_row_tiles(...)][N]
};

hipError_t launch_gemm(const GemmArgs &g, bool at, bool bt, int epi, int splits, hipStream_t st);
// The number of row tiles (grid rows) launch_gemm uses for these arguments.
int32_t gemm_row_tiles(const GemmArgs &g, bool at, bool bt, int epi);
hipError_t launch_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t st);
hipError_t launch_pad_rows_bf16(const float *in, int32_t cols_in, uint16_t *out, int32_t cols_out, int64_t rows,
                                hipStream_t st);
hipError_t launch_gather_rows(const float *src, const int32_t *idx, int32_t cols, float *dst, int32_t rows,
                              hipStream_t st);
hipError_t launch_rel_l2(const float *o, const float *t, const float *w, int32_t rows, int32_t len, float scale,
                         float *loss, float *dF, uint16_t *dB, hipStream_t st);
// Bias gradient: column sums of x[rows][cols] folded onto mod channels.
// part needs bias_grad_part_floats(rows, cols) floats for any row count up
// to rows; cols is a multiple of 4 that divides 256, or a multiple of 256.
int64_t bias_grad_part_floats(int64_t max_rows, int32_t cols);
hipError_t launch_bias_grad(const float *x, int64_t rows, int32_t cols, int32_t mod, float *part, float *out,
                            hipStream_t st);
// out[i] = sum over the splits of part[s][i]; part is overwritten (it holds
// partial sums afterwards).
hipError_t launch_sum_splits(float *part, int32_t splits, int64_t n, float *out, hipStream_t st);
// Slices summed in one pass by launch_sum_splits (more take two passes).
constexpr int32_t kSumOnePass = 64;
// Two launch_sum_splits jobs; in one launch when both take one pass (the
// same sums, in the same order, as two calls).
hipError_t launch_sum_splits_pair(float *pa, int32_t sa, int64_t na, float *oa, float *pb, int32_t sb, int64_t nb,
                                  float *ob, hipStream_t st);
hipError_t launch_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st);
// Per-layer BF16 transposes in one launch: layer i's in[off + n*k_i + k]
// ([n_i][k_i]) goes to out[off + k*n_i + n].  first_tile[i] counts the 16 x 16
// tiles of the layers before i (first_tile[layers] = all of them).
constexpr int kMaxTransposeLayers = 16;
struct TransposeTable {
  int32_t layers = 0;
  int64_t off[kMaxTransposeLayers] = {};
  int32_t n[kMaxTransposeLayers] = {}, k[kMaxTransposeLayers] = {};
  int32_t first_tile[kMaxTransposeLayers + 1] = {};
};
hipError_t launch_transpose_layers(const uint16_t *in, uint16_t *out, const TransposeTable &t, hipStream_t st);
hipError_t launch_denormalize(const float *o, int32_t bins, const double *scale, const double *delay, double df,
                              int32_t k_lo, int32_t k_hi, int32_t polar, float *out, int32_t rows, hipStream_t st);

// Inference without per-layer launches.  A layer's BF16 weights W[N][K]
// (N a multiple of 32, K of 16) packed in MFMA fragment order: 16-byte
// element (nt * K/16 + ks) * 64 + lane holds W[nt*32 + lane%32][ks*16 +
// 8*(lane/32) .. +7], so a wavefront's B operand for one 32-column tile and
// one 16-deep k-step is one contiguous 1 KB load.
hipError_t launch_pack_fragments(const uint16_t *w, int32_t n, int32_t k, uint16_t *out, hipStream_t st);

// The whole network in one kernel: for each block of 32 queries, x0 [M][16]
// (BF16 features) -> GELU(x W0^T + b0) -> five 512-wide GELU layers ->
// GELU(h W6^T + b6), read as 64 positions x 32 channels -> the two
// transposed convolutions (32 -> 2 x 32 channels with GELU, then 32 -> 4 x 2
// outputs) -> zout [M][1024] (FP32).  The block's activations stay in LDS
// between layers; the packed weights stream from L2; each wavefront runs the
// convolutions on every dense-layer position it forms, in its epilogue.
// With spec set, the output is de-normalized in the same epilogue (as
// launch_denormalize would: spec [spec_rows][512] complex) instead of
// written to zout.  With both null the kernel stops after the dense layer
// and writes its output to out [M][2048] (BF16).  Same arithmetic as the
// per-layer GEMMs (k-steps in order, bias and GELU in FP32, one BF16
// rounding per output).
struct FusedArgs {
  const uint16_t *x0 = nullptr;
  // xf set: the features as FP32 [M][xf_cols] (xf_cols < 16), rounded to
  // BF16 and zero-padded to 16 in the kernel, instead of x0.
  const float *xf = nullptr;
  int32_t xf_cols = 0;
  int32_t M = 0;
  const uint16_t *wp[8] = {};  // layers 0-7 packed in fragment order
  const float *bias[9] = {};
  const uint16_t *w8 = nullptr;  // layer 8's BF16 weights [8][32], as stored
  uint16_t *out = nullptr;
  float *zout = nullptr;
  float *spec = nullptr;
  int32_t spec_rows = 0;
  const double *scale = nullptr, *delay = nullptr;
  double df = 0;
  int32_t k_lo = 0, k_hi = 0, polar = 0;
};
hipError_t launch_encoder_fused(const FusedArgs &a, hipStream_t st);

}  // namespace gfsk
