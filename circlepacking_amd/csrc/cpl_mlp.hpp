// The learned proposal: an 8 -> 64 -> 64 -> 8 ReLU MLP per family.
//
// No counterpart in the reference (whose kernels evaluate every cell in a
// fixed order).  The network only proposes: output 0 is the logit of "the FP32
// prefilter proves this cell", used to decide whether the FP32 screen runs
// before the FP64 criteria; outputs 1..4 score the four split axes.  Every
// verdict is still decided by the interval evaluation.
//
// Parameters (FP32, this order): W1[8][64], b1[64], W2[64][64], b2[64],
// W3[64][8], b3[8]; W[i][j] connects input i to unit j.
//
// Inference runs on waves of 64 cells (k_propose, ahead of each round) with
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulate).  Activations are
// kept transposed (units x cells), so a layer's output registers are exactly
// the next layer's B operand: a lane of group g = lane / 16 holds, for M-tile
// t, units 16 t + 4 g + r (r = 0..3) of cell 16 j + lane % 16, and the next
// layer's k-step s, element i contracts over unit 16 (2 s + i / 4) + 4 g + i % 4.
// The host packs the weights into that per-lane order once (MlpPacked).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cpl {

constexpr int MLP_IN = 8, MLP_H = 64, MLP_OUT = 8;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_IN * MLP_H, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_OUT, MLP_NPARAM = MLP_B3 + MLP_OUT;
static_assert(MLP_NPARAM == 5256, "parameter count");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Per-lane packed operands of one family's network (64 lanes each).
struct MlpPacked {
  bf16x8 a1[4][64];     // layer 1 A operand per M-tile (W1 transposed, K padded 8 -> 32)
  f32x4 b1[4][64];      // layer 1 bias of the lane's output rows
  bf16x8 a2[4][2][64];  // layer 2 A operand per M-tile and k-step
  f32x4 b2[4][64];
  bf16x8 a3[2][64];     // layer 3 A operand per k-step (8 outputs padded to 16 rows)
  f32x4 b3[64];
};

// Unit index contracted by k-step s, lane group g, element i.
__host__ __device__ constexpr int mlp_unit(int s, int g, int i) { return 16 * (2 * s + i / 4) + 4 * g + (i % 4); }

// Features of a cell (la, r1, r2, p3 endpoints): per axis the midpoint mapped
// to [-1, 1] over the run's box and log2(width / box extent) / 16 + 1, clamped
// to [-1, 1] (-1 for a point; independent of the root grid, so a proposal
// trained on one grid applies to another over the same box).
__device__ __forceinline__ void mlp_features(const double* e, const double* box_lb, const double* box_ub, float* f) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double span = box_ub[a] - box_lb[a];
    const double mid = 0.5 * (e[2 * a] + e[2 * a + 1]);
    const double w = e[2 * a + 1] - e[2 * a];
    f[2 * a] = span > 0.0 ? (float)(2.0 * (mid - box_lb[a]) / span - 1.0) : 0.0f;
    float lw = -1.0f;
    if (span > 0.0 && w > 0.0) lw = __log2f((float)(w / span)) * 0.0625f + 1.0f;
    lw = lw < -1.0f ? -1.0f : lw;
    f[2 * a + 1] = lw > 1.0f ? 1.0f : lw;
  }
}

__device__ __forceinline__ float mlp_shfl(float v, int src) { return __shfl(v, src, 64); }

// One lane's operands of a family's network, held in registers so that a wave
// can run many 64-cell groups on one load of the weights.
struct MlpRegs {
  bf16x8 a1[4];
  f32x4 b1[4];
  bf16x8 a2[4][2];
  f32x4 b2[4];
  bf16x8 a3[2];
  f32x4 b3;
};
__device__ __forceinline__ void mlp_load(const MlpPacked* __restrict__ M, MlpRegs& W) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    W.a1[t] = M->a1[t][lane];
    W.b1[t] = M->b1[t][lane];
    W.a2[t][0] = M->a2[t][0][lane];
    W.a2[t][1] = M->a2[t][1][lane];
    W.b2[t] = M->b2[t][lane];
  }
  W.a3[0] = M->a3[0][lane];
  W.a3[1] = M->a3[1][lane];
  W.b3 = M->b3[lane];
}

// One wave: every lane passes its cell's features (zeros for idle lanes) and
// receives outputs 0..4 of its own cell.  Must be called by all 64 lanes.
__device__ __forceinline__ void mlp_infer_wave(const MlpRegs& W, const float* f, float* out5) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, m = lane & 15;
  float res[4][5];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // layer 1 B operand: features of cell 16 j + m in lane group 0, zeros elsewhere
    bf16x8 x;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float v = mlp_shfl(f[i], 16 * j + m);
      x[i] = (__bf16)(g == 0 ? v : 0.0f);
    }
    f32x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = W.b1[t];
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a1[t], x, c, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) c[r] = c[r] > 0.0f ? c[r] : 0.0f;
      h1[t] = c;
    }
    bf16x8 b[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) b[s][i] = (__bf16)h1[2 * s + i / 4][i % 4];
    f32x4 h2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = W.b2[t];
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a2[t][0], b[0], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a2[t][1], b[1], c, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) c[r] = c[r] > 0.0f ? c[r] : 0.0f;
      h2[t] = c;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) b[s][i] = (__bf16)h2[2 * s + i / 4][i % 4];
    f32x4 y = W.b3;
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a3[0], b[0], y, 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a3[1], b[1], y, 0, 0, 0);
    // outputs 0..3 of cell 16 j + m sit in lane m, output 4 in lane 16 + m
    const int src = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) res[j][r] = mlp_shfl(y[r], src);
    res[j][4] = mlp_shfl(y[0], 16 + src);
  }
  const int mine = lane >> 4;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    float v = res[0][k];
    v = mine == 1 ? res[1][k] : v;
    v = mine == 2 ? res[2][k] : v;
    v = mine == 3 ? res[3][k] : v;
    out5[k] = v;
  }
}
__device__ __forceinline__ void mlp_infer_wave(const MlpPacked* __restrict__ M, const float* f, float* out5) {
  MlpRegs W;
  mlp_load(M, W);
  mlp_infer_wave(W, f, out5);
}

}  // namespace cpl

// Host: packs FP32 parameters into an MlpPacked image (defined in cpl_mlp.hip).
void cpl_mlp_pack(const float* params, void* out_packed);
