// The learned proposal: an 8 -> 64 -> 64 -> 8 ReLU MLP per family.
//
// No counterpart in the reference (whose kernels evaluate every cell in a
// fixed order).  The network only proposes: output 0 is the logit of "the FP32
// prefilter proves this cell", used to decide whether the FP32 screen runs
// before the FP64 criteria; outputs 1..4 score the four split axes; output 5 is
// the logit of "the second check of a two-check family fails first" (its
// first check holds), used to order the checks.  Every verdict is still
// decided by the interval evaluation.
//
// Parameters (FP32, this order): W1[8][64], b1[64], W2[64][64], b2[64],
// W3[64][8], b3[8]; W[i][j] connects input i to unit j.
//
// Inference runs on waves of 64 cells (k_propose, ahead of each round) with
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulate), one pass j = 0..3
// per lane group: pass j evaluates the 16 cells of lanes 16 j .. 16 j + 15.
// Activations are kept transposed (units x cells), so a layer's output
// registers are exactly the next layer's B operand: a lane of group g =
// lane / 16 holds, for M-tile t, units 16 t + 4 g + r (r = 0..3) of pass j's
// cell lane % 16, and the next layer's k-step s, element i contracts over unit
// 16 (2 s + i / 4) + 4 g + i % 4.  Layer 1 contracts over 32 = 4 x 8 inputs:
// W1 is repeated in each 8-wide K slice and pass j takes the features of lane
// group j in slice j (zeros in the others), so every lane supplies its own
// cell's features without a shuffle.  Layer 3's rows are placed per pass:
// outputs 0..3 in rows 4 j .. 4 j + 3 (so they land in the lanes of group j,
// the cells' own lanes) and outputs 4..7 in the rows of group (j + 1) mod 4.
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
  bf16x8 a1[4][64];        // layer 1 A operand per M-tile (W1 transposed, repeated in the four K slices)
  f32x4 b1[4][64];         // layer 1 bias of the lane's output rows
  bf16x8 a2[4][2][64];     // layer 2 A operand per M-tile and k-step
  f32x4 b2[4][64];
  bf16x8 a3[4][2][64];     // layer 3 A operand per pass and k-step (8 outputs placed in 16 rows)
  f32x4 b3[4][64];         // layer 3 bias per pass
};

// Unit index contracted by k-step s, lane group g, element i.
__host__ __device__ constexpr int mlp_unit(int s, int g, int i) { return 16 * (2 * s + i / 4) + 4 * g + (i % 4); }

// Reciprocals of the box extents (0 for an empty extent), computed once per
// kernel for mlp_features.
struct MlpBox {
  double lb[4], inv[4];
};
__device__ __forceinline__ MlpBox mlp_box(const double* box_lb, const double* box_ub) {
  MlpBox b;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double span = box_ub[a] - box_lb[a];
    b.lb[a] = box_lb[a];
    b.inv[a] = span > 0.0 ? 1.0 / span : 0.0;
  }
  return b;
}

// Features of a cell (la, r1, r2, p3 endpoints): per axis the midpoint mapped
// to [-1, 1] over the run's box and log2(width / box extent) / 16 + 1, clamped
// to [-1, 1] (-1 for a point; independent of the root grid, so a proposal
// trained on one grid applies to another over the same box).  An empty box
// extent gives features 0 and -1.
__device__ __forceinline__ void mlp_features(const double* e, const MlpBox& B, float* f) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double mid = 0.5 * (e[2 * a] + e[2 * a + 1]);
    const double w = e[2 * a + 1] - e[2 * a];
    f[2 * a] = B.inv[a] > 0.0 ? (float)(2.0 * (mid - B.lb[a]) * B.inv[a] - 1.0) : 0.0f;
    float lw = -1.0f;
    if (B.inv[a] > 0.0 && w > 0.0) lw = __log2f((float)(w * B.inv[a])) * 0.0625f + 1.0f;
    lw = lw < -1.0f ? -1.0f : lw;
    f[2 * a + 1] = lw > 1.0f ? 1.0f : lw;
  }
}

// One lane's operands of a family's network, held in registers so that a wave
// can run many 64-cell groups on one load of the weights.
struct MlpRegs {
  bf16x8 a1[4];
  f32x4 b1[4];
  bf16x8 a2[4][2];
  f32x4 b2[4];
  bf16x8 a3[4][2];
  f32x4 b3[4];
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
    W.a3[t][0] = M->a3[t][0][lane];
    W.a3[t][1] = M->a3[t][1][lane];
    W.b3[t] = M->b3[t][lane];
  }
}

// ReLU as a signed-integer maximum of the bits: a negative value (sign bit
// set) becomes +0, others are kept (one instruction; a float maximum would
// first quieten NaNs).
__device__ __forceinline__ float mlp_relu(float v) {
  const int b = __builtin_bit_cast(int, v);
  return __builtin_bit_cast(float, b > 0 ? b : 0);
}

// One wave: every lane passes its cell's features (zeros for idle lanes) and
// receives outputs 0..5 of its own cell.  Must be called by all 64 lanes.
__device__ __forceinline__ void mlp_infer_wave(const MlpRegs& W, const float* f, float* out6) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  bf16x8 own;
#pragma unroll
  for (int i = 0; i < 8; ++i) own[i] = (__bf16)f[i];
  const bf16x8 none = {};
  float o[4] = {0.0f, 0.0f, 0.0f, 0.0f}, o4 = 0.0f, o5 = 0.0f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // layer 1 B operand: K slice j holds the features of lane group j's cells
    const bf16x8 x = g == j ? own : none;
    f32x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a1[t], x, W.b1[t], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) c[r] = mlp_relu(c[r]);
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
      f32x4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a2[t][0], b[0], W.b2[t], 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a2[t][1], b[1], c, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) c[r] = mlp_relu(c[r]);
      h2[t] = c;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int i = 0; i < 8; ++i) b[s][i] = (__bf16)h2[2 * s + i / 4][i % 4];
    f32x4 y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a3[j][0], b[0], W.b3[j], 0, 0, 0);
    y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a3[j][1], b[1], y, 0, 0, 0);
    // outputs 0..3 of lane group j's cells sit in group j itself, outputs 4, 5 in group (j + 1) mod 4
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = g == j ? y[r] : o[r];
    o4 = g == ((j + 1) & 3) ? y[0] : o4;
    o5 = g == ((j + 1) & 3) ? y[1] : o5;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) out6[r] = o[r];
  out6[4] = __shfl(o4, (lane + 16) & 63, 64);
  out6[5] = __shfl(o5, (lane + 16) & 63, 64);
}
__device__ __forceinline__ void mlp_infer_wave(const MlpPacked* __restrict__ M, const float* f, float* out6) {
  MlpRegs W;
  mlp_load(M, W);
  mlp_infer_wave(W, f, out6);
}

}  // namespace cpl

// Host: packs FP32 parameters into an MlpPacked image (defined in cpl_mlp.hip).
void cpl_mlp_pack(const float* params, void* out_packed);
