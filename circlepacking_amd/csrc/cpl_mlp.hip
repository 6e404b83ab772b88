// Training and evaluation of the learned proposal's MLP (8 -> 64 -> 64 -> 8,
// ReLU) on the GPU.  No counterpart in the reference.
//
// Loss per record (features x, targets t):
//   BCE(y0, t0)                                  gate: "the FP32 prefilter proves the cell"
//   + t5 * sum_{a=1..4} (y_a - t_a)^2 / 4        split-axis scores, CRITICAL cells only
//   + BCE(y5, s)                                 check order: s = 1 when the record's FP64
//                                                verdict byte t6 is CRITICAL with failing
//                                                criterion 2 (its first check holds)
// averaged over the batch.  Kernels:
//   k_mlp_fwd_bwd   one thread per record: forward, loss, output and hidden
//                   gradients (weights staged in LDS)
//   k_mlp_wgrad     weight gradients as sums over records of activation x
//                   gradient products, one partial per record chunk
//   k_mlp_reduce    sums the chunk partials in a fixed order (deterministic)
//   k_mlp_adam      Adam step with bias correction and decoupled weight decay
//   k_mlp_infer     the in-kernel BF16 MFMA inference on caller features
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cpl_host.hpp"
#include "cpl_mlp.hpp"

using namespace cpl;
using cpl_host::DevBuf;
using cpl_host::fail;

namespace {

constexpr int REC_CHUNK = 512;  // records per weight-gradient partial
constexpr int INFER_OUT = 6;    // outputs per record of k_mlp_infer (0..5)

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float* __restrict__ X, const float* __restrict__ T, int n,
                                                     const float* __restrict__ W, float inv_n, float* __restrict__ A1,
                                                     float* __restrict__ D1, float* __restrict__ A2,
                                                     float* __restrict__ D2, float* __restrict__ DY,
                                                     float* __restrict__ Y, float* __restrict__ loss_part) {
  __shared__ float w[MLP_NPARAM];
  __shared__ float red[256];
  for (int k = threadIdx.x; k < MLP_NPARAM; k += blockDim.x) w[k] = W[k];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  float loss = 0.0f;
  if (b < n) {
    float x[MLP_IN], t[8];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) x[i] = X[(size_t)b * MLP_IN + i];
#pragma unroll
    for (int i = 0; i < 8; ++i) t[i] = T[(size_t)b * 8 + i];
    float h1[MLP_H], h2[MLP_H], y[MLP_OUT];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
      float s = w[MLP_B1 + j];
#pragma unroll
      for (int i = 0; i < MLP_IN; ++i) s = __builtin_fmaf(x[i], w[MLP_W1 + i * MLP_H + j], s);
      h1[j] = s > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
      float s = w[MLP_B2 + j];
#pragma unroll
      for (int i = 0; i < MLP_H; ++i) s = __builtin_fmaf(h1[i], w[MLP_W2 + i * MLP_H + j], s);
      h2[j] = s > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) {
      float s = w[MLP_B3 + k];
#pragma unroll
      for (int i = 0; i < MLP_H; ++i) s = __builtin_fmaf(h2[i], w[MLP_W3 + i * MLP_OUT + k], s);
      y[k] = s;
    }
    // loss and output gradients
    float dy[MLP_OUT];
    const float y0 = y[0];
    loss = fmaxf(y0, 0.0f) - y0 * t[0] + log1pf(expf(-fabsf(y0)));
    dy[0] = (1.0f / (1.0f + expf(-y0)) - t[0]) * inv_n;
#pragma unroll
    for (int a = 1; a <= 4; ++a) {
      const float d = y[a] - t[a];
      loss += t[5] * d * d * 0.25f;
      dy[a] = t[5] * d * 0.5f * inv_n;
    }
    {
      const int vb = (int)t[6];
      const float s5 = ((vb & 7) == 1 && ((vb >> 4) & 15) == 2) ? 1.0f : 0.0f;
      const float y5 = y[5];
      loss += fmaxf(y5, 0.0f) - y5 * s5 + log1pf(expf(-fabsf(y5)));
      dy[5] = (1.0f / (1.0f + expf(-y5)) - s5) * inv_n;
    }
#pragma unroll
    for (int k = 6; k < MLP_OUT; ++k) dy[k] = 0.0f;
    loss *= inv_n;
    float dh2[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < MLP_OUT; ++k) s = __builtin_fmaf(dy[k], w[MLP_W3 + j * MLP_OUT + k], s);
      dh2[j] = h2[j] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
      float s = 0.0f;
#pragma unroll
      for (int j = 0; j < MLP_H; ++j) s = __builtin_fmaf(dh2[j], w[MLP_W2 + i * MLP_H + j], s);
      D1[(size_t)b * MLP_H + i] = h1[i] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
      A1[(size_t)b * MLP_H + j] = h1[j];
      A2[(size_t)b * MLP_H + j] = h2[j];
      D2[(size_t)b * MLP_H + j] = dh2[j];
    }
#pragma unroll
    for (int k = 0; k < MLP_OUT; ++k) {
      DY[(size_t)b * MLP_OUT + k] = dy[k];
      if (Y) Y[(size_t)b * MLP_OUT + k] = y[k];
    }
  }
  red[threadIdx.x] = loss;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Parameter p's gradient over records [c * REC_CHUNK, ...) of chunk c = blockIdx.y.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float* __restrict__ X, const float* __restrict__ A1,
                                                   const float* __restrict__ D1, const float* __restrict__ A2,
                                                   const float* __restrict__ D2, const float* __restrict__ DY, int n,
                                                   float* __restrict__ part) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const int b0 = blockIdx.y * REC_CHUNK, b1 = min(n, b0 + REC_CHUNK);
  const float* A = nullptr;
  const float* D;
  int ia = 0, sa = 0, jd, sd;
  if (p < MLP_B1) {
    A = X; sa = MLP_IN; ia = (p - MLP_W1) / MLP_H; D = D1; sd = MLP_H; jd = (p - MLP_W1) % MLP_H;
  } else if (p < MLP_W2) {
    D = D1; sd = MLP_H; jd = p - MLP_B1;
  } else if (p < MLP_B2) {
    A = A1; sa = MLP_H; ia = (p - MLP_W2) / MLP_H; D = D2; sd = MLP_H; jd = (p - MLP_W2) % MLP_H;
  } else if (p < MLP_W3) {
    D = D2; sd = MLP_H; jd = p - MLP_B2;
  } else if (p < MLP_B3) {
    A = A2; sa = MLP_H; ia = (p - MLP_W3) / MLP_OUT; D = DY; sd = MLP_OUT; jd = (p - MLP_W3) % MLP_OUT;
  } else {
    D = DY; sd = MLP_OUT; jd = p - MLP_B3;
  }
  float s = 0.0f;
  if (A) {
    for (int b = b0; b < b1; ++b) s = __builtin_fmaf(A[(size_t)b * sa + ia], D[(size_t)b * sd + jd], s);
  } else {
    for (int b = b0; b < b1; ++b) s += D[(size_t)b * sd + jd];
  }
  part[(size_t)blockIdx.y * MLP_NPARAM + p] = s;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int nchunks, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float s = 0.0f;
  for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * MLP_NPARAM + p];
  grad[p] = s;
}

__global__ void k_mlp_adam(float* __restrict__ W, const float* __restrict__ grad, float* __restrict__ m1,
                           float* __restrict__ m2, float lr, float beta1, float beta2, float eps, float wd, float c1,
                           float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = grad[p];
  const float a = beta1 * m1[p] + (1.0f - beta1) * g;
  const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
  m1[p] = a;
  m2[p] = v;
  W[p] -= lr * ((a / c1) / (sqrtf(v / c2) + eps) + wd * W[p]);
}

// In-kernel inference on caller features: one wave per 64 records.
__global__ __launch_bounds__(64) void k_mlp_infer(const MlpPacked* __restrict__ M, const float* __restrict__ X,
                                                  size_t n, float* __restrict__ out) {
  const size_t b = (size_t)blockIdx.x * 64 + threadIdx.x;
  float f[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) f[i] = b < n ? X[b * MLP_IN + i] : 0.0f;
  float y[INFER_OUT];
  mlp_infer_wave(M, f, y);
  if (b < n)
#pragma unroll
    for (int k = 0; k < INFER_OUT; ++k) out[b * INFER_OUT + k] = y[k];
}

struct MlpBufs {
  DevBuf X, T, W, A1, D1, A2, D2, DY, Y, loss, part, grad, m1, m2, packed;
};
MlpBufs mb;

int ensure_train_bufs(size_t nb) {
  int rc;
  const size_t f4 = sizeof(float);
  const size_t nblk = (nb + 255) / 256, nch = (nb + REC_CHUNK - 1) / REC_CHUNK;
  if ((rc = mb.W.ensure(MLP_NPARAM * f4)) || (rc = mb.A1.ensure(nb * MLP_H * f4)) ||
      (rc = mb.D1.ensure(nb * MLP_H * f4)) || (rc = mb.A2.ensure(nb * MLP_H * f4)) ||
      (rc = mb.D2.ensure(nb * MLP_H * f4)) || (rc = mb.DY.ensure(nb * MLP_OUT * f4)) ||
      (rc = mb.Y.ensure(nb * MLP_OUT * f4)) || (rc = mb.loss.ensure(nblk * f4)) ||
      (rc = mb.part.ensure(nch * MLP_NPARAM * f4)) || (rc = mb.grad.ensure(MLP_NPARAM * f4)) ||
      (rc = mb.m1.ensure(MLP_NPARAM * f4)) || (rc = mb.m2.ensure(MLP_NPARAM * f4)))
    return rc;
  return CPL_OK;
}

// Forward + backward of records [off, off + nb) of the resident set; writes the
// gradient into mb.grad and the loss partials into mb.loss (nblk entries).
void enqueue_grad(hipStream_t s, size_t off, size_t nb, float inv_n, bool want_y) {
  const unsigned nblk = (unsigned)((nb + 255) / 256), nch = (unsigned)((nb + REC_CHUNK - 1) / REC_CHUNK);
  const float* X = mb.X.as<float>() + off * MLP_IN;
  const float* T = mb.T.as<float>() + off * 8;
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblk), dim3(256), 0, s, X, T, (int)nb, mb.W.as<float>(), inv_n,
                     mb.A1.as<float>(), mb.D1.as<float>(), mb.A2.as<float>(), mb.D2.as<float>(), mb.DY.as<float>(),
                     want_y ? mb.Y.as<float>() : nullptr, mb.loss.as<float>());
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((MLP_NPARAM + 255) / 256, nch), dim3(256), 0, s, X, mb.A1.as<float>(),
                     mb.D1.as<float>(), mb.A2.as<float>(), mb.D2.as<float>(), mb.DY.as<float>(), (int)nb,
                     mb.part.as<float>());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, mb.part.as<float>(), (int)nch,
                     mb.grad.as<float>());
}

int upload_records(const float* features, const float* targets, size_t n) {
  int rc;
  if ((rc = mb.X.ensure(n * MLP_IN * sizeof(float))) || (rc = mb.T.ensure(n * 8 * sizeof(float)))) return rc;
  hipStream_t s = cpl_host::stream();
  HIP_TRY(hipMemcpyAsync(mb.X.p, features, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, s));
  if (targets) HIP_TRY(hipMemcpyAsync(mb.T.p, targets, n * 8 * sizeof(float), hipMemcpyHostToDevice, s));
  return CPL_OK;
}

uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan kept
  u += 0x7FFFu + ((u >> 16) & 1u);                                    // round to nearest even
  return (uint16_t)(u >> 16);
}

}  // namespace

// Host-side packing of FP32 parameters into the per-lane MFMA operand order.
void cpl_mlp_pack(const float* w, void* out_packed) {
  struct Host {
    uint16_t a1[4][64][8];
    float b1[4][64][4];
    uint16_t a2[4][2][64][8];
    float b2[4][64][4];
    uint16_t a3[4][2][64][8];
    float b3[4][64][4];
  };
  static_assert(sizeof(Host) == sizeof(MlpPacked), "packed layout");
  Host h;
  std::memset(&h, 0, sizeof(h));
  for (int lane = 0; lane < 64; ++lane) {
    const int g = lane / 16, m = lane % 16;
    for (int t = 0; t < 4; ++t) {
      for (int i = 0; i < 8; ++i) {
        // layer 1 contracts over feature i in every K slice g (one slice per pass)
        h.a1[t][lane][i] = bf16_bits(w[MLP_W1 + i * MLP_H + 16 * t + m]);
        for (int s = 0; s < 2; ++s) h.a2[t][s][lane][i] = bf16_bits(w[MLP_W2 + mlp_unit(s, g, i) * MLP_H + 16 * t + m]);
      }
      for (int r = 0; r < 4; ++r) {
        h.b1[t][lane][r] = w[MLP_B1 + 16 * t + 4 * g + r];
        h.b2[t][lane][r] = w[MLP_B2 + 16 * t + 4 * g + r];
      }
    }
    // layer 3, pass j: row 4 j + r is output r, row 4 ((j + 1) mod 4) + r output 4 + r, other rows zero
    auto out_of_row = [](int j, int row) {
      if (row / 4 == j) return row % 4;
      if (row / 4 == ((j + 1) & 3)) return 4 + row % 4;
      return -1;
    };
    for (int j = 0; j < 4; ++j) {
      const int om = out_of_row(j, m);
      for (int s = 0; s < 2; ++s)
        for (int i = 0; i < 8; ++i)
          h.a3[j][s][lane][i] = bf16_bits(om >= 0 ? w[MLP_W3 + mlp_unit(s, g, i) * MLP_OUT + om] : 0.0f);
      for (int r = 0; r < 4; ++r) {
        const int orow = out_of_row(j, 4 * g + r);
        h.b3[j][lane][r] = orow >= 0 ? w[MLP_B3 + orow] : 0.0f;
      }
    }
  }
  std::memcpy(out_packed, &h, sizeof(h));
}

extern "C" {

int cpl_mlp_grad(const float* features, const float* targets, size_t n, const float* params, float* grad_out,
                 float* loss_out, float* outputs) {
  int rc = cpl_host::ensure_init();
  if (rc) return rc;
  if (!features || !targets || !params || !grad_out || n == 0 || n > (1u << 24))
    return fail(CPL_ERR_ARG, "cpl_mlp_grad: bad argument");
  if ((rc = upload_records(features, targets, n)) || (rc = ensure_train_bufs(n))) return rc;
  hipStream_t s = cpl_host::stream();
  HIP_TRY(hipMemcpyAsync(mb.W.p, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
  enqueue_grad(s, 0, n, 1.0f / (float)n, outputs != nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(grad_out, mb.grad.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
  const size_t nblk = (n + 255) / 256;
  std::vector<float> lp(nblk);
  HIP_TRY(hipMemcpyAsync(lp.data(), mb.loss.p, nblk * sizeof(float), hipMemcpyDeviceToHost, s));
  if (outputs) HIP_TRY(hipMemcpyAsync(outputs, mb.Y.p, n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (loss_out) {
    double acc = 0.0;
    for (float v : lp) acc += v;
    *loss_out = (float)acc;
  }
  return CPL_OK;
}

int cpl_mlp_train(const float* features, const float* targets, size_t n, float* params,
                  const cpl_mlp_train_cfg* cfg, float* epoch_loss) {
  int rc = cpl_host::ensure_init();
  if (rc) return rc;
  if (!features || !targets || !params || !cfg || n == 0 || cfg->epochs < 1 || cfg->batch < 1 || cfg->batch > (1 << 20))
    return fail(CPL_ERR_ARG, "cpl_mlp_train: bad argument");
  if ((rc = upload_records(features, targets, n)) || (rc = ensure_train_bufs((size_t)cfg->batch))) return rc;
  hipStream_t s = cpl_host::stream();
  HIP_TRY(hipMemcpyAsync(mb.W.p, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemsetAsync(mb.m1.p, 0, MLP_NPARAM * sizeof(float), s));
  HIP_TRY(hipMemsetAsync(mb.m2.p, 0, MLP_NPARAM * sizeof(float), s));
  const size_t B = (size_t)cfg->batch;
  const unsigned pg = (MLP_NPARAM + 255) / 256;
  int step = 0;
  std::vector<float> lp((B + 255) / 256);
  for (int ep = 0; ep < cfg->epochs; ++ep) {
    double ep_loss = 0.0;
    for (size_t off = 0; off < n; off += B) {
      const size_t nb = std::min(B, n - off);
      enqueue_grad(s, off, nb, 1.0f / (float)nb, false);
      ++step;
      const float c1 = 1.0f - std::pow(cfg->beta1, (float)step), c2 = 1.0f - std::pow(cfg->beta2, (float)step);
      hipLaunchKernelGGL(k_mlp_adam, dim3(pg), dim3(256), 0, s, mb.W.as<float>(), mb.grad.as<float>(),
                         mb.m1.as<float>(), mb.m2.as<float>(), cfg->lr, cfg->beta1, cfg->beta2, cfg->eps,
                         cfg->weight_decay, c1, c2);
      HIP_TRY(hipGetLastError());
      if (epoch_loss) {
        const size_t nblk = (nb + 255) / 256;
        HIP_TRY(hipMemcpyAsync(lp.data(), mb.loss.p, nblk * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        double acc = 0.0;
        for (size_t k = 0; k < nblk; ++k) acc += lp[k];
        ep_loss += acc * (double)nb;
      }
    }
    if (epoch_loss) epoch_loss[ep] = (float)(ep_loss / (double)n);
  }
  HIP_TRY(hipMemcpyAsync(params, mb.W.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CPL_OK;
}

int cpl_mlp_infer(const float* features, size_t n, const float* params, float* out) {
  int rc = cpl_host::ensure_init();
  if (rc) return rc;
  if (!features || !params || !out || n == 0) return fail(CPL_ERR_ARG, "cpl_mlp_infer: bad argument");
  if ((rc = upload_records(features, nullptr, n)) || (rc = mb.Y.ensure(n * INFER_OUT * sizeof(float))) ||
      (rc = mb.packed.ensure(sizeof(MlpPacked))))
    return rc;
  std::vector<char> host(sizeof(MlpPacked));
  cpl_mlp_pack(params, host.data());
  hipStream_t s = cpl_host::stream();
  HIP_TRY(hipMemcpyAsync(mb.packed.p, host.data(), host.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, mb.packed.as<MlpPacked>(),
                     mb.X.as<float>(), n, mb.Y.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, mb.Y.p, n * INFER_OUT * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return CPL_OK;
}

}  // extern "C"
