// cpl_cost_probe: issue cost of the FP64 criteria and of their parts on the
// device (no counterpart in the reference; a measurement tool for the kernels
// of libcpl_hip.so, built from the same headers).
//
// Cells: random middle-family cells (la in [0.5, 0.88], r1 near the lane-width
// bound, r2 <= r1, r3 position in [0, 1]) at depth-8 widths of the 4-parameter
// box, evaluated on the device with the product's admissible-range clipping
// (eval_synthetic); the first 4096 that reach the criteria are kept, tightened
// as the kernel body tightens them.  Each measured routine then runs in its own
// kernel (k_cost<K>, one dispatch each), so a counter pass attributes
// instructions per call:
//   rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_WAVES ... -- cpl_cost_probe
// Counts per call = counter / (SQ_WAVES x reps), reps printed below.
//
// Printed: single-wave cycles per call (s_memtime around the loop, one wave on
// the device, so the dependency latency of the routine) and, for the criteria,
// cells per microsecond with CPL_PROBE_OCC waves per SIMD (the kernel k_occ is
// compiled with __launch_bounds__(64, CPL_PROBE_OCC) and is the only caller of
// the collision-angle functions, whose register budget follows it; build with
// -DCPL_PROBE_OCC=N to compare occupancies).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <random>
#include <utility>
#include <vector>

#include "cpl_lane.hpp"

#ifndef CPL_PROBE_OCC
#define CPL_PROBE_OCC 3
#endif

using namespace cpl;
namespace {

constexpr int F = F_MIDDLE;
constexpr int KEEP = 4096;  // cells kept; 64 per lane of one wave

__global__ __launch_bounds__(256, CPL_PROBE_OCC) void k_prepare(const double* c, size_t n, uint8_t* verdict, double* tight) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* e = c + 8 * i;
  const IV la(e[0], e[1]), r1(e[2], e[3]), r2(e[4], e[5]), p3(e[6], e[7]);
  const double lbrsp = lower_bound_radius_single_placement(la.lb);
  verdict[i] = eval_synthetic<F>(la, r1, r2, p3, true, lbrsp);
  // the kernel body's inputs: clipped as eval_synthetic, tightened as eval_kernel_body
  IV r1c = r1.intersect(compute_r1_range(la, lbrsp));
  IV r2c = r2.intersect(r2_range<F>(la, r1c));
  const IV r3 = r3_from_position(r3_range<F>(la, r2c), p3);
  r2c.tighten_lb(r3.lb);
  r1c.tighten_lb(r2c.lb);
  double* t = tight + 8 * i;
  t[0] = la.lb; t[1] = la.ub; t[2] = r1c.lb; t[3] = r1c.ub; t[4] = r2c.lb; t[5] = r2c.ub; t[6] = r3.lb; t[7] = r3.ub;
}

enum : int {
  K_CRITERIA, K_CRITERIA_F32, K_F32_THEN_F64, K_SIX_ANGLES, K_IO_12, K_II_13, K_OI_23, K_OI_12, K_OO_13, K_IO_23,
  K_DASIN_RD, K_ASIN_RN, K_DIV_RD, K_IV_MUL, K_IV_DIV, K_IV_SQUARE, K_LOADS, K_COUNT
};
const char* const kNames[K_COUNT] = {
    "criteria (FP64)", "criteria (FP32 enclosure)", "FP32 criteria, FP64 where not proved", "six collision angles",
    "inner_outer(la, r1, r2)", "inner_inner(la, r1, r3)", "outer_inner(la, r2, r3)", "outer_inner(la, r1, r2)",
    "outer_outer(r1, r3)", "inner_outer(la, r2, r3)", "dasin_rd", "asin_rn", "div_rd", "interval x interval",
    "interval / interval", "interval square", "loads only (loop baseline)"};

// Lane l of the wave walks cells l, l + 64, ... (reps of them).
// Every kernel here that calls the criteria carries the same launch bound, so
// the collision-angle functions get the register budget of CPL_PROBE_OCC waves
// per SIMD (3, the round kernels' budget, by default).
template <int K>
__global__ __launch_bounds__(64, CPL_PROBE_OCC) void k_cost(const double* tight, int reps, long long* cycles, double* sink) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  const long long t0 = __builtin_amdgcn_s_memtime();
  for (int k = 0; k < reps; ++k) {
    const double* e = tight + 8 * ((size_t)k * 64 + lane);
    const IV la(e[0], e[1]), r1(e[2], e[3]), r2(e[4], e[5]), r3(e[6], e[7]);
    int trap = 0;
    if constexpr (K == K_CRITERIA) {
      acc += (double)criteria<F, double>(la, r1, r2, r3);
    } else if constexpr (K == K_CRITERIA_F32) {
      acc += (double)criteria<F, float>(to_f32(la), to_f32(r1), to_f32(r2), to_f32(r3));
    } else if constexpr (K == K_F32_THEN_F64) {
      const int c32 = criteria<F, float>(to_f32(la), to_f32(r1), to_f32(r2), to_f32(r3));
      acc += c32 == 0 ? 0.0 : (double)criteria<F, double>(la, r1, r2, r3);
    } else if constexpr (K == K_SIX_ANGLES) {
      const IV a = collision_angle_inner_outer(la, r1, r2, trap), b = collision_angle_inner_inner(la, r1, r3, trap);
      const IV c = collision_angle_outer_inner(la, r2, r3, trap), d = collision_angle_outer_inner(la, r1, r2, trap);
      const IV f = collision_angle_outer_outer(r1, r3, trap), g = collision_angle_inner_outer(la, r2, r3, trap);
      acc += a.ub + b.ub + c.ub + d.ub + f.ub + g.ub + trap;
    } else if constexpr (K == K_IO_12) {
      acc += collision_angle_inner_outer(la, r1, r2, trap).ub + trap;
    } else if constexpr (K == K_II_13) {
      acc += collision_angle_inner_inner(la, r1, r3, trap).ub + trap;
    } else if constexpr (K == K_OI_23) {
      acc += collision_angle_outer_inner(la, r2, r3, trap).ub + trap;
    } else if constexpr (K == K_OI_12) {
      acc += collision_angle_outer_inner(la, r1, r2, trap).ub + trap;
    } else if constexpr (K == K_OO_13) {
      acc += collision_angle_outer_outer(r1, r3, trap).ub + trap;
    } else if constexpr (K == K_IO_23) {
      acc += collision_angle_inner_outer(la, r2, r3, trap).ub + trap;
    } else if constexpr (K == K_DASIN_RD) {
      acc += dasin_rd(r2.lb / r1.ub);  // a collision quotient in (0, 1]
    } else if constexpr (K == K_ASIN_RN) {
      acc += asin_rn(r2.lb / r1.ub);
    } else if constexpr (K == K_DIV_RD) {
      acc += div_rd(r2.lb, r1.ub);
    } else if constexpr (K == K_IV_MUL) {
      const IV p = (la + r1) * (r2 + r3);
      acc += p.lb + p.ub;
    } else if constexpr (K == K_IV_DIV) {
      const IV p = (la + r1) / (r2 + la);
      acc += p.lb + p.ub;
    } else if constexpr (K == K_IV_SQUARE) {
      const IV p = (la + r1).square();
      acc += p.lb + p.ub;
    } else {
      acc += la.lb + r1.ub + r2.lb + r3.ub;
    }
  }
  const long long t1 = __builtin_amdgcn_s_memtime();
  sink[(size_t)blockIdx.x * 64 + lane] = acc;
  if (lane == 0) cycles[blockIdx.x] = t1 - t0;
}

__global__ __launch_bounds__(64, CPL_PROBE_OCC) void k_occ(const double* tight, int reps, double* sink) {
  const int lane = threadIdx.x & 63;
  double acc = 0.0;
  for (int k = 0; k < reps; ++k) {
    const double* e = tight + 8 * ((size_t)k * 64 + lane);
    acc += (double)criteria<F, double>(IV(e[0], e[1]), IV(e[2], e[3]), IV(e[4], e[5]), IV(e[6], e[7]));
  }
  sink[(size_t)blockIdx.x * 64 + lane] = acc;
}

template <int K>
void launch_cost(int blocks, const double* tight, int reps, long long* cycles, double* sink) {
  hipLaunchKernelGGL(k_cost<K>, dim3(blocks), dim3(64), 0, 0, tight, reps, cycles, sink);
}
template <int... Ks>
void run_all(std::integer_sequence<int, Ks...>, const double* tight, int reps, long long* cycles, double* sink) {
  auto one = [&](auto tag, int k) {
    constexpr int K = decltype(tag)::value;
    launch_cost<K>(1, tight, reps, cycles, sink);
    long long c = 0;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&c, cycles, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess) {
      std::printf("HIP error\n");
      return;
    }
    std::printf("k_cost<%2d> %-40s %10.1f cycles per call (one wave)\n", k, kNames[k], (double)c / reps);
  };
  (one(std::integral_constant<int, Ks>{}, Ks), ...);
}

#define CHECK(x)                                                            \
  do {                                                                      \
    const hipError_t err_ = (x);                                            \
    if (err_ != hipSuccess) {                                               \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(err_)); \
      return 1;                                                             \
    }                                                                       \
  } while (0)

}  // namespace

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  const size_t N = size_t(1) << 20;
  std::vector<double> cells(8 * N);
  for (size_t i = 0; i < N; ++i) {
    const double la = 0.5 + 0.38 * U(rng), r1 = 0.25 * (1 - la) * (0.95 + 0.35 * U(rng));
    const double r2 = r1 * (0.3 + 0.75 * U(rng)), p = 0.99 * U(rng);
    double* e = &cells[8 * i];
    e[0] = la; e[1] = la + 0.49 / 512;   // depth-8 widths of the 4-parameter box
    e[2] = r1; e[3] = r1 + 0.45 / 2048;
    e[4] = r2; e[5] = r2 + 0.45 / 2048;
    e[6] = p;  e[7] = p + 1.0 / 256;
  }
  double *d_cells = nullptr, *d_tight = nullptr;
  uint8_t* d_verdict = nullptr;
  CHECK(hipMalloc(&d_cells, cells.size() * sizeof(double)));
  CHECK(hipMalloc(&d_tight, cells.size() * sizeof(double)));
  CHECK(hipMalloc(&d_verdict, N));
  CHECK(hipMemcpy(d_cells, cells.data(), cells.size() * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_prepare, dim3((unsigned)(N / 256)), dim3(256), 0, 0, d_cells, N, d_verdict, d_tight);
  CHECK(hipGetLastError());
  std::vector<uint8_t> verdict(N);
  std::vector<double> tight(8 * N);
  CHECK(hipMemcpy(verdict.data(), d_verdict, N, hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(tight.data(), d_tight, tight.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> kept;
  size_t proven = 0, critical = 0;
  for (size_t i = 0; i < N && kept.size() < 8 * (size_t)KEEP; ++i) {
    const int v = verdict[i] & 7;
    if (v != V_PROVEN && v != V_CRITICAL) continue;
    (v == V_PROVEN ? proven : critical)++;
    kept.insert(kept.end(), &tight[8 * i], &tight[8 * i] + 8);
  }
  const int reps = (int)(kept.size() / 8 / 64);
  std::printf("cells reaching the criteria: %zu kept (%zu PROVEN, %zu CRITICAL); reps = %d per lane\n", kept.size() / 8,
              proven, critical, reps);
  if (reps < 1) return 1;
  double *d_kept = nullptr, *sink = nullptr;
  long long* d_cycles = nullptr;
  const int occ_blocks = CPL_PROBE_OCC * 1024;  // CPL_PROBE_OCC waves on each of the 1024 SIMDs
  CHECK(hipMalloc(&d_kept, kept.size() * sizeof(double)));
  CHECK(hipMemcpy(d_kept, kept.data(), kept.size() * sizeof(double), hipMemcpyHostToDevice));
  CHECK(hipMalloc(&d_cycles, sizeof(long long)));
  CHECK(hipMalloc(&sink, (size_t)occ_blocks * 64 * sizeof(double)));

  run_all(std::make_integer_sequence<int, K_COUNT>{}, d_kept, reps, d_cycles, sink);

  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
    CHECK(hipEventRecord(a));
    hipLaunchKernelGGL(k_occ, dim3(occ_blocks), dim3(64), 0, 0, d_kept, reps, sink);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    if (rep == 1)
      std::printf("k_occ: criteria at %d waves per SIMD: %.3f ms, %.1f cells/us\n", CPL_PROBE_OCC, ms,
                  (double)occ_blocks * 64 * reps / (ms * 1e3));
  }
  return 0;
}
