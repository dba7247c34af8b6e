// Sustained MFMA rate of the board: a stream of independent MFMAs with pseudo-random
// operands on every CU, timed after the clock has settled under the power limit. It is the
// reference point for the contraction kernels' rooflines: the published peaks assume the
// maximum clock, which a dense MFMA load does not keep.
//
// BF16: v_mfma_f32_16x16x32_bf16 (16384 flop), FP64: v_mfma_f64_16x16x4_f64 (2048 flop).
// Each wave keeps 8 accumulators (independent chains) and cycles through 4 operand sets so
// that the multiplier inputs change from one instruction to the next, as real data does.
#include <chrono>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

namespace tdscha {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kStreamBlocks = 2048, kStreamThreads = 256;

__device__ inline float rnd(unsigned &x) {
  x = x * 1664525u + 1013904223u;
  return (float)(x >> 8) * (1.0f / 16777216.0f) - 0.5f;
}

__global__ __launch_bounds__(kStreamThreads, 2) void k_stream_bf16(float *out, int iters) {
  v8bf a[4], b[4];
  unsigned x = 2654435761u * (threadIdx.x + 1) + blockIdx.x * 97u;
  for (int q = 0; q < 4; ++q)
    for (int i = 0; i < 8; ++i) {
      a[q][i] = (__bf16)rnd(x);
      b[q][i] = (__bf16)rnd(x);
    }
  v4f acc[8];
  for (int j = 0; j < 8; ++j) acc[j] = (v4f){0.f, 0.f, 0.f, 0.f};
  for (int it = 0; it < iters; it += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j & 3], b[(j + u) & 3], acc[j], 0, 0,
                                                         0);
  }
  float s = 0.f;
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(kStreamThreads, 2) void k_stream_f64(float *out, int iters) {
  double a[4], b[4];
  unsigned x = 2654435761u * (threadIdx.x + 1) + blockIdx.x * 97u;
  for (int q = 0; q < 4; ++q) {
    a[q] = (double)rnd(x) + 1e-9 * rnd(x);
    b[q] = (double)rnd(x) + 1e-9 * rnd(x);
  }
  v4d acc[8];
  for (int j = 0; j < 8; ++j) acc[j] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int it = 0; it < iters; it += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j & 3], b[(j + u) & 3], acc[j], 0, 0, 0);
  }
  double s = 0.0;
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
}

}  // namespace
}  // namespace tdscha

using namespace tdscha;

extern "C" int tdscha_mfma_stream(int precision, double seconds, double *tflops,
                                  double *ms_per_launch) {
  if ((precision != TDSCHA_PREC_BF16X3 && precision != TDSCHA_PREC_FP64) || !(seconds > 0) ||
      !tflops) {
    set_error("tdscha_mfma_stream: precision TDSCHA_PREC_BF16X3 or TDSCHA_PREC_FP64, "
              "seconds > 0 and a result pointer");
    return 1;
  }
  const bool bf = precision == TDSCHA_PREC_BF16X3;
  // about 10 ms per launch at the sustained rate
  const int iters = bf ? 8192 : 4096;
  const double waves = (double)kStreamBlocks * kStreamThreads / 64;
  const double flop = (bf ? 16384.0 : 2048.0) * 8.0 * iters * waves;
  float *out = nullptr;
  if (hipMalloc(&out, (size_t)kStreamBlocks * kStreamThreads * sizeof(float)) != hipSuccess) {
    set_error("tdscha_mfma_stream: hipMalloc failed");
    return 11;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = 10;
  double timed_ms = 0.0;
  long long timed = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (!rc) {
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (el >= seconds) break;
    const bool time_it = el >= 0.5 * seconds;   // the second half, after the clock settles
    if (time_it) (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < 10; ++r) {
      if (bf)
        hipLaunchKernelGGL(k_stream_bf16, dim3(kStreamBlocks), dim3(kStreamThreads), 0, nullptr,
                           out, iters);
      else
        hipLaunchKernelGGL(k_stream_f64, dim3(kStreamBlocks), dim3(kStreamThreads), 0, nullptr,
                           out, iters);
    }
    if (time_it) {
      (void)hipEventRecord(e1, nullptr);
      if (hipEventSynchronize(e1) != hipSuccess) { rc = 10; break; }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      timed_ms += ms;
      timed += 10;
    } else if (hipDeviceSynchronize() != hipSuccess) {
      rc = 10;
    }
  }
  if (!rc && hipGetLastError() != hipSuccess) rc = 10;
  if (rc) set_error("tdscha_mfma_stream: device execution failed");
  if (!rc && timed == 0) {
    set_error("tdscha_mfma_stream: no launch was timed (seconds too short)");
    rc = 1;
  }
  if (!rc) {
    *tflops = flop * timed / (timed_ms * 1e-3) / 1e12;
    if (ms_per_launch) *ms_per_launch = timed_ms / timed;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}
