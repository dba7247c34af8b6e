// Device helpers shared by the BF16x3 contraction kernels (gemm_bf16.hip, gemm_bf16_rp.hip).
//
// a*b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi with hi = RNE_bf16(v), lo = RNE_bf16(v - hi):
// three v_mfma_f32_16x16x32_bf16 per operand pair, FP32 accumulation.
#pragma once

#include "tdscha_internal.h"

namespace tdscha {
namespace bfx {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));

constexpr int TK = 32;           // k slice (one MFMA deep)

// Pairs of FP32 values split into packed bf16 hi / lo (v_cvt_pk_bf16_f32, nearest even).
__device__ inline void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
  const v2f f = {a, b};
  const v2bf h = __builtin_convertvector(f, v2bf);
  const v2f hf = __builtin_convertvector(h, v2f);
  const v2bf l = __builtin_convertvector(f - hf, v2bf);
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, l);
}

// Workgroup id -> logical tile id, so that XCD x (workgroups are dealt out to the 8 XCDs
// round-robin by id) runs logical ids [x * per, (x + 1) * per). The grid has 8 * per
// workgroups; ids past `total` return.
__device__ inline int xcd_tile(int total) {
  const int per = (total + 7) / 8;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Number of 16-wide blocks, out of `cap` starting at `start`, below `valid`; wave-uniform.
__device__ inline int clamp_blocks(int valid, int start, int cap) {
  int b = (valid - start + 15) / 16;
  b = b < 0 ? 0 : (b > cap ? cap : b);
  return __builtin_amdgcn_readfirstlane(b);
}

__device__ inline v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace bfx
}  // namespace tdscha
