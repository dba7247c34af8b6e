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

// Tile `id` of a contraction over sb image slots with n_it row tiles and n_ct column tiles.
// Slots go in groups of sg (the last group may be shorter); inside a group the column tile
// is slowest, then the row tile, and the slot fastest, so that the tiles an XCD runs together
// share the row operand and any slot-independent epilogue operands and, per slot, the column
// panel.
struct SlotTile {
  int sl, it, ct;
};
__device__ inline SlotTile slot_group_tile(int id, int sb, int sg, int n_it, int n_ct) {
  const int per_g = sg * n_it * n_ct;
  const int grp = id / per_g, s0 = grp * sg;
  const int gsz = min(sg, sb - s0);
  const int rest = id - grp * per_g;
  const int ct = rest / (n_it * gsz);
  const int r2 = rest - ct * n_it * gsz;
  const int it = r2 / gsz;
  return SlotTile{s0 + (r2 - it * gsz), it, ct};
}

// Number of 16-wide blocks, out of `cap` starting at `start`, below `valid`; wave-uniform.
__device__ inline int clamp_blocks(int valid, int start, int cap) {
  int b = (valid - start + 15) / 16;
  b = b < 0 ? 0 : (b > cap ? cap : b);
  return __builtin_amdgcn_readfirstlane(b);
}

// Sum of a double over each 16-lane row of the wave, in every lane of the row, with DPP
// moves (VALU, no LDS): pairs within quads (quad_perm [1,0,3,2], [2,3,0,1]), then the
// mirrored quad of the half row (row_half_mirror) and the mirrored half (row_mirror). Every
// lane adds the same two partial sums at each step, so all lanes hold the same value.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const int2 u = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_mov_dpp(u.x, CTRL, 0xF, 0xF, false);
  r.y = __builtin_amdgcn_mov_dpp(u.y, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}

__device__ inline double row16_sum(double v) {
  v += dpp_f64<0xB1>(v);    // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E>(v);    // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141>(v);   // row_half_mirror
  v += dpp_f64<0x140>(v);   // row_mirror
  return v;
}

__device__ inline v4f mfma_bf16(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace bfx
}  // namespace tdscha
