// Device side of the projective Gray walk (shared by the batch, single-code
// and serving kernels).  See gt_common.h for the row layout and the piece /
// chunk decomposition.
#pragma once
#include <hip/hip_runtime.h>
#include "gt_common.h"

namespace gt {

struct RowV {
  uint32_t w[6];
};

// Rows that no thread writes while the kernel runs: plain loads, which the
// compiler turns into scalar loads for wave-uniform addresses.
struct ConstRows {
  const uint32_t* __restrict__ R;
  __device__ __forceinline__ RowV operator()(uint32_t r) const {
    const uint32_t* p = R + (size_t)r * kRowWords;
    const uint4 a = *reinterpret_cast<const uint4*>(p);
    const uint2 b = *reinterpret_cast<const uint2*>(p + 4);
    RowV v;
    v.w[0] = a.x; v.w[1] = a.y; v.w[2] = a.z; v.w[3] = a.w; v.w[4] = b.x; v.w[5] = b.y;
    return v;
  }
  // The same row held in VGPRs.  On gfx950 v_xor_b32 with two VGPR operands
  // issues at twice the rate of the form with an SGPR operand (measured:
  // 68 against 38 T lane-ops/s), so the rows of the inner walk, which are
  // wave-uniform and would live in SGPRs, are copied to VGPRs once.
  __device__ __forceinline__ RowV vgpr(const RowV& s) const {
    RowV v;
#pragma unroll
    for (int i = 0; i < 6; ++i) asm volatile("v_mov_b32 %0, %1" : "=v"(v.w[i]) : "s"(s.w[i]));
    return v;
  }
  // The same copy, ordered after `after` is computed (an operand the copy
  // does not read), so it is not hoisted above the work before it.
  __device__ __forceinline__ RowV vgpr_after(const RowV& s, uint32_t after) const {
    RowV v;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      asm volatile("v_mov_b32 %0, %1" : "=v"(v.w[i]) : "s"(s.w[i]), "v"(after));
    return v;
  }
};

// Rows in LDS (a wave's own area, written before the walk): plain loads,
// which land in VGPRs already.
struct LdsRows {
  const uint32_t* R;
  __device__ __forceinline__ RowV operator()(uint32_t r) const {
    const uint32_t* p = R + r * kRowWords;
    RowV v;
#pragma unroll
    for (int i = 0; i < 6; ++i) v.w[i] = p[i];
    return v;
  }
  __device__ __forceinline__ RowV vgpr(const RowV& v) const { return v; }
  __device__ __forceinline__ RowV vgpr_after(const RowV& v, uint32_t) const { return v; }
};

__device__ __forceinline__ void xor_row(RowV& v, const RowV& r) {
#pragma unroll
  for (int i = 0; i < 6; ++i) v.w[i] ^= r.w[i];
}

__device__ __forceinline__ void xor_row_masked(RowV& v, const RowV& r, uint32_t m) {
#pragma unroll
  for (int i = 0; i < 6; ++i) v.w[i] ^= r.w[i] & m;
}

__device__ __forceinline__ uint32_t weight(const RowV& v) {
  return __builtin_popcount(v.w[0] | v.w[2] | v.w[4]) +
         __builtin_popcount(v.w[1] | v.w[3] | v.w[5]);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off, kWave));
  return x;
}

// Two Gray steps: add row a, evaluate, add row b, evaluate.  The running
// minimum absorbs both weights in one v_min3_u32 (associated this way the
// compiler fuses it; min(best, min(wa, wb)) costs 1.5 instructions per pair).
#define GT_STEP2(A, B)                              \
  {                                                 \
    xor_row(v, A);                                  \
    const uint32_t wa = weight(v);                  \
    xor_row(v, B);                                  \
    const uint32_t wb = weight(v);                  \
    best = min(min(best, wa), wb);                  \
  }

// One chunk of Gray piece j: lanes own Gray indices (c*64 + lane) * 2^L and
// walk 2^L consecutive indices.  The walk also evaluates one codeword past the
// chunk end (row L added at the last step); it is a nonzero codeword of the
// same piece, so the minimum is unchanged.  With THR the walk stops once a
// lane of this wave, or any wave (through *best_g), reaches weight <= thr.
template <bool THR, class Rows>
__device__ __forceinline__ uint32_t gray_chunk(const Rows& row, uint32_t j, uint32_t L,
                                               uint64_t c, uint32_t lane, int thr,
                                               uint32_t check_mask, const uint32_t* best_g) {
  RowV base = row(3 * j);
  // Chunk-uniform Gray bits above the lane bits.
  uint64_t gc = c ^ (c >> 1);
  for (uint32_t r = L + kLaneBits; gc != 0; gc >>= 1, ++r)
    if (gc & 1) xor_row(base, row(r));
  // Lane bits: gray(c*64 + lane) contributes lane ^ (lane >> 1) ^ (c & 1) << 5
  // at bit L, and gray() of the shifted index sets bit L-1 when lane is odd.
  const uint32_t lp = lane ^ (lane >> 1) ^ ((uint32_t)(c & 1) << 5);
  RowV v = base;
#pragma unroll
  for (int t = 0; t < kLaneBits; ++t) xor_row_masked(v, row(L + t), 0u - ((lp >> t) & 1u));
  xor_row_masked(v, row(L - 1), 0u - (lane & 1u));

  const RowV s0 = row.vgpr(row(0)), s1 = row.vgpr(row(1)), s2 = row.vgpr(row(2)),
             s3 = row.vgpr(row(3));
  uint32_t best = weight(v);
  const uint32_t nb = 1u << (L - kU);
  for (uint32_t m = 0; m < nb; ++m) {
    // The block's last row is loaded at its top and copied to VGPRs only
    // before its last step, so the scalar load's latency is covered.
    const RowV dl = row(kU + __builtin_ctz(m + 1));
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s2)
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s3)
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s2)
    GT_STEP2(s0, s1)
    const RowV d = row.vgpr_after(dl, best);
    GT_STEP2(s0, d)
    if (THR && (m & check_mask) == check_mask) {
      if (__any((int)best <= thr)) break;
      if ((int)__hip_atomic_load(best_g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= thr) break;
    }
  }
  return best;
}
static_assert(kU == 4, "gray_chunk unrolls 16 steps");

// Blocks [m0, m1) of Gray piece j (j > kTinyTop), walked lane-major: lane l
// owns the Gray indices x = l * 2^P + s, s in [0, 2^P), P = 3j - 6, and walks
// s = 16 m0 .. 16 m1 (one past the run; a nonzero codeword of the piece, so
// the minimum is unchanged).  gray(x) = gray(l) << P ^ (l & 1) << (P - 1) ^
// gray(s), so the start is h_j, the rows of gray(16 m0) (wave-uniform), and
// the lane's rows P - 1 .. P + 5; step s -> s + 1 adds row ctz(s + 1).
template <class Rows>
__device__ __forceinline__ uint32_t gray_blocks(const Rows& row, uint32_t j, uint64_t m0,
                                                uint64_t m1, uint32_t lane) {
  const uint32_t P = 3 * j - kLaneBits;
  RowV v = row(3 * j);
  uint64_t gs = (m0 << kU) ^ (m0 << (kU - 1));
  for (uint32_t r = kU - 1; gs >> r != 0; ++r)
    if ((gs >> r) & 1) xor_row(v, row(r));
  const uint32_t lp = lane ^ (lane >> 1);
#pragma unroll
  for (int t = 0; t < kLaneBits; ++t) xor_row_masked(v, row(P + t), 0u - ((lp >> t) & 1u));
  xor_row_masked(v, row(P - 1), 0u - (lane & 1u));

  const RowV s0 = row.vgpr(row(0)), s1 = row.vgpr(row(1)), s2 = row.vgpr(row(2)),
             s3 = row.vgpr(row(3));
  uint32_t best = weight(v);
  // m0 < m1; the loop tests m != m1, which stays a scalar compare for 64-bit m.
  for (uint64_t m = m0; m != m1; ++m) {
    const RowV dl = row(kU + (uint32_t)__builtin_ctzll(m + 1));
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s2)
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s3)
    GT_STEP2(s0, s1)
    GT_STEP2(s0, s2)
    GT_STEP2(s0, s1)
    const RowV d = row.vgpr_after(dl, best);
    GT_STEP2(s0, d)
  }
  return best;
}

// Pieces 0..jt (jt <= kTinyTop = 3): piece j is row 3j plus the rows below
// it by its index bits.  One partial sum per lane serves every piece: P3 =
// rows 0..2 by the lane's bits 0..2, P = P3 + rows 3..5 by bits 3..5.  Piece
// 0 is row 0; piece 1 (lanes < 8) row 3 + P3; piece 2 row 6 + P; piece 3
// (512 codewords) row 9 + P, then 7 Gray steps over rows 6..8.  (Six masked
// row additions in all, against one set per piece: the tiny part of a
// dimension-4 subcode costs about 160 VALU instead of 250.)
template <class Rows>
__device__ __forceinline__ uint32_t tiny_chunk(const Rows& row, int jt, uint32_t lane) {
  uint32_t best = weight(row(0));
  if (jt == 0) return best;
  RowV p = {{0u, 0u, 0u, 0u, 0u, 0u}};
#pragma unroll
  for (int r = 0; r < 3; ++r) xor_row_masked(p, row(r), 0u - ((lane >> r) & 1u));
  {
    RowV v = row(3);
    xor_row(v, p);
    const uint32_t wt = weight(v);
    if (lane < 8u) best = min(best, wt);
  }
  if (jt == 1) return best;
#pragma unroll
  for (int r = 3; r < 6; ++r) xor_row_masked(p, row(r), 0u - ((lane >> r) & 1u));
  const RowV r6 = row(6);
  RowV v = r6;
  xor_row(v, p);
  best = min(best, weight(v));
  if (jt == 2) return best;
  const RowV r7 = row(7), r8 = row(8);
  v = row(9);
  xor_row(v, p);
  best = min(best, weight(v));
  xor_row(v, r6); best = min(best, weight(v));
  xor_row(v, r7); best = min(best, weight(v));
  xor_row(v, r6); best = min(best, weight(v));
  xor_row(v, r8); best = min(best, weight(v));
  xor_row(v, r6); best = min(best, weight(v));
  xor_row(v, r7); best = min(best, weight(v));
  xor_row(v, r6); best = min(best, weight(v));
  return best;
}
static_assert(kTinyTop == 3, "tiny_chunk enumerates pieces 0..3");

// Map a code-local chunk id of a walk of dimension w to (piece j, chunk c);
// j == tiny_top(w) means tiny.
__device__ __forceinline__ void locate_chunk(int w, int L, uint64_t local, int& j, uint64_t& c) {
  const int jt = tiny_top(w);
  j = w - 1;
  for (; j > jt; --j) {
    const uint64_t pc = piece_chunks(j, L);
    if (local < pc) break;
    local -= pc;
  }
  c = local;
}

// One chunk of the walk of dimension w over rows 0 .. 3w-1.
template <bool THR, class Rows>
__device__ __forceinline__ uint32_t run_chunk(const Rows& row, int w, int L, uint64_t local,
                                              uint32_t lane, int thr, uint32_t check_mask,
                                              const uint32_t* best_g) {
  int j;
  uint64_t c;
  locate_chunk(w, L, local, j, c);
  if (j > tiny_top(w))
    return gray_chunk<THR>(row, (uint32_t)j, (uint32_t)piece_L(j, L), c, lane, thr, check_mask,
                           best_g);
  return tiny_chunk(row, tiny_top(w), lane);
}

}  // namespace gt
