// Anchored subcodes on the device (gt_common.h): bit rows of the subcodes
// vanishing at a level's representative point sets, by Gaussian elimination
// on the codewords, and the whole exact search of one small code by one
// wave (rows in LDS).  Shared by the batch kernels and the serving kernel.
#pragma once
#include <hip/hip_runtime.h>
#include "gt_common.h"
#include "gt_walk.h"

namespace gt {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Result of a code whose minimum weight is b (max_zeros clamped at tmax + 1).
__device__ __forceinline__ int32_t md_result(int b, int tmax) {
  return (b <= 48 - tmax) ? tmax + 1 : 49 - b;
}

// ---- kernel bases of the anchored subcodes (gt_common.h) -------------------
// GF(8) elements are 3-bit polynomials; alpha^e (e = 0..6, alpha = x) and
// inverses packed three bits per entry.
constexpr uint32_t kF8Exp = 1u | 2u << 3 | 4u << 6 | 3u << 9 | 6u << 12 | 7u << 15 | 5u << 18;
constexpr uint32_t kF8Inv = 0u | 1u << 3 | 5u << 6 | 6u << 9 | 7u << 12 | 2u << 15 | 3u << 18 |
                            4u << 21;

// Product in GF(8) = GF(2)[x]/(x^3 + x + 1) of 3-bit polynomials.
GT_HD constexpr uint32_t f8_mul(uint32_t a, uint32_t b) {
  uint32_t p = ((b & 1u) ? a : 0u) ^ ((b & 2u) ? a << 1 : 0u) ^ ((b & 4u) ? a << 2 : 0u);
  if (p & 16u) p ^= 0x16u;   // x^4 = x^2 + x
  if (p & 8u) p ^= 0x0Bu;    // x^3 = x + 1
  return p;
}
// The products v * b for v = 0..7, three bits each (bits 3v .. 3v + 2).
GT_HD constexpr uint32_t f8_mul_row(uint32_t b) {
  uint32_t t = 0u;
  for (uint32_t v = 1; v < 8; ++v) t |= f8_mul(v, b) << (3 * v);
  return t;
}

// Evaluation-matrix column (point index) of the exponent vector r d on the
// line of direction d: (1,0), (0,1), (1,1), ..., (1,6) for d = 0..7.  Point
// j is (x, y) = (j / 7 + 1, j % 7 + 1) as GF(8) elements, (alpha^u, alpha^v)
// = (kF8Exp[u], kF8Exp[v]).
__device__ __forceinline__ int line_point(int d, int r) {
  const int u = (d == 1) ? 0 : r;
  const int v = (d == 0) ? 0 : (d == 1) ? r : (r * (d - 1)) % 7;
  const int x = (int)((kF8Exp >> (3 * u)) & 7u), y = (int)((kF8Exp >> (3 * v)) & 7u);
  return (x - 1) * 7 + (y - 1);
}

// Value of the codeword h (bit planes) at point p (wave-uniform: a uniform
// branch picks the words, three bit-field extracts read the planes).
__device__ __forceinline__ uint32_t cw_at(const uint32_t h[6], int p) {
  const uint32_t bit = (uint32_t)p & 31u;
  uint32_t a, b, c;
  if (p >= 32) {
    a = h[1]; b = h[3]; c = h[5];
  } else {
    a = h[0]; b = h[2]; c = h[4];
  }
  return __builtin_amdgcn_ubfe(a, bit, 1u) | (__builtin_amdgcn_ubfe(b, bit, 1u) << 1) |
         (__builtin_amdgcn_ubfe(c, bit, 1u) << 2);
}

// One elimination step of the wave's basis (lane t holds h_t, act: still in
// the basis): the first active lane whose codeword is nonzero at point p
// becomes the pivot and leaves the basis; every other active lane subtracts
// the multiple of the pivot codeword that zeroes its own value at p.
// Returns false when no active codeword is nonzero at p (the condition is
// dependent on the earlier ones: not of full rank).
__device__ __forceinline__ bool cw_eliminate(uint32_t h[6], bool& act, int p, uint32_t lane) {
  const uint32_t e = cw_at(h, p);
  const uint64_t cand = __ballot(act && e != 0u);
  if (cand == 0ull) return false;
  const int t0 = __builtin_ctzll(cand);
  const uint32_t e0 = (uint32_t)__builtin_amdgcn_readlane((int)e, t0);
  // h_t(p) / h_t0(p): row 1 / e0 of the multiplication table (wave-uniform),
  // read at e.
  static constexpr uint32_t kRows[8] = {f8_mul_row(0), f8_mul_row(1), f8_mul_row(2),
                                        f8_mul_row(3), f8_mul_row(4), f8_mul_row(5),
                                        f8_mul_row(6), f8_mul_row(7)};
  const uint32_t row = kRows[(kF8Inv >> (3 * e0)) & 7u];
  const uint32_t c = __builtin_amdgcn_ubfe(row, 3u * e, 3u);
  const bool upd = act && lane != (uint32_t)t0;
  uint32_t H[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) H[q] = (uint32_t)__builtin_amdgcn_readlane((int)h[q], t0);
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const uint32_t msk = upd ? 0u - ((c >> b) & 1u) : 0u;
#pragma unroll
    for (int q = 0; q < 6; ++q) h[q] ^= H[q] & msk;
    mul_x(H);
  }
  act = upd;
  return true;
}

// Bit rows of one code at level M (one wave; lane i < k holds index g of its
// exponent point): subcode s (s < level_nsub(M)) of the representative R_s
// takes rows s * 3w + 3t + b = x^b * h_t (w = k - M, t < w), where h_0 ..
// h_{w-1} span the codewords of C(S) that vanish at the points r d, r in
// R_s, in the first direction d where every R_s gives M independent
// conditions.  Returns false when no direction does.
//
// The basis comes from Gaussian elimination on the codewords themselves:
// lane i starts with g_i and each point of the representative removes one
// pivot lane (cw_eliminate); the prefix {0, .., M-2} shared by the
// representatives is eliminated once per direction, then each
// representative's last point on a copy.  (M is a template argument so the
// prefix loop unrolls.)
template <int M>
__device__ __forceinline__ bool md_rows_level(int g, int k, uint32_t lane,
                                              const uint32_t* __restrict__ mplanes,
                                              uint32_t* __restrict__ base) {
  constexpr int nsub = level_nsub(M);
  const int w = k - M;
  uint32_t own[6];
  {
    const uint4 r0 = *reinterpret_cast<const uint4*>(mplanes + g * kRowWords);
    const uint2 r1 = *reinterpret_cast<const uint2*>(mplanes + g * kRowWords + 4);
    own[0] = r0.x; own[1] = r0.y; own[2] = r0.z; own[3] = r0.w; own[4] = r1.x; own[5] = r1.y;
  }
#pragma unroll 1
  for (int d = 0; d < 8; ++d) {
    uint32_t h[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) h[q] = own[q];
    bool act = lane < (uint32_t)k;
    bool ok = true;
#pragma unroll
    for (int r = 0; r < M - 1; ++r) ok = ok && cw_eliminate(h, act, line_point(d, r), lane);
    if (!ok) continue;
#pragma unroll 1
    for (int s = 0; s < nsub; ++s) {
      uint32_t hs[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) hs[q] = h[q];
      bool as = act;
      if (!cw_eliminate(hs, as, line_point(d, level_last(M, s)), lane)) {
        ok = false;
        break;
      }
      const uint64_t live = __ballot(as);
      if (as) {
        const int t = __builtin_popcountll(live & ((1ull << lane) - 1ull));
        uint32_t* out_rows = base + (size_t)(s * 3 * w + 3 * t) * kRowWords;
#pragma unroll
        for (int bp = 0; bp < 3; ++bp) {
          uint4* rr = reinterpret_cast<uint4*>(out_rows + bp * kRowWords);
          rr[0] = make_uint4(hs[0], hs[1], hs[2], hs[3]);
          rr[1] = make_uint4(hs[4], hs[5], 0u, 0u);
          mul_x(hs);
        }
      }
    }
    if (ok) return true;
  }
  return false;
}

// ---- one wave per code (k <= kFusedMaxK) -----------------------------------
// A code of at most 8 generators walks at most 45 rows (level 3: 3 subcodes
// of 3 x 5 rows; any other level fewer) and at most 3 (8^5 - 1) / 7
// codewords at its default level, so one wave builds its rows in LDS and
// walks them itself: no plan, no cross-wave fold, one launch.  A maximum z
// found below the level's certificate for found maxima, 7m - 7 (or a level
// without a direction of full rank), is followed, in the same wave, by the
// highest lower level whose certificate the bound max(z, k - 1) already
// meets or whose found certificate z meets (level 1 at worst: 21 rows,
// 299,593 codewords for k = 8).  Codes with more generators are appended to the list of the
// planned search (md_prepl, md_main).  (Dealing the walks of a workgroup's
// four codes over its four waves measured slower: 26.9 against 23.6 us on
// the 4096-code mixed batch, the dealing's scalar work outweighing the
// balance it bought.)
constexpr int kFusedRows = 45;   // 3 subcodes x 3 (kFusedMaxK - 3) rows
static_assert(kFusedRows == 3 * 3 * (kFusedMaxK - 3), "the level-3 rows of the largest fused code");
constexpr int kFusedWaves = 4;   // waves per workgroup of the one-wave search

// Rows of level m into `base` (md_rows_level); false: no full-rank direction.
__device__ __forceinline__ bool md_rows_switch(int m, int g, int k, uint32_t lane,
                                               const uint32_t* __restrict__ mplanes,
                                               uint32_t* __restrict__ base) {
  switch (m) {
    case 1: return md_rows_level<1>(g, k, lane, mplanes, base);
    case 2: return md_rows_level<2>(g, k, lane, mplanes, base);
    case 3: return md_rows_level<3>(g, k, lane, mplanes, base);
    case 4: return md_rows_level<4>(g, k, lane, mplanes, base);
    case 5: return md_rows_level<5>(g, k, lane, mplanes, base);
    default: return md_rows_level<6>(g, k, lane, mplanes, base);   // kMaxLevel
  }
}

// Per-lane minimum weight over a whole w-dimensional walk (rows in LDS).
__device__ __forceinline__ uint32_t walk_all(const LdsRows& row, int w, uint32_t lane) {
  uint32_t best = 64u;
  for (int j = w - 1; j > kTinyTop; --j)
    best = min(best, gray_blocks(row, (uint32_t)j, 0ull, piece_blocks(j), lane));
  return min(best, tiny_chunk(row, tiny_top(w), lane));
}

// LDS written by lanes of this wave and read by all of them next (and the
// other way round before the area is rewritten): wait for the wave's LDS
// operations and keep the compiler from moving loads or stores across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The whole exact search of one code of 2 <= k <= kFusedMaxK generators by
// one wave (lane i < k holds the code's index g of point i): the first
// level's rows built in `base` (kFusedRows rows of the wave's LDS), walked,
// certified (z >= 7m - 7); below that (or without a full-rank direction) a
// lower level, in the same wave, until a level certifies its maximum.
// Returns the minimum weight: exact, or <= 48 - tmax once the threshold is
// exceeded (the result is md_result of it); adds the projective codewords of
// the subcodes it walked to `walked`.
__device__ __forceinline__ uint32_t fused_search(int g, int k, int tmax, uint32_t lane,
                                                 const uint32_t* __restrict__ mplanes,
                                                 const uint8_t* __restrict__ level,
                                                 uint32_t* __restrict__ base, uint64_t& walked) {
  const int thr = 48 - tmax;
  int m = __builtin_amdgcn_readfirstlane(code_level(level, k, false));
  for (;;) {
    wave_lds_sync();
    const bool ok = md_rows_switch(m, g, k, lane, mplanes, base);
    wave_lds_sync();
    uint32_t best = 64u;
    if (ok) {
      const int w = k - m;
      for (int s = 0; s < level_nsub(m) && (int)best > thr; ++s) {
        best = wave_min(min(best, walk_all(LdsRows{base + (size_t)s * 3 * w * kRowWords}, w,
                                           lane)));
        walked += proj_codewords(w);
      }
    }
    const int z = kNpts - (int)best;
    if (m == 1 || (ok && (z > tmax || z >= found_cert(m)))) return best;
    // The highest lower level q that max(z, k - 1) certifies beforehand
    // (7q - 6), or whose found certificate 7q - 7 the walk's z already meets
    // (the level-q walk in the same direction finds at least z; its own
    // maximum is checked again).
    const int lb = max(ok ? z : 0, k - 1);
    int m2 = 1;
    for (int q = m - 1; q >= 2; --q)
      if (level_cert(q) <= lb || (ok && found_cert(q) <= z)) { m2 = q; break; }
    m = m2;
  }
}

}  // namespace gt
