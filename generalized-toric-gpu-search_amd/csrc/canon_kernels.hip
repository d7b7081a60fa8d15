// AGL2(F7) canonical form of exponent sets S in (Z/7)^2 on gfx950.
//
// Replaces canonical_forms (kernel_canon.py:44-129) and the CPU path
// canonical_forms_batch (champion_search_canonical.py:115-140).
//
// Definition kept from the reference: the key is the minimum, over the 2016
// matrices A of GL2(F7) and the k anchors t of S, of the sorted tuple
// rel(A s_t, A s_i) = A (s_i - s_t) packed 6 bits per position.  For sets of
// one size the packed order equals the order of the 49-bit image masks
// (both compare the largest element first), so the kernel minimises masks:
//   image(A, t) = translate(mask(A S), -(A s_t)),
// held in a byte-row layout (below) that makes the translation cheap.
// Only images that contain point 0 are candidates (the anchored images); the
// minimum over all 98,784 affine images is a different function.  The number
// of (A, t) pairs attaining the minimum is |Stab_AGL2(S)|.
//
// Search (one wavefront per set).  Write A = (r1; r2): r1 = (a, b) gives the
// row u = r1.p of an image point, r2 = (c, d) its column v.  The highest bit of
// an image is in its top occupied row, and after translating by anchor t the
// rows are u_i - u_t (mod 7), so the top row is 6 - (the run of empty rows
// just below row u_t), which depends on r1 and t only.  Phase 1: lane r1
// (48 nonzero first rows) finds the best top row its r1 allows (6 - the
// longest cyclic run of empty rows of its row-occupancy set); the wave minimum
// U* is the top row of the canonical image.  Phase 2: for every first row
// reaching U* (typically a handful of the 48), lane r2 builds the image mask
// for A = (r1; r2) (42 valid second rows) and translates it by every anchor
// whose row starts the longest empty run; those are the only (A, t) pairs
// whose image can be minimal, so both the minimum and the count of pairs that
// attain it (|Stab|) are exact.  A wave reduction combines the lanes.  Sets are
// bucketed by size, largest first, and walked by the waves with a fixed stride.
#include <algorithm>

#include "gt_launch.h"

namespace gt {

constexpr int kCanonBlock = 256;
constexpr int kCanonMaxBlocks = 2048;  // persistent grid: 8 waves per SIMD on 256 CUs
constexpr int kRows = 48;              // nonzero vectors (a, b) of F7^2, index 7a + b - 1
constexpr int kRowStride = 52;

// Byte-row layout.  Inside the kernel a subset of the 7x7 torus is a 56-bit
// value with row u in byte u and column v in bit v of that byte (position
// 8u + v; bit 7 of every byte and byte 7 stay 0).  Integer order on this
// layout equals integer order on the 49-bit mask (position 7u + v), because
// both orders compare row 6 first and, inside a row, column 6 first.
//
// Translating by (-u, -v) is then a byte rotation of the 7 rows (two
// v_perm_b32, whose selectors are the row numbers (r + u) mod 7 and 12 for a
// zero byte) and a rotation of every 7-bit row by v: v_bfi(lo, x >> v,
// x << (7 - v)) with lo = bits 0..6-v of every row, then an AND that clears
// bit 7 of every row and byte 7, where the left shift leaves stray bits.
// One 16-byte entry per anchor point (one ds_read_b128 per image).
struct Anchor {
  uint32_t sel_lo, sel_hi;   // v_perm selectors for rows 0..3 and 4..6
  uint32_t vv;               // v | (7 - v) << 8: the 64-bit shifts read bits 0..5
  uint32_t lo;               // bits 0..6-v of every byte
};

constexpr uint32_t kRowBitsLo = 0x7F7F7F7Fu;   // rows 0..3
constexpr uint32_t kRowBitsHi = 0x007F7F7Fu;   // rows 4..6

__device__ __forceinline__ uint64_t translate(uint64_t M, const Anchor& e) {
  const uint32_t m0 = (uint32_t)M, m1 = (uint32_t)(M >> 32);
  const uint64_t X = ((uint64_t)__builtin_amdgcn_perm(m1, m0, e.sel_hi) << 32) |
                     __builtin_amdgcn_perm(m1, m0, e.sel_lo);
  const uint64_t a = X >> (e.vv & 63u), b = X << ((e.vv >> 8) & 63u);
  // v_bitop3 truth table 0xE4: (a & lo) | (b & ~lo).
  const uint32_t r0 = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, e.lo, 0xE4) &
                      kRowBitsLo;
  const uint32_t r1 = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32), e.lo,
                                                  0xE4) & kRowBitsHi;
  return ((uint64_t)r1 << 32) | r0;
}

__host__ __device__ __forceinline__ Anchor make_anchor(uint32_t u, uint32_t v) {
  Anchor e{};
  uint32_t sel[8];
  for (uint32_t r = 0; r < 7; ++r) sel[r] = (r + u) % 7u;
  sel[7] = 12u;
  e.sel_lo = sel[0] | (sel[1] << 8) | (sel[2] << 16) | (sel[3] << 24);
  e.sel_hi = sel[4] | (sel[5] << 8) | (sel[6] << 16) | (sel[7] << 24);
  e.vv = v | ((7u - v) << 8);
  e.lo = (0x7Fu >> v) * 0x01010101u;
  return e;
}

// Tables of one workgroup (4.5 KB of LDS).
struct CanonTables {
  uint8_t rowval[kRows * kRowStride];  // rowval[r][p] = (a x + b y) mod 7, r = 7a + b - 1, p = 7x + y
  uint64_t par[kRows];                 // bit r2 set when vector r2 is parallel to vector r
  uint8_t gap[128 * 8];                // gap[R][u]: empty rows just below row u (cyclic) in the
                                       // row set R; gap[R][7]: the longest such run over u in R
  Anchor anc[56];                      // translation by byte-row position 8u + v
};

__device__ void build_canon_tables(CanonTables& T) {
  for (int e = threadIdx.x; e < kRows * kRowStride; e += blockDim.x) {
    const int r = e / kRowStride, p = e % kRowStride;
    const int a = (r + 1) / 7, b = (r + 1) % 7;
    T.rowval[e] = (uint8_t)((p < kNpts) ? (a * (p / 7) + b * (p % 7)) % 7 : 0);
  }
  for (int r = threadIdx.x; r < kRows; r += blockDim.x) {
    const int a = (r + 1) / 7, b = (r + 1) % 7;
    uint64_t m = 0;
    for (int r2 = 0; r2 < kRows; ++r2) {
      const int c = (r2 + 1) / 7, d = (r2 + 1) % 7;
      if ((a * d - b * c + 49) % 7 == 0) m |= 1ull << r2;
    }
    T.par[r] = m;
  }
  for (int R = threadIdx.x; R < 128; R += blockDim.x) {
    int longest = 0;
    for (int u = 0; u < 7; ++u) {
      int g = 0;
      while (g < 6 && !((R >> ((u + 6 - g) % 7)) & 1)) ++g;
      T.gap[R * 8 + u] = (uint8_t)g;
      if ((R >> u) & 1) longest = max(longest, g);
    }
    T.gap[R * 8 + 7] = (uint8_t)longest;
  }
  if (threadIdx.x < 56) {
    const uint32_t u = threadIdx.x >> 3, v = threadIdx.x & 7u;
    T.anc[threadIdx.x] = (u < 7 && v < 7) ? make_anchor(u, v) : Anchor{};
  }
}

// Minimum byte-row image and the number of (A, t) attaining it, for the set
// with points p[0..K-1] (wave-uniform); every lane returns the same pair.
// par_mine = T.par[lane] (the relation is symmetric: bit r1 says whether this
// lane's vector is parallel to r1).
template <int K>
__device__ __forceinline__ void canon_set(const CanonTables& T, const uint32_t* p, uint32_t lane,
                                          uint64_t par_mine, uint64_t& best_out,
                                          uint32_t& cnt_out) {
  // Phase 1, lane r1 < 48: row values w_i = r1.p_i, their row set, the best top
  // row r1 allows, and the anchors whose row starts the longest empty run.
  // The same w_i are the column values z_i when the lane stands for r2.
  const bool row_lane = lane < (uint32_t)kRows;
  const uint32_t base = (row_lane ? lane : 0u) * kRowStride;
  uint32_t wl[K];
  uint32_t R = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    wl[i] = T.rowval[base + p[i]];
    R |= 1u << wl[i];
  }
  const uint32_t longest = T.gap[R * 8 + 7];
  uint32_t amask = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) amask |= (uint32_t)(T.gap[R * 8 + wl[t]] == longest) << t;
  const uint32_t mytop = row_lane ? 6u - longest : 7u;
  uint32_t top = mytop;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) top = min(top, (uint32_t)__shfl_xor((int)top, off, kWave));
  uint64_t cand = __ballot(mytop == top);
  // Phase 2, lane r2 < 48 (not parallel to r1), for every r1 reaching the top
  // row: image of A = (r1; r2), translated by each candidate anchor.
  uint64_t best = ~0ull;
  uint32_t cnt = 0;
  while (cand != 0) {
    const int r1 = __builtin_ctzll(cand);
    cand &= cand - 1;
    const uint32_t am = __builtin_amdgcn_readlane(amask, r1);
    uint32_t w[K];
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = __builtin_amdgcn_readlane(wl[i], r1);
    if (row_lane && !((par_mine >> r1) & 1ull)) {
      uint64_t MA = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) MA |= 1ull << (8u * w[i] + wl[i]);
#pragma unroll
      for (int t = 0; t < K; ++t) {
        if ((am >> t) & 1u) {
          const uint64_t img = translate(MA, T.anc[8u * w[t] + wl[t]]);
          const bool lt = img < best;
          cnt = lt ? 1u : cnt + (uint32_t)(img == best);
          best = lt ? img : best;
        }
      }
    }
  }
  // Combine the lanes: minimum, and the counts of the lanes holding it.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t ob = __shfl_xor(best, off, kWave);
    const uint32_t oc = (uint32_t)__shfl_xor((int)cnt, off, kWave);
    cnt = (ob < best) ? oc : cnt + ((ob == best) ? oc : 0u);
    best = (ob < best) ? ob : best;
  }
  best_out = best;
  cnt_out = cnt;
}

// Byte-row layout -> 49-bit mask.
__device__ __forceinline__ uint64_t rows_to_mask(uint64_t x) {
  uint64_t m = 0;
#pragma unroll
  for (int u = 0; u < 7; ++u) m |= ((x >> (8 * u)) & 0x7Full) << (7 * u);
  return m;
}

// Counting sort of set indices by |S| (1..16); bad sizes are flagged.  Each
// workgroup counts its grid-stride share in LDS and adds one global atomic per
// size: a global atomic per set would queue a million atomics on 17 addresses
// (about 3 ms per pass for 10^6 sets).
constexpr int kSortBlock = 256;
constexpr int kSortGrid = 512;

__global__ __launch_bounds__(kSortBlock) void canon_hist_kernel(const uint64_t* __restrict__ masks,
                                                                int32_t n, uint32_t* hist,
                                                                uint32_t* ctrl) {
  __shared__ uint32_t lh[kMaxKCanon + 1];
  if (threadIdx.x <= kMaxKCanon) lh[threadIdx.x] = 0u;
  __syncthreads();
  bool bad = false;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t m = masks[i];
    const int k = __popcll(m);
    if (k < 1 || k > kMaxKCanon || (m >> 49) != 0) { bad = true; continue; }
    atomicAdd(&lh[k], 1u);
  }
  if (bad) atomicOr(&ctrl[2], kErrBadK);
  __syncthreads();
  if (threadIdx.x <= kMaxKCanon && lh[threadIdx.x] != 0u) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

__global__ void canon_scan_kernel(uint32_t* hist) {
  // hist[0..16] counts -> hist[17..33] running offsets (single thread; 17 bins).
  // Largest sets first: their waves run longest, so starting them first keeps
  // the end of the launch to short waves.
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t run = 0;
    for (int k = kMaxKCanon; k >= 0; --k) {
      hist[17 + k] = run;
      // Round each bucket up to whole waves so no wave mixes two sizes.
      run += (hist[k] + kWave - 1) / kWave * kWave;
    }
    hist[34] = run;
  }
}

// Same grid and the same grid-stride share as canon_hist_kernel: each
// workgroup counts again, reserves one range per size with one global atomic,
// then places its sets inside its ranges (LDS atomics; the order inside a
// bucket does not change any result).
__global__ __launch_bounds__(kSortBlock) void canon_scatter_kernel(
    const uint64_t* __restrict__ masks, int32_t n, uint32_t* hist, uint32_t* order) {
  __shared__ uint32_t lh[kMaxKCanon + 1];
  __shared__ uint32_t base[kMaxKCanon + 1];
  if (threadIdx.x <= kMaxKCanon) lh[threadIdx.x] = 0u;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t m = masks[i];
    const int k = __popcll(m);
    if (k < 1 || k > kMaxKCanon || (m >> 49) != 0) continue;
    atomicAdd(&lh[k], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= kMaxKCanon) {
    const uint32_t c = lh[threadIdx.x];
    base[threadIdx.x] = (c != 0u) ? atomicAdd(&hist[17 + threadIdx.x], c) : 0u;
    lh[threadIdx.x] = 0u;
  }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t m = masks[i];
    const int k = __popcll(m);
    if (k < 1 || k > kMaxKCanon || (m >> 49) != 0) continue;
    order[base[k] + atomicAdd(&lh[k], 1u)] = (uint32_t)i;
  }
}

__global__ void canon_fill_kernel(uint32_t* order, uint32_t len, uint32_t* hist) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) order[i] = 0xFFFFFFFFu;
  if (i < 35) hist[i] = 0u;
}

__global__ __launch_bounds__(kCanonBlock) void canon_kernel(
    const uint64_t* __restrict__ masks, const uint32_t* __restrict__ order, uint32_t slots,
    uint64_t* __restrict__ key_lo, uint64_t* __restrict__ key_hi, uint64_t* __restrict__ canon,
    int32_t* __restrict__ stab) {
  __shared__ CanonTables T;
  build_canon_tables(T);
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint64_t par_mine = T.par[lane < (uint32_t)kRows ? lane : 0u];
  const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
  for (uint32_t j = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; j < slots;
       j += nwaves) {
    const uint32_t idx = __builtin_amdgcn_readfirstlane(order[j]);
    if (idx == 0xFFFFFFFFu) continue;
    const uint64_t mask = masks[idx];
    const int k = __popcll(mask);
    uint32_t pts[kMaxKCanon];
    {
      uint64_t m = mask;
#pragma unroll
      for (int i = 0; i < kMaxKCanon; ++i) {
        const uint32_t q = (uint32_t)__builtin_ctzll(m | (1ull << 63));
        pts[i] = (q > 48) ? 0u : q;
        m &= m - 1;
      }
    }
    uint64_t best = ~0ull;
    uint32_t cnt = 0;
    switch (k) {
#define GT_CANON_CASE(KK) \
  case KK: canon_set<KK>(T, pts, lane, par_mine, best, cnt); break;
      GT_CANON_CASE(1) GT_CANON_CASE(2) GT_CANON_CASE(3) GT_CANON_CASE(4)
      GT_CANON_CASE(5) GT_CANON_CASE(6) GT_CANON_CASE(7) GT_CANON_CASE(8)
      GT_CANON_CASE(9) GT_CANON_CASE(10) GT_CANON_CASE(11) GT_CANON_CASE(12)
      GT_CANON_CASE(13) GT_CANON_CASE(14) GT_CANON_CASE(15) GT_CANON_CASE(16)
#undef GT_CANON_CASE
      default: continue;   // sizes outside 1..16 never reach the order array
    }
    if (lane != 0) continue;
    best = rows_to_mask(best);
    // Sorted positions of the canonical image, packed as kernel_canon.py:84-104:
    // k <= 10 packs positions 0..k-1, k >= 11 skips position 0 (always 0).
    const int pack_from = (k <= 10) ? 0 : 1;
    uint64_t lo = 0, hi = 0, m = best;
    for (int i = 0; i < k; ++i) {
      const uint64_t v = (uint64_t)__builtin_ctzll(m);
      m &= m - 1;
      if (i < pack_from) continue;
      const int off = 6 * (i - pack_from);
      if (off + 6 <= 64) lo |= v << off;
      else if (off < 64) { lo |= v << off; hi |= v >> (64 - off); }
      else hi |= v << (off - 64);
    }
    key_lo[idx] = lo;
    key_hi[idx] = hi;
    canon[idx] = best;
    stab[idx] = (int32_t)cnt;
  }
}

// (n, k) sorted index rows -> masks; a repeated or out-of-range index is flagged.
__global__ void sets_to_masks_kernel(const int32_t* __restrict__ sets, int32_t n, int32_t k,
                                     uint64_t* __restrict__ masks, uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t m = 0;
  bool bad = false;
  for (int j = 0; j < k; ++j) {
    const int s = sets[(size_t)i * k + j];
    if (s < 0 || s >= kNpts || ((m >> s) & 1ull)) { bad = true; continue; }
    m |= 1ull << s;
  }
  if (bad) { atomicOr(&ctrl[2], kErrBadIndex); m = 0; }
  masks[i] = m;
}

hipError_t launch_sets_to_masks(const int32_t* d_sets, int32_t n, int32_t k, uint64_t* d_masks,
                                uint32_t* d_ctrl, hipStream_t s) {
  hipLaunchKernelGGL(sets_to_masks_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_sets, n, k,
                     d_masks, d_ctrl);
  return hipGetLastError();
}

// d_order must hold n + 16 * 64 entries; d_hist 35 entries.
hipError_t launch_canon(const uint64_t* d_masks, int32_t n,
                        uint32_t* d_order, uint32_t* d_hist, uint64_t* d_key_lo,
                        uint64_t* d_key_hi, uint64_t* d_canon, int32_t* d_stab,
                        uint32_t* d_ctrl, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const uint32_t slots = (uint32_t)n + (uint32_t)(kMaxKCanon + 1) * kWave;
  hipLaunchKernelGGL(canon_fill_kernel, dim3((slots + 255) / 256), dim3(256), 0, s, d_order,
                     slots, d_hist);
  const int sort_grid = std::min(kSortGrid, (n + kSortBlock - 1) / kSortBlock);
  hipLaunchKernelGGL(canon_hist_kernel, dim3(sort_grid), dim3(kSortBlock), 0, s, d_masks, n,
                     d_hist, d_ctrl);
  hipLaunchKernelGGL(canon_scan_kernel, dim3(1), dim3(64), 0, s, d_hist);
  hipLaunchKernelGGL(canon_scatter_kernel, dim3(sort_grid), dim3(kSortBlock), 0, s, d_masks, n,
                     d_hist, d_order);
  const uint32_t waves = (uint32_t)n;
  const int grid = (int)std::min<uint32_t>(kCanonMaxBlocks,
                                           (waves + kCanonBlock / kWave - 1) / (kCanonBlock / kWave));
  hipLaunchKernelGGL(canon_kernel, dim3(grid), dim3(kCanonBlock), 0, s, d_masks, d_order, slots,
                     d_key_lo, d_key_hi, d_canon, d_stab);
  return hipGetLastError();
}

}  // namespace gt
