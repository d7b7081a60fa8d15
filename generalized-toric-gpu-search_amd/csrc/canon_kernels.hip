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
//   image(A, t) = translate(mask(A S), -(A s_t)).
// Only images that contain point 0 are candidates (the anchored images); the
// minimum over all 98,784 affine images is a different function.  The number
// of (A, t) pairs attaining the minimum is |Stab_AGL2(S)|.
//
// Layout: one lane per set, A uniform across the wave (2016 iterations), the
// GL2 permutation table staged in LDS in tiles.  Sets are first bucketed by
// size so that every wave runs a loop specialised for its k.
#include "gt_launch.h"

namespace gt {

constexpr int kPermStride = 52;        // 49 bytes per matrix, padded to 13 dwords
constexpr int kPermTile = 504;         // matrices per LDS tile (4 tiles)
constexpr int kCanonBlock = 256;
constexpr uint64_t kMask49 = (1ull << 49) - 1ull;

int canon_perm_bytes() { return 2016 * kPermStride; }

// Bits whose position inside their 7-bit row is < 7 - v, plus every bit above
// 49 (so that v_bfi keeps those bits from the right-shifted operand, which has
// them clear).
__device__ __forceinline__ uint64_t keep_low(int v) {
  const uint64_t rowbits = (0x7Full >> v);
  uint64_t m = 0;
#pragma unroll
  for (int g = 0; g < 7; ++g) m |= rowbits << (7 * g);
  return m | ~kMask49;
}

// Translate a 49-bit mask on the 7x7 torus by (-u0, -v0).
__device__ __forceinline__ uint64_t translate(uint64_t M, uint32_t s7, uint32_t v0, uint64_t K) {
  const uint64_t R = ((M >> s7) | (M << (49u - s7))) & kMask49;
  const uint64_t a = R >> v0, b = R << (7u - v0);
  return (a & K) | (b & ~K);
}

template <int K>
__device__ __forceinline__ void canon_tile(const uint8_t* __restrict__ lperm, int na,
                                           const uint32_t* pts, const uint64_t* lkeep,
                                           uint64_t& best, uint32_t& cnt) {
  for (int a = 0; a < na; ++a) {
    const uint8_t* pa = lperm + a * kPermStride;
    uint32_t q[K];
    uint64_t MA = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      q[i] = pa[pts[i]];
      MA |= 1ull << q[i];
    }
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const uint32_t u0 = (q[t] * 37u) >> 8;      // q / 7 for q <= 48
      const uint32_t v0 = q[t] - 7u * u0;
      const uint64_t img = translate(MA, 7u * u0, v0, lkeep[v0]);
      if (img < best) { best = img; cnt = 1u; }
      else if (img == best) { ++cnt; }
    }
  }
}


// Counting sort of set indices by |S| (1..16); bad sizes are flagged.
__global__ void canon_hist_kernel(const uint64_t* __restrict__ masks, int32_t n,
                                  uint32_t* hist, uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t m = masks[i];
  const int k = __popcll(m);
  if (k < 1 || k > kMaxKCanon || (m >> 49) != 0) { atomicOr(&ctrl[2], kErrBadK); return; }
  atomicAdd(&hist[k], 1u);
}

__global__ void canon_scan_kernel(uint32_t* hist) {
  // hist[0..16] counts -> hist[17..33] running offsets (single thread; 17 bins).
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k <= kMaxKCanon; ++k) {
      hist[17 + k] = run;
      // Round each bucket up to whole waves so no wave mixes two sizes.
      run += (hist[k] + kWave - 1) / kWave * kWave;
    }
    hist[34] = run;
  }
}

__global__ void canon_scatter_kernel(const uint64_t* __restrict__ masks, int32_t n,
                                     uint32_t* hist, uint32_t* order) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t m = masks[i];
  const int k = __popcll(m);
  if (k < 1 || k > kMaxKCanon || (m >> 49) != 0) return;
  const uint32_t pos = atomicAdd(&hist[17 + k], 1u);
  order[pos] = (uint32_t)i;
}

__global__ void canon_fill_kernel(uint32_t* order, uint32_t len, uint32_t* hist) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) order[i] = 0xFFFFFFFFu;
  if (i < 35) hist[i] = 0u;
}

__global__ __launch_bounds__(kCanonBlock) void canon_kernel(
    const uint64_t* __restrict__ masks, const uint32_t* __restrict__ order, uint32_t slots,
    const uint8_t* __restrict__ gperm, uint64_t* __restrict__ key_lo,
    uint64_t* __restrict__ key_hi, uint64_t* __restrict__ canon, int32_t* __restrict__ stab) {
  __shared__ uint32_t lperm_words[kPermTile * kPermStride / 4];
  __shared__ uint64_t lkeep[8];
  uint8_t* lperm = reinterpret_cast<uint8_t*>(lperm_words);
  if (threadIdx.x < 7) lkeep[threadIdx.x] = keep_low(threadIdx.x);
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t idx = (slot < slots) ? order[slot] : 0xFFFFFFFFu;
  const bool active = idx != 0xFFFFFFFFu;
  const uint64_t mask = active ? masks[idx] : 0ull;
  // Buckets are wave aligned: the first active lane's size is the wave's size.
  int k = active ? __popcll(mask) : 0;
  const uint64_t act = __ballot(active);
  if (act != 0) {
    const int src = __builtin_ctzll(act);
    k = __shfl(k, src, kWave);
  }
  const int kw = __builtin_amdgcn_readfirstlane(k);
  uint32_t pts[kMaxKCanon];
  {
    uint64_t m = mask;
#pragma unroll
    for (int i = 0; i < kMaxKCanon; ++i) {
      const uint32_t p = (uint32_t)__builtin_ctzll(m | (1ull << 63));
      pts[i] = (p > 48) ? 0u : p;
      m &= m - 1;
    }
  }
  uint64_t best = ~0ull;
  uint32_t cnt = 0;
  // Every thread of the block loads its share of every tile, whatever its
  // wave's k (or none), before the wave runs the loop specialised for k.
  for (int t0 = 0; t0 < 2016; t0 += kPermTile) {
    __syncthreads();
    const uint32_t* src = reinterpret_cast<const uint32_t*>(gperm + t0 * kPermStride);
    for (int e = threadIdx.x; e < kPermTile * kPermStride / 4; e += blockDim.x)
      lperm_words[e] = src[e];
    __syncthreads();
    if (!active) continue;
    switch (kw) {
#define GT_CANON_CASE(KK) \
  case KK: canon_tile<KK>(lperm, kPermTile, pts, lkeep, best, cnt); break;
      GT_CANON_CASE(1) GT_CANON_CASE(2) GT_CANON_CASE(3) GT_CANON_CASE(4)
      GT_CANON_CASE(5) GT_CANON_CASE(6) GT_CANON_CASE(7) GT_CANON_CASE(8)
      GT_CANON_CASE(9) GT_CANON_CASE(10) GT_CANON_CASE(11) GT_CANON_CASE(12)
      GT_CANON_CASE(13) GT_CANON_CASE(14) GT_CANON_CASE(15) GT_CANON_CASE(16)
#undef GT_CANON_CASE
      default: break;
    }
  }
  if (!active) return;
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
hipError_t launch_canon(const uint64_t* d_masks, int32_t n, const uint8_t* d_perm,
                        uint32_t* d_order, uint32_t* d_hist, uint64_t* d_key_lo,
                        uint64_t* d_key_hi, uint64_t* d_canon, int32_t* d_stab,
                        uint32_t* d_ctrl, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const uint32_t slots = (uint32_t)n + (uint32_t)(kMaxKCanon + 1) * kWave;
  hipLaunchKernelGGL(canon_fill_kernel, dim3((slots + 255) / 256), dim3(256), 0, s, d_order,
                     slots, d_hist);
  hipLaunchKernelGGL(canon_hist_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_masks, n,
                     d_hist, d_ctrl);
  hipLaunchKernelGGL(canon_scan_kernel, dim3(1), dim3(64), 0, s, d_hist);
  hipLaunchKernelGGL(canon_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_masks, n,
                     d_hist, d_order);
  hipLaunchKernelGGL(canon_kernel, dim3((slots + kCanonBlock - 1) / kCanonBlock),
                     dim3(kCanonBlock), 0, s, d_masks, d_order, slots, d_perm, d_key_lo,
                     d_key_hi, d_canon, d_stab);
  return hipGetLastError();
}

}  // namespace gt
