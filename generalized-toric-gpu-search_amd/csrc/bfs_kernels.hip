// One level of the canonical breadth-first search on the device: expand
// parents by one point, deduplicate canonical forms against everything seen
// (open-addressing hash set in HBM), compact the new classes in input order,
// and turn masks into index rows for the min-distance path.
//
// Replaces the host loops of champion_search_canonical.py:496-506 (expansion)
// and :308-345 (first-occurrence dedup of canonical keys).
#include "gt_bfs.h"

namespace gt {

// Children of parent i are parent | bit p for the points p outside it, in
// increasing p, at positions i * (49 - kp) + rank(p).
__global__ void bfs_expand_kernel(const uint64_t* __restrict__ parents, int32_t n, int32_t kp,
                                  uint64_t* __restrict__ children, uint32_t* ctrl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int per = kNpts - kp;
  if (t >= (int64_t)n * per) return;
  const int i = (int)(t / per), r = (int)(t % per);
  const uint64_t m = parents[i];
  if (__popcll(m) != kp || (m >> 49) != 0) {
    atomicOr(&ctrl[2], kErrBadK);
    children[t] = 0;
    return;
  }
  // The r-th clear bit of m (0-based) among bits 0..48.
  uint64_t free = ~m & ((1ull << 49) - 1ull);
  for (int s = 0; s < r; ++s) free &= free - 1;
  children[t] = m | (1ull << __builtin_ctzll(free));
}

// Insert keys; tag = (call << 32) | index, and each slot keeps the smallest
// tag, so the first occurrence of a key (earliest call, then lowest index)
// owns it.  Keys must be nonzero (0 marks an empty slot).
__global__ void hs_insert_kernel(const uint64_t* __restrict__ keys, int32_t n,
                                 uint64_t* __restrict__ table, uint64_t* __restrict__ tags,
                                 uint64_t mask, uint64_t call, uint32_t* count,
                                 uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  if (key == 0) { atomicOr(&ctrl[2], kErrBadK); return; }
  uint64_t h = mix64(key) & mask;
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const uint64_t cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == key) break;
    if (cur == 0) {
      const uint64_t prev = atomicCAS((unsigned long long*)&table[h], 0ull,
                                      (unsigned long long)key);
      if (prev == 0) { atomicAdd(count, 1u); break; }
      if (prev == key) break;
    }
    h = (h + 1) & mask;
    if (probe == mask) { atomicOr(&ctrl[2], kErrFull); return; }
  }
  atomicMin((unsigned long long*)&tags[h], (unsigned long long)((call << 32) | (uint64_t)i));
}

// flag[i] = 1 when key i is the first occurrence of its key in this call and
// was not present before it.
__global__ void hs_flag_kernel(const uint64_t* __restrict__ keys, int32_t n,
                               const uint64_t* __restrict__ table,
                               const uint64_t* __restrict__ tags, uint64_t mask, uint64_t call,
                               uint32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = keys[i];
  uint32_t f = 0;
  if (key != 0) {
    uint64_t h = mix64(key) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
      const uint64_t cur = table[h];
      if (cur == key) { f = (tags[h] == ((call << 32) | (uint64_t)i)) ? 1u : 0u; break; }
      if (cur == 0) break;
      h = (h + 1) & mask;
    }
  }
  flag[i] = f;
}

__global__ void hs_clear_kernel(uint64_t* table, uint64_t* tags, uint64_t cap, uint32_t* count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) { table[i] = 0ull; tags[i] = ~0ull; }
  if (i == 0) *count = 0u;
}

// Ordered compaction: per-block counts, one-block scan of the counts, scatter.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void compact_count_kernel(const uint32_t* __restrict__ flag,
                                                                   int32_t n,
                                                                   uint32_t* __restrict__ bsum) {
  __shared__ uint32_t s[kScanBlock / kWave];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t f = (i < n) ? (flag[i] != 0u) : 0u;
  const uint64_t b = __ballot(f);
  if ((threadIdx.x & (kWave - 1)) == 0) s[threadIdx.x / kWave] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) t += s[w];
    bsum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void compact_scan_kernel(uint32_t* bsum, int32_t nb,
                                                            uint32_t* total) {
  __shared__ uint32_t s[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b = min(nb, tid * per), e = min(nb, b + per);
  uint32_t sum = 0;
  for (int i = b; i < e; ++i) sum += bsum[i];
  s[tid] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t t = (tid >= off) ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  uint32_t run = s[tid] - sum;
  for (int i = b; i < e; ++i) {
    const uint32_t c = bsum[i];
    bsum[i] = run;
    run += c;
  }
  if (tid == 1023) *total = s[1023];
}

__global__ __launch_bounds__(kScanBlock) void compact_scatter_kernel(
    const uint32_t* __restrict__ flag, int32_t n, const uint32_t* __restrict__ bsum,
    uint32_t* __restrict__ out_idx) {
  __shared__ uint32_t s[kScanBlock / kWave];
  const int i = blockIdx.x * kScanBlock + threadIdx.x;
  const uint32_t f = (i < n) ? (flag[i] != 0u) : 0u;
  const uint64_t b = __ballot(f);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t below = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) s[threadIdx.x / kWave] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t off = bsum[blockIdx.x];
  for (int w = 0; w < (int)(threadIdx.x / kWave); ++w) off += s[w];
  if (f) out_idx[off + below] = (uint32_t)i;
}

__global__ void gather_u64_kernel(const uint64_t* __restrict__ src,
                                  const uint32_t* __restrict__ idx, const uint32_t* count,
                                  uint64_t* __restrict__ dst, int32_t cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap || (uint32_t)i >= *count) return;
  dst[i] = src[idx[i]];
}

// Masks of size k -> sorted index rows and offsets (i * k).
__global__ void masks_to_rows_kernel(const uint64_t* __restrict__ masks, int32_t n, int32_t k,
                                     int32_t* __restrict__ idx, int32_t* __restrict__ offsets,
                                     uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  offsets[i] = i * k;
  if (i == n) return;
  uint64_t m = masks[i];
  if (__popcll(m) != k || (m >> 49) != 0) { atomicOr(&ctrl[2], kErrBadK); m = 0; }
  for (int j = 0; j < k; ++j) {
    idx[(size_t)i * k + j] = m ? __builtin_ctzll(m) : 0;
    if (m) m &= m - 1;
  }
}

// ------------------------------------------------------------- launchers --

hipError_t launch_bfs_expand(const uint64_t* d_par, int32_t n, int32_t kp, uint64_t* d_child,
                             uint32_t* d_ctrl, hipStream_t s) {
  const int64_t tot = (int64_t)n * (kNpts - kp);
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(bfs_expand_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     d_par, n, kp, d_child, d_ctrl);
  return hipGetLastError();
}

hipError_t launch_hs_clear(uint64_t* table, uint64_t* tags, uint64_t cap, uint32_t* count,
                           hipStream_t s) {
  hipLaunchKernelGGL(hs_clear_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, table,
                     tags, cap, count);
  return hipGetLastError();
}

hipError_t launch_hs_dedup(const uint64_t* d_keys, int32_t n, uint64_t* table, uint64_t* tags,
                           uint64_t cap, uint64_t call, uint32_t* count, uint32_t* d_flag,
                           uint32_t* d_ctrl, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(hs_insert_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_keys, n, table,
                     tags, cap - 1, call, count, d_ctrl);
  hipLaunchKernelGGL(hs_flag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_keys, n, table,
                     tags, cap - 1, call, d_flag);
  return hipGetLastError();
}

hipError_t launch_hs_insert(const uint64_t* d_keys, int32_t n, uint64_t* table, uint64_t* tags,
                           uint64_t cap, uint64_t call, uint32_t* count, uint32_t* d_ctrl,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(hs_insert_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_keys, n, table,
                     tags, cap - 1, call, count, d_ctrl);
  return hipGetLastError();
}

hipError_t launch_compact(const uint32_t* d_flag, int32_t n, uint32_t* d_bsum, uint32_t* d_out,
                          uint32_t* d_total, hipStream_t s) {
  const int nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(compact_count_kernel, dim3(nb), dim3(kScanBlock), 0, s, d_flag, n, d_bsum);
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, s, d_bsum, nb, d_total);
  hipLaunchKernelGGL(compact_scatter_kernel, dim3(nb), dim3(kScanBlock), 0, s, d_flag, n, d_bsum,
                     d_out);
  return hipGetLastError();
}

hipError_t launch_gather_u64(const uint64_t* src, const uint32_t* idx, const uint32_t* count,
                             uint64_t* dst, int32_t cap, hipStream_t s) {
  if (cap <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_u64_kernel, dim3((cap + 255) / 256), dim3(256), 0, s, src, idx, count,
                     dst, cap);
  return hipGetLastError();
}

hipError_t launch_masks_to_rows(const uint64_t* d_masks, int32_t n, int32_t k, int32_t* d_idx,
                                int32_t* d_offsets, uint32_t* d_ctrl, hipStream_t s) {
  hipLaunchKernelGGL(masks_to_rows_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, s, d_masks,
                     n, k, d_idx, d_offsets, d_ctrl);
  return hipGetLastError();
}

}  // namespace gt
