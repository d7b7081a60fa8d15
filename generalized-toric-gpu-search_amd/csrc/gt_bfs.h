// Device-side breadth-first search helpers (expansion, hash-set dedup,
// ordered compaction, mask -> index rows).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"

namespace gt {

#if defined(__HIPCC__)
// Slot hash of the open-addressing hash sets (the murmur3 64-bit finaliser).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
#endif

hipError_t launch_bfs_expand(const uint64_t* d_par, int32_t n, int32_t kp, uint64_t* d_child,
                             uint32_t* d_ctrl, hipStream_t s);
hipError_t launch_hs_clear(uint64_t* table, uint64_t* tags, uint64_t cap, uint32_t* count,
                           hipStream_t s);
hipError_t launch_hs_dedup(const uint64_t* d_keys, int32_t n, uint64_t* table, uint64_t* tags,
                           uint64_t cap, uint64_t call, uint32_t* count, uint32_t* d_flag,
                           uint32_t* d_ctrl, hipStream_t s);
// Insert only (tags as for launch_hs_dedup; call 0 gives tag = index).
hipError_t launch_hs_insert(const uint64_t* d_keys, int32_t n, uint64_t* table, uint64_t* tags,
                           uint64_t cap, uint64_t call, uint32_t* count, uint32_t* d_ctrl,
                           hipStream_t s);
hipError_t launch_compact(const uint32_t* d_flag, int32_t n, uint32_t* d_bsum, uint32_t* d_out,
                          uint32_t* d_total, hipStream_t s);
hipError_t launch_gather_u64(const uint64_t* src, const uint32_t* idx, const uint32_t* count,
                             uint64_t* dst, int32_t cap, hipStream_t s);
hipError_t launch_masks_to_rows(const uint64_t* d_masks, int32_t n, int32_t k, int32_t* d_idx,
                                int32_t* d_offsets, uint32_t* d_ctrl, hipStream_t s);

}  // namespace gt
