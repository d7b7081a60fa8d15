// Host-side launch wrappers for the gfx950 kernels (one translation unit per
// kernel family; the C-ABI in gt_api.hip calls only these).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"

namespace gt {

// Record `msg` as the calling thread's gt_last_error() message; returns -1.
int set_error(const char* msg);

// ---- min distance -------------------------------------------------------
// ctrl layout (uint32): two slots of kCtrlSlotWords, each [0] chunk counter,
// [1] unused, [2] error bits, [3] total chunks of the last plan; slot 0 at
// word 0 (device-pointer calls), slot 1 at word kCtrlHostSlot (host-pointer
// calls).
constexpr int kCtrlSlotWords = 8;
constexpr int kCtrlHostSlot = 8;
constexpr int kCtrlWords = 16;

hipError_t launch_md_rows(const int32_t* d_idx, int64_t n_idx, const uint32_t* d_mplanes,
                          uint32_t* d_rows, uint32_t* d_ctrl, hipStream_t s);
// Per-call words of the exact search.  big_count / big_list: the codes
// md_fused leaves to the list search (k > kfused), reset by the first-pass
// md_main; aux[code] = 1 when a list code's pass found no direction of
// full rank; mid_count / mid_list and fb_count / fb_list: the codes handed
// to the middle and the final pass, both reset by the first-pass md_prepl;
// walked (may be null): per code, the projective codewords of the subcodes
// its passes walked (zeroed by md_order, written by md_fused, added to by
// each md_main pass that finishes the code).
struct MdAux {
  uint32_t* aux;
  uint32_t* fb_count;
  uint32_t* fb_list;
  uint32_t* big_count;
  uint32_t* big_list;
  uint32_t* mid_count;
  uint32_t* mid_list;
  uint64_t* walked;
};
// Exact (or thresholded) search of one batch: md_order (sizes checked,
// results of codes without a walk, larger codes listed, the others sorted
// by size into d_order) and md_fused (a wave per code of k <= kfused, rows
// in LDS, results written), then,
// unless the caller knows that no code has more than kfused generators, the
// list search: md_prepl and md_main at the codes' levels (kPassFirst), then
// over the codes the first could not certify, one level lower (kPassMid,
// when some k of the batch has a middle level) and at second_level(k)
// (kPassFinal).
// d_rows: kRowsPerIdx row slots per index.
// d_claims: kClaimShards counters, kClaimStride words apart (reset by the plan).
// d_order, d_roff: n entries each (code index and first index per plan position).
// d_state: one word per code (minimum weight, chunks finished), initialised
// by md_prepl for its listed codes; md_main writes a listed code's result
// when its last chunk is folded, or lists it for a later pass.
hipError_t launch_md_fused(const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets,
                           int32_t n, const uint32_t* d_mplanes, const MdParams* d_params,
                           const int32_t* d_tmax, int32_t tmax, int32_t* d_out, const MdAux& ax,
                           uint32_t* d_order, uint32_t* d_fused_count, int kfused, int kmax,
                           uint32_t* d_ctrl, hipStream_t s, hipEvent_t ev_start = nullptr,
                           hipEvent_t ev_stop = nullptr);
hipError_t launch_md_prepl(int pass, const int32_t* d_idx, int64_t n_idx,
                           const uint32_t* d_mplanes, uint32_t* d_rows, const int32_t* d_offsets,
                           int32_t n, const MdParams* d_params, MdPlan* d_plan, uint32_t* d_order,
                           uint32_t* d_roff, uint32_t* d_claims, uint64_t* d_state,
                           int32_t* d_out, const int32_t* d_tmax, int32_t tmax, const MdAux& ax,
                           uint32_t* d_ctrl, hipStream_t s);
hipError_t launch_md_main(bool thresholded, int grid, const uint32_t* d_rows,
                          const MdPlan* d_plan, const uint32_t* d_order, const uint32_t* d_roff,
                          uint32_t* d_claims, uint64_t* d_state, const int32_t* d_tmax,
                          int32_t tmax, int32_t* d_out, const MdParams* d_params, const MdAux& ax,
                          int pass, hipStream_t s, hipEvent_t ev_start = nullptr,
                          hipEvent_t ev_stop = nullptr);
// One code, one launch: grid = code_chunks(w, L) single-wave workgroups, the
// 3w rows (w <= 20) in the kernel arguments.  d_state: {64, 0} before the
// first launch; each launch leaves it so.  h_result: a host-visible word
// (pinned memory) that receives tag << 32 | min(max_zeros, tmax + 1) (the
// low half as int32) when the launch's last wave finishes.
struct alignas(16) SingleRows {
  uint32_t r[3 * (kMaxKExact - 1) * kRowWords];
};
hipError_t launch_md_single(bool thresholded, const SingleRows& rows, int w, int L,
                            int check_log2, int thr, int tmax, uint32_t tag, uint32_t* d_state,
                            uint64_t* h_result, hipStream_t s);
// set_period > 0: set ids restart every set_period codes (0: set id = index).
// d_gen (may be null): a search's completed-generation counter; the seed
// becomes sample_seed + (d_gen[0] + 1) * 1,000,003.
hipError_t launch_md_sampled(const uint32_t* d_rows, const int32_t* d_offsets, int32_t n,
                             const int32_t* d_tmax, int32_t tmax, int64_t sample_count,
                             uint64_t sample_seed, uint32_t set_period, const uint32_t* d_gen,
                             int32_t* d_zmax, uint32_t* d_stop, int32_t* d_out, uint32_t* d_ctrl,
                             hipStream_t s);
int md_main_blocks_per_cu(bool thresholded);

// ---- canonical form ------------------------------------------------------
// One wave per set in input order; a mask with bits above 48 or a size
// outside 1..16 is flagged (kErrBadK) and its outputs are unspecified.
hipError_t launch_canon(const uint64_t* d_masks, int32_t n, uint64_t* d_key_lo,
                        uint64_t* d_key_hi, uint64_t* d_canon, int32_t* d_stab, uint32_t* d_ctrl,
                        hipStream_t s);
// ---- memory ---------------------------------------------------------------
hipError_t launch_fill_u32(uint32_t* d, int64_t n, uint32_t value, hipStream_t s);
hipError_t launch_copy16(void* dst, const void* src, int64_t bytes, hipStream_t s);

hipError_t launch_sets_to_masks(const int32_t* d_sets, int32_t n, int32_t k, uint64_t* d_masks,
                                uint32_t* d_ctrl, hipStream_t s);

}  // namespace gt
