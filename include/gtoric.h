/*
 * gtoric.h — C ABI of libgtoric.so, the gfx950 exact-search core for GF(8)
 * generalized toric codes of length 49.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative value on error;
 *     gt_last_error() then returns a message (thread-local).
 *   - Host-pointer entry points copy in, compute and copy out before they
 *     return (the reference oracles are synchronous: kernel.py:321-323,
 *     kernel_cuda_bp.py:322-323, kernel_canon.py:174-176).
 *   - *_dev entry points take device pointers and a hipStream_t (passed as
 *     void*, NULL = the context's stream); they only enqueue work.  Call
 *     gt_sync() to wait and to collect device-side input errors.
 *   - A context owns one HIP stream and grow-only device buffers.  Calls on one
 *     context are serialised by a mutex; separate contexts run concurrently.
 *   - Min-distance results are max_zeros values.  With tmax = 49 - target_distance,
 *     a result r <= tmax is the exact maximum number of zeros of a nonzero
 *     codeword; on early abort r is exactly tmax + 1 (the reference returns a
 *     scheduling-dependent value > tmax there; "kernel backup.py":116-125).
 */
#ifndef GTORIC_H
#define GTORIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gt_ctx gt_ctx;

/* Message for the last failing call on this thread. */
const char* gt_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int gt_device_count(int32_t* n);

/* Open a context on HIP device `device` (replaces the device set-up of
 * init_cuda_oracles_bp, kernel_cuda_bp.py:330-347, and the constructor
 * DistanceOracleCUDABP.__init__, kernel_cuda_bp.py:251-257). */
int gt_open(int32_t device, gt_ctx** out);
int gt_close(gt_ctx* ctx);

/* Compute-unit count, peak shader clock (kHz), LDS bytes per workgroup and the
 * device name; callers size batches from the CU count
 * (champion_search_canonical.py:626, evolve_search.py:197). */
int gt_device_info(gt_ctx* ctx, int32_t* cu_count, int32_t* clock_khz,
                   int32_t* lds_per_block, char* name, int32_t name_len);

/* Per-k launch parameters (autotuner): Gray walk bits per lane L (4..30) and
 * the early-exit check interval 2^check_log2 blocks of 16 steps. */
int gt_set_md_params(gt_ctx* ctx, int32_t k, int32_t L, int32_t check_log2);
int gt_get_md_params(gt_ctx* ctx, int32_t k, int32_t* L, int32_t* check_log2);

/* Batched minimum distance, mixed k (replaces DistanceOracle.max_zeros_batch,
 * kernel.py:276-323, and DistanceOracleCUDABP.max_zeros_batch,
 * kernel_cuda_bp.py:290-323; kernel eval_min_distance_batch_bp,
 * kernel.py:97-203).  Set s has indices idx[offsets[s] .. offsets[s+1]).
 * tmax_per_set (may be NULL) overrides the scalar tmax.  sample_count > 0 runs
 * the sampled stream of kernel.py:136-178 with set_id = s.  Exact mode needs
 * k <= 21, sampled k <= 49.  out[s] = max_zeros as documented above. */
int gt_md_batch(gt_ctx* ctx, const int32_t* idx, const int32_t* offsets, int32_t n,
                const int32_t* tmax_per_set, int32_t tmax, int64_t sample_count,
                uint64_t sample_seed, int32_t* out);

/* Same on device pointers; n_idx = offsets[n].  thresholded = 0 promises that
 * every tmax >= 48 (exact search; no early-exit checks). */
int gt_md_batch_dev(gt_ctx* ctx, const int32_t* d_idx, int64_t n_idx,
                    const int32_t* d_offsets, int32_t n, const int32_t* d_tmax_per_set,
                    int32_t tmax, int32_t thresholded, int64_t sample_count,
                    uint64_t sample_seed, int32_t* d_out, void* stream);

/* One code, one kernel launch (replaces DistanceOracle.max_zeros,
 * kernel.py:241-274, and DistanceOracleCUDABP.max_zeros,
 * kernel_cuda_bp.py:266-288).  k = 0 gives 0; k > 21 is an error. */
int gt_md_single(gt_ctx* ctx, const int32_t* idx, int32_t k, int32_t tmax, int32_t* out);

/* Canonical form under AGL2(F7) of n sets of size k given as sorted index rows
 * (replaces CanonicalOracle.compute, kernel_canon.py:151-176, and
 * canonical_forms_batch, champion_search_canonical.py:115-140).  Outputs the
 * reference key (lo, hi) — positions 0..k-1 for k <= 10, positions 1..k-1 for
 * k >= 11, 6 bits each — the canonical 49-bit image mask, and |Stab_AGL2(S)|.
 * 1 <= k <= 16.  Any output pointer except key_lo may be NULL. */
int gt_canon_batch(gt_ctx* ctx, const int32_t* sets, int32_t n, int32_t k,
                   uint64_t* key_lo, uint64_t* key_hi, uint64_t* canon_mask, int32_t* stab);

/* Same for 49-bit masks of mixed size 1..16 (bit i = point i). */
int gt_canon_masks(gt_ctx* ctx, const uint64_t* masks, int32_t n, uint64_t* key_lo,
                   uint64_t* key_hi, uint64_t* canon_mask, int32_t* stab);
int gt_canon_masks_dev(gt_ctx* ctx, const uint64_t* d_masks, int32_t n, uint64_t* d_key_lo,
                       uint64_t* d_key_hi, uint64_t* d_canon_mask, int32_t* d_stab,
                       void* stream);

/* Wait for `stream` (NULL = context stream) and report device-side input
 * errors (bad index, bad size) raised by *_dev calls since the last sync. */
int gt_sync(gt_ctx* ctx, void* stream);

/* Kernel timing with HIP events recorded around the dominant kernel of each
 * call ("md_main", "md_single", "canon").  gt_prof_read waits for the events,
 * returns count and total milliseconds, and clears the record. */
int gt_prof_enable(gt_ctx* ctx, int32_t on);
int gt_prof_read(gt_ctx* ctx, const char* kernel, int32_t* count, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* GTORIC_H */
