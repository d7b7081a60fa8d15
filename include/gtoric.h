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
 *     context are serialised by a mutex, and their device work too: a call on
 *     another stream than the previous call's waits (an event) for the work
 *     already enqueued there, since both use the context's buffers.  Separate
 *     contexts run concurrently.
 *   - Device-side input errors of *_dev calls, graph replays and searches are
 *     reported by gt_sync (gt_evolve_sync); a host-pointer call reports only
 *     its own.
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

/* Per-k launch parameters (autotuner) of the thresholded batch walk and the
 * one-code launch: Gray walk bits per lane L (4..30; chunks of 64 x 2^L
 * codewords) and the early-exit check interval 2^check_log2 blocks of 16
 * steps.  The exact batch walk deals equal runs of blocks and reads neither;
 * results never depend on them. */
int gt_set_md_params(gt_ctx* ctx, int32_t k, int32_t L, int32_t check_log2);
int gt_get_md_params(gt_ctx* ctx, int32_t k, int32_t* L, int32_t* check_log2);

/* First-pass level of the exact and thresholded batch search for codes of k
 * generators (3 <= k <= 21; k <= 2 walks at level 1); part of the per-k
 * parameters above, with the same effect on results (none).  A level-m pass
 * walks the subcodes that vanish at m points of a line through the torus
 * point (1,1) (1 subcode for m = 1, 2, 5, 6, 3 for m = 3, 4, of dimension
 * k - m); a maximum it finds of at least 7m - 7 zeros is exact, and so is
 * every level-m maximum of a code known to have at least 7m - 6.  Codes
 * below that are searched again at a lower level: in the same wavefront for
 * k <= 8, else at m - 1 when that level is certified by what the first pass
 * found, and at level 2 (k >= 9, certified by the k - 1 zeros every code
 * has) or level 1 otherwise (the subcode through (1,1), which the
 * reference's enumeration of all 8^k - 1 coefficient vectors, kernel.py:
 * 128-139, amounts to).  1 <= level <= min(k - 1, 6), and k - level must not
 * decrease with k.  Defaults: 1 for k <= 3, 2 for k = 4..6, 3 for k = 7, 8,
 * 4 above. */
int gt_set_md_level(gt_ctx* ctx, int32_t k, int32_t level);
int gt_get_md_level(gt_ctx* ctx, int32_t k, int32_t* level);

/* The level table itself (host only, no device): representative s of level
 * m as a bit mask over r (the points r d of a line, bit r set <=> r in the
 * set), the number of representatives, and the certificate 7m - 6 (the
 * zeros that make level m exact beforehand; a found maximum is exact one
 * zero earlier).  No counterpart in the reference; for checking the table. */
int gt_md_level_info(int32_t m, int32_t s, uint32_t* rep_mask, int32_t* nsub, int32_t* cert);

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

/* gt_md_batch_dev for a batch whose sets have at most kmax points (a larger
 * set is an input error: result 0, reported by gt_sync).  Codes of at most
 * 8 points are searched one wavefront per code in one launch; with kmax <= 8
 * (and the one-wave bound at its default, gt_set_md_fused_max_k) that
 * launch is the whole call.  gt_md_batch_dev is this call with kmax = 21;
 * host-pointer calls take kmax from their offsets.  Replaces the same
 * reference interface as gt_md_batch (kernel.py:276-323). */
int gt_md_batch_dev2(gt_ctx* ctx, const int32_t* d_idx, int64_t n_idx,
                     const int32_t* d_offsets, int32_t n, const int32_t* d_tmax_per_set,
                     int32_t tmax, int32_t thresholded, int32_t kmax, int64_t sample_count,
                     uint64_t sample_seed, int32_t* d_out, void* stream);

/* Per-code record of the work of exact calls: while on (default off), each
 * exact or thresholded call on the context's own scratch (not the
 * allocation-free call) writes, for each code, the projective codewords of
 * the subcodes its passes walked (exact mode: the codewords evaluated;
 * thresholded: an upper bound, chunks stop early); the getter copies the
 * record of the last such call, whose batch size n must be, after waiting
 * for it.  No counterpart in the reference (whose enumeration walks 8^k - 1
 * coefficient vectors per code); for measuring the search. */
int gt_set_md_walk_record(gt_ctx* ctx, int32_t on);
int gt_get_md_walk_record(gt_ctx* ctx, uint64_t* out, int32_t n);

/* Largest set size searched one wavefront per code (1..8, default 8); larger
 * codes go through the planned walk over all CUs.  Results never depend on
 * it (1 sends every code with a walk through the planned walk). */
int gt_set_md_fused_max_k(gt_ctx* ctx, int32_t k);
int gt_get_md_fused_max_k(gt_ctx* ctx, int32_t* k);

/* Allocation-free variant of gt_md_batch_dev (exact or thresholded; no
 * sampling).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its scratch (plan,
 * per-code minima, bit rows) is one file-scope device array per GPU, held
 * until the process ends; the call itself calls no hipMalloc/hipFree.
 * Limits: n <= 16384 codes and n_idx <= 65536 indices (larger calls fail
 * with an error; gt_md_batch_dev_noalloc_limits reports them).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other contexts, other threads) race and must be ordered by the
 * caller.  Device-side input errors are reported by gt_sync, as for the
 * *_dev calls. */
int gt_md_batch_dev_noalloc(gt_ctx* ctx, const int32_t* d_idx, int64_t n_idx,
                            const int32_t* d_offsets, int32_t n, const int32_t* d_tmax_per_set,
                            int32_t tmax, int32_t thresholded, int32_t* d_out, void* stream);
int gt_md_batch_dev_noalloc_limits(int32_t* max_codes, int64_t* max_idx);

/* One code (replaces DistanceOracle.max_zeros, kernel.py:241-274, and
 * DistanceOracleCUDABP.max_zeros, kernel_cuda_bp.py:266-288): one kernel
 * launch over the subcode through (1,1) for k <= 11; from k = 12 on the
 * code is searched as a batch of one (gt_md_batch: anchored subcodes, a few
 * launches, far fewer codewords).  k = 0 gives 0; k > 21 is an error. */
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

/* Graph capture of *_dev calls on the context stream (pass stream = NULL to
 * them while capturing).  Buffers must already have their size: run the same
 * calls once eagerly first.  A graph records raw buffer addresses, so a later
 * call that grows a context buffer invalidates it (gt_graph_launch then fails).
 * No memset nodes are recorded: every reset is a kernel.  No reference
 * counterpart (the reference launches eagerly, kernel_cuda_bp.py:314-322). */
typedef struct gt_graph gt_graph;
int gt_capture_begin(gt_ctx* ctx);
int gt_capture_end(gt_ctx* ctx, gt_graph** out);
int gt_graph_launch(gt_graph* graph, void* stream);
int gt_graph_destroy(gt_graph* graph);

/* Continuously batched serving: one persistent kernel on its own stream
 * (no reference counterpart; it serves the synchronous max_zeros_batch calls
 * of champion_search_canonical.py:217-234 and evolve_search.py:223-263 as a
 * stream).  slots (power of two) bounds the requests in flight (memory
 * budget); work_budget bounds the projective codewords their exact walks
 * visit ((8^(k-1) - 1) / 7 for k points; a request larger than the budget
 * runs alone).  L_by_k / check_by_k (22 entries, k = 0..21, may be NULL) are
 * the walk parameters.  The kernel stops itself after idle_s seconds without
 * requests or life_s seconds in total; requests it had not taken then are
 * handed to a new launch by the next submit or poll, so none is lost.
 * Requests: k <= 21 points (the exact limit of kernel.py:246-247) and a
 * threshold tmax = 48 - d_stop; results as for gt_md_batch (k = 0 gives 0). */
typedef struct gt_server gt_server;
int gt_serve_start(int32_t device, int32_t slots, int64_t work_budget, const int32_t* L_by_k,
                   const int32_t* check_by_k, double idle_s, double life_s, gt_server** out);
/* Non-blocking admission: *accepted = 0 when no slot or budget is free. */
int gt_serve_submit(gt_server* srv, const int32_t* idx, int32_t k, int32_t tmax, uint64_t tag,
                    int32_t* accepted);
/* Load-dependent chunk length (no reference counterpart): a request admitted
 * while fewer than below_work projective codewords are outstanding walks
 * chunks of L_short_by_k[k] steps (22 entries, 4..30; used only where
 * shorter than the start call's L_by_k), so a lightly loaded server answers
 * large requests sooner and a loaded one keeps its long, cheaper chunks.
 * Results do not change.  NULL turns it off. */
int gt_serve_set_short_chunks(gt_server* srv, const int32_t* L_short_by_k, int64_t below_work);
/* Completed requests (tag, result), at most max_n per call. */
int gt_serve_poll(gt_server* srv, uint64_t* tags, int32_t* results, int32_t max_n,
                  int32_t* n_done);
int gt_serve_in_flight(gt_server* srv, int64_t* n);
/* Run a whole arrival stream from the calling thread: request i is admitted
 * no earlier than arrival_s[i] seconds after the call starts; latency_s[i] is
 * completion minus scheduled arrival (queueing included). */
int gt_serve_run_stream(gt_server* srv, const int32_t* idx, const int32_t* offsets, int32_t n,
                        const int32_t* tmax, const double* arrival_s, int32_t* results,
                        double* latency_s, double* wall_s, double timeout_s);
/* Stop the kernel (only with nothing in flight; the next submit relaunches
 * it) and return its counters: out[0] dispatcher passes, [1] requests
 * dispatched, [2] dispatcher busy seconds, [3] chunks walked, [4] chunks
 * skipped, [5..8] worker seconds waiting / finding the request and staging
 * rows / walking / folding and claiming (summed over waves), [9] row
 * stagings, [10] dispatcher lifetime seconds, [11] worker waves.  Writes
 * min(n, 12) values.  No reference counterpart (observability). */
int gt_serve_stats(gt_server* srv, double* out, int32_t n);

/* Stop the kernel (waits for it to exit) and free the server. */
int gt_serve_stop(gt_server* srv);

/* Surrogate ranking MLP (no reference counterpart: it ranks the proposals
 * of evolve_search.py:509-523 before exact evaluation and never reports a
 * distance).  Input: 49-bit masks; features = 49 mask bits, |S|/16 and the 48
 * bins of the ordered difference multiset divided by |S|.  128 -> 128 ReLU ->
 * 128 ReLU -> 1, fp32 master weights with bf16 copies, fp32 accumulation
 * (bf16 MFMA), loss 0.5 * mean((y - target)^2), Adam (0.9, 0.999, 1e-8).
 * Parameters are one flat fp32 vector: W1[128][128], b1[128], W2[128][128],
 * b2[128], w3[128], b3. */
typedef struct gt_mlp gt_mlp;
int gt_mlp_num_params(int32_t* n);
int gt_mlp_create(int32_t device, uint64_t seed, gt_mlp** out);
int gt_mlp_destroy(gt_mlp* mlp);
int gt_mlp_get_params(gt_mlp* mlp, float* theta);
int gt_mlp_set_params(gt_mlp* mlp, const float* theta);
/* bf16 features, n x 128 (host output). */
int gt_mlp_features(gt_mlp* mlp, const uint64_t* masks, int32_t n, uint16_t* feat);
/* Gradient of the loss at the current parameters (no update). */
int gt_mlp_grads(gt_mlp* mlp, const uint64_t* masks, const float* target, int32_t n,
                 float* grads, float* loss);
/* One Adam step on a batch; *loss is the loss before the step. */
int gt_mlp_train_step(gt_mlp* mlp, const uint64_t* masks, const float* target, int32_t n,
                      float lr, float* loss);
/* Adam steps on consecutive batches of `batch` rows (the last may be
 * shorter), queued back to back with one synchronisation at the end;
 * losses[b] (may be NULL) is batch b's loss before its step.  Same steps as
 * gt_mlp_train_step on each batch in turn. */
int gt_mlp_train_batches(gt_mlp* mlp, const uint64_t* masks, const float* target, int32_t n,
                         int32_t batch, float lr, float* losses);
int gt_mlp_train_step_dev(gt_mlp* mlp, const uint64_t* d_masks, const float* d_target,
                          int32_t n, float lr, void* stream);
int gt_mlp_predict(gt_mlp* mlp, const uint64_t* masks, int32_t n, float* out);
int gt_mlp_predict_dev(gt_mlp* mlp, const uint64_t* d_masks, int32_t n, float* d_out,
                       void* stream);
/* The MLP's bf16 MFMA tile product as a plain GEMM, C = A B (row-major, fp32
 * out; M, N, K multiples of 16); exposed so its lane layout is testable. */
int gt_bf16_gemm(int32_t device, const uint16_t* A, const uint16_t* B, float* C, int32_t M,
                 int32_t N, int32_t K);

/* Device hash set of nonzero 64-bit keys (canonical masks) and one canonical
 * BFS level on the device (replaces the host loops of
 * champion_search_canonical.py:496-506 and :308-345).  gt_dedup marks
 * flag[i] = 1 when key i is new: not inserted by an earlier call and the
 * first occurrence within this call; every key is inserted. */
typedef struct gt_hashset gt_hashset;
int gt_hashset_create(gt_ctx* ctx, int32_t log2_capacity, gt_hashset** out);
int gt_hashset_destroy(gt_hashset* set);
int gt_hashset_clear(gt_hashset* set, void* stream);
int gt_hashset_count(gt_hashset* set, int64_t* n);
int gt_dedup(gt_hashset* set, const uint64_t* keys, int32_t n, uint32_t* flag);
int gt_dedup_dev(gt_hashset* set, const uint64_t* d_keys, int32_t n, uint32_t* d_flag,
                 void* stream);
/* Ordered compaction: d_out_idx[0 .. *d_count) = indices i with d_flag[i] != 0. */
int gt_compact_dev(gt_ctx* ctx, const uint32_t* d_flag, int32_t n, uint32_t* d_out_idx,
                   uint32_t* d_count, void* stream);
/* Children of n_par parent masks of size kp (each parent plus one outside
 * point), their canonical masks, the classes not yet in `seen` (first
 * occurrence in child order) into d_out_masks, and their max_zeros at
 * threshold tmax into d_out_mz.  *n_new = number of new classes (error if
 * above cap).  Synchronous. */
int gt_bfs_level(gt_ctx* ctx, gt_hashset* seen, const uint64_t* d_parents, int32_t n_par,
                 int32_t kp, int32_t tmax, uint64_t* d_out_masks, int32_t* d_out_mz,
                 int32_t cap, int32_t* n_new);
/* The same level with the reference's other modes
 * (champion_search_canonical.py:386-397, 483-512): canonical = 0 is its raw
 * sorted-set BFS above k_canonical_max (children deduplicated as they are;
 * kp + 1 <= 21 exact, <= 49 sampled); sample_count > 0 is its sampled
 * screening (kernel.py:136-178 stream, seed sample_seed; set ids restart
 * every set_period classes when set_period > 0, as for the reference's
 * batches of batch_size sets). */
int gt_bfs_level_ex(gt_ctx* ctx, gt_hashset* seen, const uint64_t* d_parents, int32_t n_par,
                    int32_t kp, int32_t tmax, int32_t canonical, int64_t sample_count,
                    uint64_t sample_seed, int32_t set_period, uint64_t* d_out_masks,
                    int32_t* d_out_mz, int32_t cap, int32_t* n_new);

/* Evolutionary search with every generation on the device (replaces the
 * generation loop of run_evolution, evolve_search.py:322-539, with exact
 * scores).  A generation: n_prop children, each a 1- or 2-point swap of a
 * size-4 tournament winner (highest 1000 d - age); their canonical masks;
 * the classes not evaluated before (first occurrence in proposal order);
 * pop of them chosen in proposal order or, with a ranker, by the ranker's
 * score (ties: proposal order), random k-sets filling a short generation;
 * exact distances of the chosen; survivors = best pop of parents and
 * children by distance, then age.  Random draws are a function of (seed,
 * generation, index), so a search is reproducible.  One captured graph is
 * replayed per generation; the population, the hash set of evaluated
 * classes and the per-generation log stay in device memory.
 *   - 2 <= k <= 16, 1 <= pop <= 4096, pop <= n_prop <= 8192.
 *   - init_masks: pop host masks of size k (the initial population, slot 0).
 *   - log2_seen: capacity of the evaluated-class set (0 = chosen from
 *     pop * (max_generations + 1)).
 *   - ranker (may be NULL): its parameters are read by every generation;
 *     update them only between gt_evolve_sync and the next gt_evolve_run.
 * Destroy a search before closing its context. */
typedef struct gt_evolve gt_evolve;
int gt_evolve_create(gt_ctx* ctx, int32_t k, int32_t pop, int32_t n_prop, uint64_t seed,
                     int32_t max_generations, int32_t log2_seen, gt_mlp* ranker,
                     const uint64_t* init_masks, gt_evolve** out);
/* The same search with the reference's sampled scores and beyond the
 * canonical limit (evolve_search.py:266-282, 350-351, where every GA score is
 * sampled): sample_count > 0 scores each chosen set by the sampled stream of
 * kernel.py:136-178 (no early stop; seed `seed` for the initial population,
 * seed + (g + 1) * 1,000,003 for generation g, set id = index among the
 * chosen), so recorded distances are sampled upper bounds, not certified;
 * canonical = 0 deduplicates sets as they are instead of by AGL2 class.
 * Limits: canonical 2 <= k <= 16; otherwise k <= 21 exact, k <= 47 sampled. */
int gt_evolve_create2(gt_ctx* ctx, int32_t k, int32_t pop, int32_t n_prop, uint64_t seed,
                      int32_t max_generations, int32_t log2_seen, gt_mlp* ranker,
                      const uint64_t* init_masks, int64_t sample_count, int32_t canonical,
                      gt_evolve** out);
/* Enqueue `generations` more generations (asynchronous). */
int gt_evolve_run(gt_evolve* ev, int32_t generations);
/* Train the search's ranker on its own log slots [slot0, slot1) (classes and
 * exact distances, already in device memory): features, targets
 * standardised over the slots, `epochs` passes of Adam steps over batches of
 * `batch` rows in log order.  Enqueued on the context stream between
 * generations, with no host copy or synchronisation; the next generation
 * ranks with the updated network.  No reference counterpart. */
int gt_evolve_train_ranker(gt_evolve* ev, int32_t slot0, int32_t slot1, int32_t batch, float lr,
                           int32_t epochs);
/* Wait for them; device-side errors are reported here. */
int gt_evolve_sync(gt_evolve* ev, int32_t* generations_done);
/* Log slots [slot0, slot1): slot 0 is the initial population, slot g the sets
 * evaluated by generation g (pop canonical masks and exact distances each),
 * best[] the population's best distance after it, n_new[] its count of new
 * classes among the proposals.  Any output may be NULL. */
int gt_evolve_log(gt_evolve* ev, int32_t slot0, int32_t slot1, uint64_t* classes,
                  int32_t* dists, int32_t* best, int32_t* n_new);
/* Current population (sorted best first) and the number of classes in the
 * evaluated set. */
int gt_evolve_population(gt_evolve* ev, uint64_t* masks, int32_t* dists, int32_t* age,
                         int64_t* seen_classes);
int gt_evolve_destroy(gt_evolve* ev);

/* Device memory of the context's device, for *_dev callers that do not bring
 * their own allocator.  Copies are ordered on `stream` (NULL = context
 * stream) and return when the copy has finished.  gt_fill_u32_dev and
 * gt_copy_dev are kernels (graph-safe; no memset nodes); gt_copy_dev needs a
 * size that is a multiple of 16 bytes. */
int gt_dev_alloc(gt_ctx* ctx, int64_t bytes, void** out);
int gt_dev_free(gt_ctx* ctx, void* d);
int gt_memcpy_h2d(gt_ctx* ctx, void* d, const void* h, int64_t bytes, void* stream);
int gt_memcpy_d2h(gt_ctx* ctx, void* h, const void* d, int64_t bytes, void* stream);
int gt_fill_u32_dev(gt_ctx* ctx, uint32_t* d, int64_t n, uint32_t value, void* stream);
int gt_copy_dev(gt_ctx* ctx, void* dst, const void* src, int64_t bytes, void* stream);

/* Issue-rate probe (no reference counterpart; the roofline's attainable
 * ceiling): times, on `device`, 8 waves per SIMD issuing a straight-line
 * stream of one instruction form — 0 v_xor_b32 VGPR-VGPR, 1 v_xor_b32 with an
 * SGPR operand, 2 v_or3_b32, 3 v_bcnt_u32_b32, 4 v_min3_u32, 5 v_mov_b32 from
 * an SGPR — or 6 the min-distance walk's own 16-codeword block (174 wave64
 * instructions), `trips` loop trips per wave.  *ms is the kernel time,
 * *wave_instr the wave64 instructions of the probed form it issued. */
int gt_issue_probe(int32_t device, int32_t form, int32_t trips, double* ms, double* wave_instr);

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
