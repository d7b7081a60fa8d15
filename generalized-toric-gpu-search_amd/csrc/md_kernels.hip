// Exact / thresholded / sampled minimum distance of GF(8) generalized toric
// codes on gfx950.
//
// Replaces eval_min_distance_batch_bp (kernel.py:97-203, kernel_cuda_bp.py:146-238)
// and eval_min_distance / eval_min_distance_single_bp (kernel.py:63-94,
// kernel_cuda_bp.py:99-127).  Work decomposition (see gt_common.h for the
// anchored subcodes, levels and their certificates):
//
//   md_fused    one wave per code of k <= 8 generators: the rows of the
//               code's first-pass subcodes built in LDS (a Gauss-Jordan
//               kernel basis per representative point set), walked by the
//               wave, certified, and if not certified the next lower level
//               in the same wave; the result written.  Larger codes are
//               listed for:
//   md_prepl    workgroup 0: the listed codes bucketed by size (MdPlan,
//               order[], roff[]), claim counters reset; workgroups 1..: a
//               wave per listed code in turn, its state word reset and the
//               bit rows of its subcodes written to global memory.
//   md_main     persistent.  Exact mode: two static deals, one share of each
//               per wave: the codes whose subcodes have no Gray piece above
//               the tiny part, whole, in equal counts; the other codes' work
//               units (16-step Gray blocks of their subcodes' pieces and
//               their tiny parts, in plan order) in equal runs.  A code
//               inside one run is finished by its wave, a code split over
//               several runs is folded into its state word by each and
//               finished by the last.  Thresholded mode: chunk ids claimed
//               from 16 counters, decoded against the plan in LDS (bucket,
//               piece, subcode, code), 64 lanes x 2^L Gray steps each, every
//               chunk folded at once (other waves stop on the folded value).
//               Finishing a code writes max_zeros clamped at tmax + 1, or,
//               below the level's certificate, lists the code for
//   md_prepl,   the same two launches over that list: the middle pass, one
//   md_main     level lower, for the codes whose first-pass maximum already
//               certifies that level; then the final pass, at the level the
//               bound z* >= k - 1 certifies (2 for k >= 9, else 1), for the
//               rest.  The middle and final passes always finish.
//   The sampled path keeps the full rows x^b * g_i of C(S) (md_rows).

// Inner step (per projective codeword, all VALU; the block's rows are
// wave-uniform, loaded with scalar loads and copied to VGPRs):
//   6 v_xor_b32 (row add, VOP2 with two VGPR operands), 2 v_or3_b32 (plane
//   OR), 2 v_bcnt_u32_b32, and one v_min3_u32 per two codewords.
#include <cstddef>

#include <hip/hip_ext.h>

#include "gt_anchor.h"
#include "gt_launch.h"

namespace gt {

// --------------------------------------------------------------- kernels --

__device__ __forceinline__ void md_rows_one(int64_t t, const int32_t* __restrict__ idx,
                                            int64_t n_idx, const uint32_t* __restrict__ mplanes,
                                            uint32_t* __restrict__ rows, uint32_t* ctrl) {
  if (t >= n_idx) return;
  int g = idx[t];
  uint32_t w[6] = {0, 0, 0, 0, 0, 0};
  if (g < 0 || g >= kNpts) {
    atomicOr(&ctrl[2], kErrBadIndex);
  } else {
#pragma unroll
    for (int i = 0; i < 6; ++i) w[i] = mplanes[g * kRowWords + i];
  }
  uint32_t* out = rows + (size_t)t * 3 * kRowWords;
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    uint4* o = reinterpret_cast<uint4*>(out + b * kRowWords);
    o[0] = make_uint4(w[0], w[1], w[2], w[3]);
    o[1] = make_uint4(w[4], w[5], 0u, 0u);
    mul_x(w);
  }
}

// Rows of one code for a pass at level m = code_level(k, pass)
// (md_rows_level), starting at slot kRowsPerIdx * offsets[code].  The wave
// also resets the code's state word, writes the result of a code without a
// walk (k <= 1), and, in the first and middle passes, sets aux[code] = 1
// when no direction has full rank (the code then goes to the final pass;
// its rows are not used for a result).  Codes with
// k > 21 or offsets outside [0, n_idx] are flagged by the plan (no row is
// written for them).
__device__ __forceinline__ void md_rows_anchor(int code, int32_t n, int64_t n_idx,
                                               const int32_t* __restrict__ idx,
                                               const int32_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ mplanes,
                                               const uint8_t* __restrict__ level, int pass,
                                               uint32_t* __restrict__ rows, uint64_t* state,
                                               uint32_t* aux, int32_t* out,
                                               const int32_t* tmax_arr, int32_t tmax_s,
                                               uint32_t* ctrl) {
  if (code >= n) return;
  const uint32_t lane = lane_id();
  const int o = offsets[code];
  const int k = offsets[code + 1] - o;
  const bool bad_k = k < 0 || k > kMaxKExact || o < 0 || (int64_t)o + k > n_idx;
  if (lane == 0) {
    state[code] = 64ull;   // minimum weight 64, no chunk finished
    aux[code] = 0u;
    // No walk: k = 0 (and bad input) gives 0, k = 1 a monomial (weight 49).
    if (bad_k || k <= 1)
      out[code] = (!bad_k && k == 1) ? md_result(49, tmax_arr ? tmax_arr[code] : tmax_s) : 0;
  }
  if (k < 1 || bad_k) return;
  int g = 0;
  bool bad = false;
  if (lane < (uint32_t)k) {
    g = idx[o + (int)lane];
    bad = g < 0 || g >= kNpts;
  }
  if (__ballot(bad) != 0ull) {
    if (lane == 0) atomicOr(&ctrl[2], kErrBadIndex);
    return;
  }
  if (k < 2) return;
  uint32_t* base = rows + (size_t)o * kRowsPerIdx * kRowWords;
  const int m = __builtin_amdgcn_readfirstlane(code_level(level, k, pass));
  bool ok;
  switch (m) {
    case 1: ok = md_rows_level<1>(g, k, lane, mplanes, base); break;
    case 2: ok = md_rows_level<2>(g, k, lane, mplanes, base); break;
    case 3: ok = md_rows_level<3>(g, k, lane, mplanes, base); break;
    case 4: ok = md_rows_level<4>(g, k, lane, mplanes, base); break;
    case 5: ok = md_rows_level<5>(g, k, lane, mplanes, base); break;
    default: ok = md_rows_level<6>(g, k, lane, mplanes, base); break;   // kMaxLevel
  }
  if (!ok) {
    if (pass == kPassFinal) {
      // Final pass, level 2 without a full-rank direction: every index is
      // the same point, the set spans the zero codeword, and zero rows give
      // its 49 zeros.
      const int w = k - m;
      for (int r = (int)lane; r < level_nsub(m) * 3 * w; r += kWave) {
        uint4* rr = reinterpret_cast<uint4*>(base + (size_t)r * kRowWords);
        rr[0] = make_uint4(0u, 0u, 0u, 0u);
        rr[1] = make_uint4(0u, 0u, 0u, 0u);
      }
    } else if (lane == 0) {
      aux[code] = 1u;
    }
  }
}

__global__ __launch_bounds__(256) void md_rows_kernel(const int32_t* __restrict__ idx,
                                                      int64_t n_idx,
                                                      const uint32_t* __restrict__ mplanes,
                                                      uint32_t* __restrict__ rows,
                                                      uint32_t* ctrl) {
  md_rows_one((int64_t)blockIdx.x * blockDim.x + threadIdx.x, idx, n_idx, mplanes, rows, ctrl);
}

// The order the one-wave search takes the codes in: one workgroup checks
// every code's size, writes the results of codes without a walk (k <= 1, and
// 0 for bad sizes, flagged), lists the codes with kfused < k <= kmax for the
// planned walk (k > kmax is an input error: a caller's bound that skips the
// planned walk), and counting-sorts the others by k, largest first, into
// order[0 .. *fused_count) (the order inside a size does not matter).
constexpr int kOrderThreads = 1024;
constexpr int kFusedCodesPerWave = 1;   // md_fused launches a wave per code,
constexpr int kFusedMaxWaves = 8192;    // at most this many (4096 codes of k in
                                        // 3..8: 1 per wave 15 us, 2: 17, 4: 22, 8: 39)

__global__ __launch_bounds__(kOrderThreads) void md_order_kernel(
    const int32_t* __restrict__ offsets, int32_t n, int64_t n_idx,
    const int32_t* __restrict__ tmax_arr, int32_t tmax_s, int32_t* __restrict__ out,
    uint32_t* big_count, uint32_t* big_list, uint32_t* __restrict__ order,
    uint32_t* fused_count, int32_t kfused, int32_t kmax, uint64_t* __restrict__ walked,
    uint32_t* ctrl) {
  __shared__ uint32_t cnt[kFusedMaxK + 1];
  const int tid = (int)threadIdx.x;
  if (tid <= kFusedMaxK) cnt[tid] = 0u;
  __syncthreads();
  uint32_t err = 0u;
  for (int i = tid; i < n; i += kOrderThreads) {
    const int o = offsets[i];
    const int k = offsets[i + 1] - o;
    const bool bad_k = k < 0 || k > kMaxKExact || k > kmax || o < 0 || (int64_t)o + k > n_idx;
    if (walked) walked[i] = 0ull;
    if (bad_k || k <= 1) {
      if (bad_k) err |= kErrBadK;
      // No walk: k = 0 (and bad input) gives 0, k = 1 a monomial (weight 49).
      out[i] = (!bad_k && k == 1) ? md_result(49, tmax_arr ? tmax_arr[i] : tmax_s) : 0;
    } else if (k > kfused) {
      const uint32_t j = atomicAdd(big_count, 1u);
      if (j < (uint32_t)n) big_list[j] = (uint32_t)i;
      else err |= kErrBadK;   // (a count left by an interrupted call)
    } else {
      atomicAdd(&cnt[k], 1u);
    }
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t base = 0u;
    for (int k = kFusedMaxK; k >= 2; --k) {
      const uint32_t t = cnt[k];
      cnt[k] = base;
      base += t;
    }
    *fused_count = base;
  }
  __syncthreads();
  for (int i = tid; i < n; i += kOrderThreads) {
    const int o = offsets[i];
    const int k = offsets[i + 1] - o;
    const bool bad_k = k < 0 || k > kMaxKExact || k > kmax || o < 0 || (int64_t)o + k > n_idx;
    if (!bad_k && k >= 2 && k <= kfused) order[atomicAdd(&cnt[k], 1u)] = (uint32_t)i;
  }
  if (err) atomicOr(&ctrl[2], err);
}

// The one-wave search over order[0 .. *fused_count): W waves; in round r
// wave w takes position r W + w (r even) or (r + 1) W - 1 - w (r odd), a
// snake over the size-sorted codes.  With a wave per code the waves are
// dispatched largest code first, so the long searches start early and the
// short ones fill in behind them (one wave per code in input order left the
// SIMDs that drew several k = 8 codes 3x the mean); with more codes than
// waves every wave gets codes of every size class in turn.  At most 96
// VGPRs (five waves per SIMD; 120 unbounded, four): a call issued on another
// stream gets waves onto the SIMDs while this call's last waves finish.
__global__ __launch_bounds__(kFusedWaves * kWave, 5) void md_fused_kernel(
    const int32_t* __restrict__ idx, const int32_t* __restrict__ offsets, int32_t nwaves,
    const uint32_t* __restrict__ order, const uint32_t* fused_count,
    const uint32_t* __restrict__ mplanes, const MdParams* __restrict__ p,
    const int32_t* __restrict__ tmax_arr, int32_t tmax_s, int32_t* __restrict__ out,
    uint64_t* __restrict__ walked, uint32_t* ctrl) {
  __shared__ __align__(16) uint32_t lrows[kFusedWaves][kFusedRows * kRowWords];
  const int wv = (int)(threadIdx.x / kWave);
  const uint32_t w = blockIdx.x * kFusedWaves + wv;
  if (w >= (uint32_t)nwaves) return;
  const uint32_t lane = lane_id();
  const uint32_t m = __builtin_amdgcn_readfirstlane(*fused_count);
  const uint32_t W = (uint32_t)nwaves;
  for (uint32_t r = 0;; ++r) {
    const uint32_t pos = (r & 1u) ? (r + 1u) * W - 1u - w : r * W + w;
    if (pos >= m) break;   // later rounds start above m too
    const int code = (int)__builtin_amdgcn_readfirstlane(order[pos]);
    const int o = offsets[code];
    const int k = offsets[code + 1] - o;
    const int tmax = tmax_arr ? tmax_arr[code] : tmax_s;
    int g = 0;
    bool bad = false;
    if (lane < (uint32_t)k) {
      g = idx[o + (int)lane];
      bad = g < 0 || g >= kNpts;
    }
    if (__ballot(bad) != 0ull) {
      if (lane == 0) {
        atomicOr(&ctrl[2], kErrBadIndex);
        out[code] = 0;
      }
      continue;
    }
    uint64_t cw = 0;
    const uint32_t best = fused_search(g, k, tmax, lane, mplanes, p->level, lrows[wv], cw);
    if (lane == 0) {
      out[code] = md_result((int)best, tmax);
      if (walked) walked[code] = cw;
    }
  }
}

// ---- plan and rows of a listed set of codes ----------------------------------
// The list search: the first pass over the codes md_fused left (k > kfused,
// at their levels), then the middle pass (one level lower) and the final
// pass (second_level(k)) over the codes the first pass could not certify,
// each over its own list.  Workgroup 0 plans the listed codes (positions by
// bucket from LDS counters; the order inside a bucket does not matter); the
// other workgroups build their rows and reset their state words, a wave per
// listed code in turn.  With an empty list the plan is empty and the walk
// that follows returns at once.
constexpr int kPlanThreads = 256;
constexpr int kPrepRowBlocks = 256;   // row builders (k = 3..10 batch, 4096 codes: 64 -> 256: 122 -> 116 us per call)

__device__ __forceinline__ void md_list_plan(const int32_t* __restrict__ offsets, int32_t n,
                                             const MdParams* __restrict__ p,
                                             MdPlan* __restrict__ plan,
                                             uint32_t* __restrict__ order,
                                             uint32_t* __restrict__ roff,
                                             uint32_t* __restrict__ claims, uint32_t nl,
                                             const uint32_t* __restrict__ list, int pass,
                                             const MdAux& ax, uint32_t* ctrl) {
  constexpr int kB = kMaxKExact;
  __shared__ uint32_t cnt[kB + 1];
  // Per-bucket chunk counts and units, staged by one thread each so that the
  // serial scan below reads LDS only (not a chain of dependent global loads).
  __shared__ uint32_t cc_s[kB], c1_s[kB];
  __shared__ uint64_t un_s[kB];
  const int tid = threadIdx.x;
  if (tid <= kB) cnt[tid] = 0u;
  if (tid < kB) {
    const int k = min(kB + 1 - tid, kB);
    cc_s[tid] = p->chunks[pass][k];
    c1_s[tid] = p->chunks1[k];
    un_s[tid] = anchor_units(k, code_level(p->level, k, pass));
  }
  if (tid < kClaimShards) claims[tid * kClaimStride] = 0u;
  if (tid == 0 && pass == kPassFirst) {   // the later lists of this call
    *ax.mid_count = 0u;
    *ax.fb_count = 0u;
  }
  __syncthreads();
  for (uint32_t i = tid; i < nl; i += blockDim.x) {
    const int code = (int)list[i];
    if (code < n) atomicAdd(&cnt[kB + 1 - (offsets[code + 1] - offsets[code])], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    // Chunks of the pass (claimed walk) and at level 1 (the call is refused
    // when those reach 2^31, the limit the host-pointer call checks, so
    // neither pass can overflow).
    uint64_t eu = 0, ec = 0, e1 = 0;
    uint32_t ep = 0;
    for (int b = 0; b < kB; ++b) {
      const uint32_t t = cnt[b];
      const uint64_t cc = cc_s[b];
      plan->ub[b] = eu;
      plan->cb[b] = (uint32_t)min(ec, 1ull << 31);
      plan->pb[b] = ep;
      plan->cc[b] = (uint32_t)cc;
      cnt[b] = ep;
      eu += (uint64_t)t * un_s[b];
      ec = min(ec + (uint64_t)t * cc, 1ull << 31);
      e1 = min(e1 + (uint64_t)t * c1_s[b], 1ull << 31);
      ep += t;
    }
    const bool refused = ec >= (1ull << 31) || e1 >= (1ull << 31);
    if (refused) {   // no units, no chunks (the call reports the error)
      for (int b = 0; b < kB; ++b) plan->ub[b] = 0ull;
      atomicOr(&ctrl[2], kErrBadK);
    }
    plan->ub[kB] = refused ? 0ull : eu;
    plan->cb[kB] = refused ? 0u : (uint32_t)ec;
    plan->pb[kB] = ep;
  }
  __syncthreads();
  for (uint32_t i = tid; i < nl; i += blockDim.x) {
    const int code = (int)list[i];
    if (code >= n) continue;
    const int o = offsets[code];
    const uint32_t q = atomicAdd(&cnt[kB + 1 - (offsets[code + 1] - o)], 1u);
    order[q] = (uint32_t)code;
    roff[q] = (uint32_t)o;
  }
}

// list_count: entries of `list` (at most n are read; the count may be left
// over from an interrupted call, whose entries, codes of this batch or
// beyond it, are then searched again or skipped, never mis-attributed).
__global__ __launch_bounds__(kPlanThreads) void md_prepl_kernel(
    const int32_t* __restrict__ idx, int64_t n_idx, const uint32_t* __restrict__ mplanes,
    uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets, int32_t n,
    const MdParams* __restrict__ p, MdPlan* __restrict__ plan, uint32_t* __restrict__ order,
    uint32_t* __restrict__ roff, uint32_t* __restrict__ claims, uint64_t* __restrict__ state,
    int32_t* __restrict__ out, const int32_t* __restrict__ tmax_arr, int32_t tmax_s, MdAux ax,
    int32_t pass, uint32_t* ctrl) {
  const uint32_t* cnt = (pass == kPassFirst) ? ax.big_count
                        : (pass == kPassMid) ? ax.mid_count : ax.fb_count;
  const uint32_t* list = (pass == kPassFirst) ? ax.big_list
                         : (pass == kPassMid) ? ax.mid_list : ax.fb_list;
  uint32_t nl = __builtin_amdgcn_readfirstlane(*cnt);
  nl = min(nl, (uint32_t)max(n, 0));
  if (blockIdx.x == 0) {
    md_list_plan(offsets, n, p, plan, order, roff, claims, nl, list, pass, ax, ctrl);
    return;
  }
  constexpr uint32_t kWaves = kPrepRowBlocks * (kPlanThreads / kWave);
  for (uint32_t i = (blockIdx.x - 1) * (kPlanThreads / kWave) + threadIdx.x / kWave; i < nl;
       i += kWaves) {
    const uint32_t code = __builtin_amdgcn_readfirstlane(list[i]);
    if (code < (uint32_t)n)
      md_rows_anchor((int)code, n, n_idx, idx, offsets, mplanes, p->level, pass, rows, state,
                     ax.aux, out, tmax_arr, tmax_s, ctrl);
  }
}

// Chunk ids.  The plan numbers chunks bucket by bucket, largest walk first;
// inside a bucket piece-major: every code's chunks of piece w-1, then of
// piece w-2, ..., then one tiny chunk per code, consecutive ids of a piece
// going to consecutive codes (code = id mod codes in the bucket).  Chunk
// sizes therefore fall along the ids, apart from steps at bucket borders.
//
// The plan and the per-k parameters are copied into LDS once per workgroup,
// so decoding an id (a binary search over 21 bucket starts and a walk over
// the pieces) reads LDS and scalar registers only; the code index and its
// row offset (two independent loads) are the only global loads before the
// rows.
struct MdShared {
  alignas(8) uint32_t plan[sizeof(MdPlan) / 4];
  uint8_t L[kMaxKExact + 1];
  uint8_t check_log2[kMaxKExact + 1];
  uint8_t level[kMaxKExact + 1];
  uint8_t small_k[kPasses];
};

__device__ __forceinline__ void md_load_shared(MdShared& sh, const MdPlan* __restrict__ plan,
                                               const MdParams* __restrict__ p) {
  static_assert(sizeof(MdPlan) / 4 <= 256, "one plan word per thread");
  if (threadIdx.x < sizeof(MdPlan) / 4)
    sh.plan[threadIdx.x] = reinterpret_cast<const uint32_t*>(plan)[threadIdx.x];
  if (threadIdx.x < kMaxKExact + 1) {
    sh.L[threadIdx.x] = p->L[threadIdx.x];
    sh.check_log2[threadIdx.x] = p->check_log2[threadIdx.x];
    sh.level[threadIdx.x] = p->level[threadIdx.x];
  }
  if (threadIdx.x < kPasses) sh.small_k[threadIdx.x] = p->small_k[threadIdx.x];
  __syncthreads();
}

struct MdChunk {
  int code;
  int m;            // level of the pass
  int w;            // walk dimension of a subcode
  int L;
  uint32_t cmask;   // early-exit check mask
  int j;            // Gray piece, or tiny_top(w) for the tiny chunk
  uint32_t c;       // chunk index inside the piece
  uint32_t s;       // subcode
  uint32_t nch;     // chunks of the code
  uint32_t roff;    // first index of the code (kRowsPerIdx row slots per index)
};

// Wave-uniform decode of chunk id g < total.  Inside a piece, ids run over
// (chunk, subcode) pairs, then codes: id = ((c * nsub + s) * codes + q).
__device__ __forceinline__ MdChunk md_decode(const MdShared& sh, uint32_t g,
                                             const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ roff, int pass) {
  const uint32_t* cb = sh.plan + offsetof(MdPlan, cb) / 4;
  const uint32_t* pb = sh.plan + offsetof(MdPlan, pb) / 4;
  const uint32_t* ccv = sh.plan + offsetof(MdPlan, cc) / 4;
  // Last bucket whose first chunk id is <= g (empty buckets tie with their
  // successor, so the last tie is the nonempty one).
  int lo = 0, hi = kMaxKExact - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cb[mid] <= g) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(cb[lo]);
  const uint32_t cc = __builtin_amdgcn_readfirstlane(ccv[lo]);
  const uint32_t nb = (__builtin_amdgcn_readfirstlane(cb[lo + 1]) - b0) / cc;   // codes
  MdChunk r;
  r.nch = cc;
  const int k = kMaxKExact + 1 - lo;
  r.m = __builtin_amdgcn_readfirstlane(code_level(sh.level, k, pass));
  const uint32_t nsub = (uint32_t)level_nsub(r.m);
  r.w = k - r.m;
  r.L = __builtin_amdgcn_readfirstlane(sh.L[k]);
  r.cmask = (1u << __builtin_amdgcn_readfirstlane(sh.check_log2[k])) - 1u;
  uint32_t rel = g - b0;
  const int jt = tiny_top(r.w);
  int j = r.w - 1;
  for (; j > jt; --j) {
    const uint32_t span = (uint32_t)piece_chunks(j, r.L) * nsub * nb;
    if (rel < span) break;
    rel -= span;
  }
  r.j = j;
  const uint32_t qi = rel / nb;
  r.s = qi % nsub;
  r.c = qi / nsub;
  const uint32_t pos = __builtin_amdgcn_readfirstlane(pb[lo]) + (rel - qi * nb);
  r.code = (int)order[pos];
  r.roff = roff[pos];
  return r;
}

template <bool THR>
__device__ __forceinline__ uint32_t md_walk(const MdChunk& ck, const uint32_t* R, uint32_t lane,
                                            int thr, const uint32_t* best_g) {
  uint32_t b;
  if (ck.j > tiny_top(ck.w))
    b = gray_chunk<THR>(ConstRows{R}, (uint32_t)ck.j, (uint32_t)piece_L(ck.j, ck.L), ck.c, lane,
                        thr, ck.cmask, best_g);
  else
    b = tiny_chunk(ConstRows{R}, tiny_top(ck.w), lane);
  return wave_min(b);
}

// Fold a wave's minimum weight b over n finished chunks of a code into its
// state word (bits 0..31 the minimum, 32..63 the chunks finished), from lane
// 0.  The minimum is a returning atomic and the count's operand depends on
// its result, with an s_waitcnt 0 between them, so the count cannot be
// performed before the minimum; both halves of one word share a coherence
// point.  The wave whose count completes the code (all its nch chunks) reads
// the final minimum there and finishes the code (md_finish).
// A pass at level m > 1 certifies its maximum z only when z >= found_cert(m)
// = 7m - 7 (or z > tmax: the threshold is exceeded by a real codeword) and
// the code had a direction of full rank.  Any other code of the first pass
// is appended to the middle-pass list when the middle level is certified
// beforehand by the bound max(z, k - 1) on z* (7q - 6) or its found
// certificate is met by z (the middle walk, in the same direction, finds at
// least z), else to the final-pass list; a middle-pass code that its walk
// does not certify goes to the final-pass list too.  The final pass's level
// is certified by z* >= k - 1.  The rest are written.  Lane 0 only.
__device__ __forceinline__ void md_finish(int code, uint32_t bw, int k, int m, int mid, int tmax,
                                          const MdAux& ax, int32_t* out, int pass) {
  if (ax.walked) ax.walked[code] += (uint64_t)level_nsub(m) * proj_codewords(k - m);
  const int z = kNpts - (int)bw;
  bool done = true;
  bool to_mid = false;
  if (pass != kPassFinal && m > 1) {
    const bool norank = ax.aux[code] != 0u;
    if (norank || (z <= tmax && z < found_cert(m))) {
      done = false;
      to_mid = pass == kPassFirst && mid > 0 &&
               (level_cert(mid) <= k - 1 || (!norank && found_cert(mid) <= z));
    }
  }
  if (done) {
    out[code] = md_result((int)bw, tmax);
  } else if (to_mid) {
    ax.mid_list[atomicAdd(ax.mid_count, 1u)] = (uint32_t)code;
  } else {
    ax.fb_list[atomicAdd(ax.fb_count, 1u)] = (uint32_t)code;
  }
}

__device__ __forceinline__ void md_fold(uint64_t* state, uint32_t b, uint32_t n, uint32_t nch,
                                        int code, int k, int m, int mid, int tmax,
                                        const MdAux& ax, int32_t* out, int pass) {
  uint32_t* st = reinterpret_cast<uint32_t*>(state);
  const uint32_t prev = __hip_atomic_fetch_min(&st[0], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);
  const uint32_t done =
      __hip_atomic_fetch_add(&st[1], n + (prev >> 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + n;
  if (done == nch) {
    const uint32_t bw = __hip_atomic_load(&st[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    md_finish(code, bw, k, m, mid, tmax, ax, out, pass);
  }
}

// Unsigned division, 32-bit when both operands fit (wave-uniform operands).
__device__ __forceinline__ uint64_t udiv64(uint64_t a, uint64_t b) {
  if (((a | b) >> 32) == 0) return (uint32_t)a / (uint32_t)b;
  return a / b;
}

// The static deal of U units over W waves: wave i takes units
// [start(i), start(i + 1)), q = U / W units each and one more for the first
// r = U % W waves.
struct UnitDeal {
  uint64_t q, r;
  __device__ __forceinline__ uint64_t start(uint64_t i) const { return i * q + min(i, r); }
  __device__ __forceinline__ uint64_t wave_of(uint64_t u) const {
    const uint64_t t = r * (q + 1);
    return (u < t) ? udiv64(u, q + 1) : r + udiv64(u - t, q);
  }
};

// Per-lane minimum weight over units [o0, o1) of a code of walk dimension w
// (its pieces from the largest down, then the tiny part, which is walked by
// the run that holds its first unit).
template <class Rows>
__device__ __forceinline__ uint32_t walk_units(const Rows& row, int w, uint64_t o0, uint64_t o1,
                                               uint32_t lane) {
  uint32_t best = 64u;
  uint64_t ps = 0;
  for (int j = w - 1; j > kTinyTop; --j) {
    const uint64_t nbk = piece_blocks(j);
    const uint64_t s = max(o0, ps), e = min(o1, ps + nbk);
    if (s < e) best = min(best, gray_blocks(row, (uint32_t)j, s - ps, e - ps, lane));
    ps += nbk;
  }
  if (o0 <= ps && ps < o1) best = min(best, tiny_chunk(row, tiny_top(w), lane));
  return best;
}

// The same over units [o0, o1) of a code's nsub subcodes (each code_units(w)
// units, rows 3w apart from R): one loop over the (subcode, piece) pairs in
// unit order, subcode by subcode, each subcode's pieces from the largest
// down and then its tiny part.
__device__ __forceinline__ uint32_t walk_code_units(const uint32_t* R, int w, int nsub,
                                                    uint64_t o0, uint64_t o1, uint32_t lane) {
  const int np = max(w - 1 - kTinyTop, 0);   // pieces above the tiny part
  uint32_t best = 64u;
  uint64_t ps = 0;
  for (int t = 0; t < nsub * (np + 1) && ps < o1; ++t) {
    const int s = t / (np + 1), q = t - s * (np + 1);
    const ConstRows row{R + (size_t)s * 3 * w * kRowWords};
    if (q < np) {
      const int j = w - 1 - q;
      const uint64_t nbk = piece_blocks(j);
      const uint64_t lo = max(o0, ps), hi = min(o1, ps + nbk);
      if (lo < hi) best = min(best, gray_blocks(row, (uint32_t)j, lo - ps, hi - ps, lane));
      ps += nbk;
    } else {
      if (o0 <= ps) best = min(best, tiny_chunk(row, tiny_top(w), lane));
      ps += kTinyUnits;
    }
  }
  return best;
}

// Exact mode: static deals, no claim atomics.  Wave i of W takes
//  * the i-th of W equal shares (to one code) of the small codes, walk
//    dimension <= kSmallW, which are the plan's last buckets: each is its
//    tiny part alone, a few hundred instructions behind a chain of dependent
//    loads, so they are spread over all waves rather than lined up in a few;
//  * the i-th of W equal runs (to one unit) of the other codes' work units,
//    numbered bucket by bucket, largest walk first, each code's units
//    consecutive.
// Every wave, and so every SIMD whatever waves land on it, gets the same
// work to within a block, a small code and the set-up of a few runs.  A code
// inside the wave's share or run is finished by it: one plain store of its
// result.  A code split over runs is folded (md_fold) by each wave whose run
// meets it, counted against the number of such waves.  (Chunk claims from
// shared counters cost about 90 ns each, serialised at the device coherence
// point: 66 us of a 4096-code pass; a snake deal of whole chunks left SIMDs
// uneven by up to a 65536-codeword chunk.)
__global__ __launch_bounds__(256) void md_main_exact_kernel(
    const uint32_t* __restrict__ rows, const MdPlan* __restrict__ plan,
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ roff, uint64_t* state,
    const int32_t* __restrict__ tmax_arr, int32_t tmax_s, int32_t* __restrict__ out,
    const MdParams* __restrict__ p, MdAux ax, int32_t pass) {
  __shared__ MdShared sh;
  // The first-pass list was read by md_prepl: empty for the next call.
  if (pass == kPassFirst && blockIdx.x == 0 && threadIdx.x == 0) *ax.big_count = 0u;
  // No listed code (the usual final pass): leave before staging the plan.
  if (__builtin_amdgcn_readfirstlane(plan->pb[kMaxKExact]) == 0u) return;
  md_load_shared(sh, plan, p);
  // First bucket of small codes (k <= small_k of the pass: subcodes without
  // a piece above the tiny part).
  const int kSmallB = kMaxKExact + 1 - (int)__builtin_amdgcn_readfirstlane(sh.small_k[pass]);
  auto lvl = [&](int k) {
    return (int)__builtin_amdgcn_readfirstlane(code_level(sh.level, k, pass));
  };
  auto midl = [&](int k) {
    return (int)__builtin_amdgcn_readfirstlane(mid_level(sh.level, k));
  };
  const uint32_t lane = lane_id();
  const uint64_t* ub = reinterpret_cast<const uint64_t*>(sh.plan + offsetof(MdPlan, ub) / 4);
  const uint32_t* pb = sh.plan + offsetof(MdPlan, pb) / 4;
  auto pbu = [&](int b) { return (uint32_t)__builtin_amdgcn_readfirstlane(pb[b]); };
  const uint64_t W = (uint64_t)gridDim.x * (blockDim.x / kWave);
  // Wave index in the deal, XCD-major: workgroups go round-robin over the 8
  // XCDs (workgroup b on XCD b % 8), so numbering the waves of XCD 0 first,
  // then XCD 1, ... gives each XCD one contiguous range of runs, and a code's
  // runs (consecutive) mostly sit on one XCD, whose L2 then fetches its rows
  // once.
  constexpr uint32_t kXcds = 8;
  const uint32_t G = gridDim.x, bx = blockIdx.x, x = bx % kXcds;
  uint32_t before = 0;
  for (uint32_t y = 0; y < x; ++y) before += (G - y + kXcds - 1) / kXcds;
  const uint64_t wid = __builtin_amdgcn_readfirstlane(
      (before + bx / kXcds) * (blockDim.x / kWave) + threadIdx.x / kWave);

  // Small codes, whole.
  {
    const uint32_t P0 = pbu(kSmallB);
    const uint64_t T = pbu(kMaxKExact) - P0;
    const uint64_t q = udiv64(T, W);
    const UnitDeal d{q, T - q * W};
    const uint32_t pe = P0 + (uint32_t)d.start(wid + 1);
    int lo = kSmallB;
    for (uint32_t pos = P0 + (uint32_t)d.start(wid); pos < pe; ++pos) {
      while (pos >= pbu(lo + 1)) ++lo;
      const int k = kMaxKExact + 1 - lo, m = lvl(k), w = k - m;
      const int code = (int)order[pos];
      const uint32_t* R = rows + (size_t)roff[pos] * kRowsPerIdx * kRowWords;
      uint32_t b = 64u;
      for (int s = 0; s < level_nsub(m); ++s)
        b = min(b, tiny_chunk(ConstRows{R + (size_t)s * 3 * w * kRowWords}, tiny_top(w), lane));
      b = wave_min(b);
      if (lane == 0)
        md_finish(code, b, k, m, midl(k), tmax_arr ? tmax_arr[code] : tmax_s, ax, out, pass);
    }
  }

  // Units of the other codes.
  const uint64_t U = uniform64(ub[kSmallB]);
  const uint64_t dq = udiv64(U, W);
  const UnitDeal deal{dq, U - dq * W};
  const uint64_t u0 = deal.start(wid), u1 = deal.start(wid + 1);
  if (u0 >= u1) return;
  // Last bucket whose first unit is <= u0 (empty buckets tie with their
  // successor, so the last tie is the nonempty one).
  int lo = 0, hi = kSmallB - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ub[mid] <= u0) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  int m = lvl(kMaxKExact + 1 - lo);
  uint64_t cu = anchor_units(kMaxKExact + 1 - lo, m);
  const uint64_t rel = u0 - uniform64(ub[lo]);
  uint32_t qc = (uint32_t)udiv64(rel, cu);
  uint64_t off = rel - (uint64_t)qc * cu;
  uint64_t u = u0;
  for (;;) {
    const int w = kMaxKExact + 1 - lo - m;
    const uint32_t pos = pbu(lo) + qc;
    const int code = (int)order[pos];
    const uint32_t* R = rows + (size_t)roff[pos] * kRowsPerIdx * kRowWords;
    const uint64_t a = u - off;           // first unit of the code
    const uint64_t e = min(a + cu, u1);   // end of this wave's part of it
    const uint32_t b = wave_min(walk_code_units(R, w, level_nsub(m), off, e - a, lane));
    const int tmax = tmax_arr ? tmax_arr[code] : tmax_s;
    const int k = kMaxKExact + 1 - lo;
    if (a >= u0 && a + cu <= u1) {
      if (lane == 0) md_finish(code, b, k, m, midl(k), tmax, ax, out, pass);
    } else {
      const uint64_t first = (a < u0) ? deal.wave_of(a) : wid;
      const uint64_t last = (a + cu > u1) ? deal.wave_of(a + cu - 1) : wid;
      if (lane == 0)
        md_fold(&state[code], b, 1u, (uint32_t)(last - first + 1), code, k, m, midl(k), tmax,
                ax, out, pass);
    }
    u = e;
    if (u >= u1) break;
    off = 0;
    if (++qc == pbu(lo + 1) - pbu(lo)) {
      // Next nonempty bucket (one exists: units remain below U).
      do {
        ++lo;
      } while (pbu(lo + 1) == pbu(lo));
      qc = 0;
      m = lvl(kMaxKExact + 1 - lo);
      cu = anchor_units(kMaxKExact + 1 - lo, m);
    }
  }
}

// Thresholded mode: chunks of codes that already reached their threshold
// are skipped, so the work per chunk is not known in advance and ids are
// claimed.  A workgroup's home shard is blockIdx.x % kClaimShards (workgroups
// go round-robin over the XCDs, so a shard's counter is mostly hit from one
// XCD).  When the home shard runs dry the wave reads all shard counters at
// once (lane t reads shard t), takes the next shard in ring order that still
// has ids, and leaves once none has.  Every chunk is folded at once, because
// other waves stop early on the folded value.
//
// Every iteration has exactly one lane-0 region, at the bottom: it folds the
// finished chunk into its code and claims the next id; there is no
// `continue`, and every value that steers the loop is made wave-uniform with
// readfirstlane.  (With a separate lane-0 claim at the loop top the compiler
// threaded the two regions together and the other 63 lanes re-read a stale
// id with lane 0 inactive, walking chunk 0 forever.)
__global__ __launch_bounds__(256) void md_main_thr_kernel(
    const uint32_t* __restrict__ rows, const MdPlan* __restrict__ plan,
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ roff, uint32_t* claims,
    uint64_t* state, const int32_t* __restrict__ tmax_arr, int32_t tmax_s,
    int32_t* __restrict__ out, const MdParams* __restrict__ p, MdAux ax, int32_t pass) {
  __shared__ MdShared sh;
  if (pass == kPassFirst && blockIdx.x == 0 && threadIdx.x == 0) *ax.big_count = 0u;
  // No listed code (the usual final pass): leave before staging the plan.
  if (__builtin_amdgcn_readfirstlane(plan->pb[kMaxKExact]) == 0u) return;
  md_load_shared(sh, plan, p);
  const uint32_t lane = lane_id();
  const uint32_t total = __builtin_amdgcn_readfirstlane(sh.plan[offsetof(MdPlan, cb) / 4 + kMaxKExact]);
  uint32_t shard = blockIdx.x % kClaimShards;
  uint32_t dry = 0;                       // shards seen without ids (bit mask)
  uint32_t first = 0;
  if (lane == 0) first = atomicAdd(&claims[shard * kClaimStride], 1u);
  uint32_t g = __builtin_amdgcn_readfirstlane(first) * kClaimShards + shard;
  for (;;) {
    const bool have = g < total;
    int code = 0, tmax = 0, m = 1, k = 0;
    bool fold = false;
    uint32_t b = 64u, nch = 0;
    if (have) {
      const MdChunk ck = md_decode(sh, g, order, roff, pass);
      code = ck.code;
      nch = ck.nch;
      m = ck.m;
      k = ck.w + ck.m;
      fold = true;   // walked or skipped, the chunk is counted
      const uint32_t* R =
          rows + ((size_t)ck.roff * kRowsPerIdx + (size_t)ck.s * 3 * ck.w) * kRowWords;
      tmax = tmax_arr ? tmax_arr[code] : tmax_s;
      const int thr = 48 - tmax;
      const uint32_t* best_g = reinterpret_cast<const uint32_t*>(&state[code]);
      const uint32_t bw = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(best_g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if ((int)bw > thr) b = md_walk<true>(ck, R, lane, thr, best_g);
    } else {
      dry |= 1u << shard;
      uint32_t c = 0xFFFFFFFFu;
      if (lane < (uint32_t)kClaimShards)
        c = __hip_atomic_load(&claims[lane * kClaimStride], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
      const bool open = lane < (uint32_t)kClaimShards && !((dry >> lane) & 1u) &&
                        (uint64_t)c * kClaimShards + lane < total;
      const uint32_t m = (uint32_t)__ballot(open);
      if (m == 0u) break;
      // Next open shard after `shard` in ring order.
      const uint32_t r = ((m >> (shard + 1)) | (m << (kClaimShards - 1 - shard))) &
                         ((1u << kClaimShards) - 1u);
      shard = (shard + 1u + (uint32_t)__builtin_ctz(r)) % kClaimShards;
    }
    uint32_t next = 0;
    if (lane == 0) {
      if (fold)
        md_fold(&state[code], b, 1u, nch, code, k, m, mid_level(sh.level, k), tmax, ax, out,
                pass);
      next = atomicAdd(&claims[shard * kClaimStride], 1u);
    }
    g = __builtin_amdgcn_readfirstlane(next) * kClaimShards + shard;
  }
}

// One code, one launch (kernel.py:241-274).  The rows travel in the kernel
// arguments (no upload); each wave folds its minimum into state[0] and counts
// itself in state[1] (the minimum first, as in md_fold), and the wave that
// completes the launch writes (tag << 32 | result) to the caller's
// host-visible word in one 64-bit store and resets the state for the next
// launch, so a call is one launch and a poll of that word.  With THR, waves
// that start after the threshold is reached only count themselves.
template <bool THR>
__global__ __launch_bounds__(64) void md_single_kernel(const SingleRows rows, int w, int L,
                                                       uint32_t check_mask, int thr, int tmax,
                                                       uint32_t tag, uint32_t* state,
                                                       uint64_t* result) {
  const uint32_t lane = lane_id();
  bool walk = true;
  if (THR)
    walk = (int)__builtin_amdgcn_readfirstlane(
               __hip_atomic_load(&state[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) > thr;
  uint32_t b = 64u;
  if (walk)
    b = wave_min(run_chunk<THR>(ConstRows{rows.r}, w, L, blockIdx.x, lane, thr, check_mask,
                                &state[0]));
  if (lane == 0) {
    const uint32_t prev =
        __hip_atomic_fetch_min(&state[0], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    const uint32_t done = __hip_atomic_fetch_add(&state[1], 1u + (prev >> 31), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (done == gridDim.x) {
      const int bw = (int)__hip_atomic_load(&state[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t word = ((uint64_t)tag << 32) | (uint32_t)md_result(bw, tmax);
      __hip_atomic_store(result, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&state[0], 64u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&state[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// ------------------------------------------------------------- sampled ----

// kernel.py:10-15 (SplitMix64 finalizer, Vigna).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// x mod (2^m - 1), 3 <= m <= 63: 2^m == 1 modulo 2^m - 1, so fold m-bit digits.
__device__ __forceinline__ uint64_t mod_mersenne(uint64_t x, int m) {
  const uint64_t M = (1ull << m) - 1ull;
  while (x > M) x = (x & M) + (x >> m);
  return (x == M) ? 0ull : x;
}

constexpr int kSampledBlock = 256;

// Sampled screening (kernel.py:136-178): work_id = 1..sample_count, set_id =
// index of the code in this call, or that index modulo set_period when it is
// nonzero (a caller that hands the reference's oracle batches of set_period
// sets restarts set ids every batch).  Several workgroups share one code; the
// result min(max_zeros, tmax + 1) does not depend on their schedule.
__global__ __launch_bounds__(kSampledBlock) void md_sampled_kernel(
    const uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets, int32_t n,
    const int32_t* __restrict__ tmax_arr, int32_t tmax_s, uint64_t sample_count,
    uint64_t seed, uint32_t set_period, const uint32_t* __restrict__ gen, uint32_t bpc,
    int32_t* zmax, uint32_t* stopf) {
  __shared__ uint32_t bp[kMaxK * 8 * kRowWords];  // [i][v] -> v * g_i planes
  __shared__ int32_t red[kSampledBlock / kWave];
  const uint32_t code = blockIdx.x / bpc, part = blockIdx.x % bpc;
  const uint64_t set_id = set_period ? code % set_period : code;
  // A search generation g (gen[0] = g completed) draws with seed
  // sample_seed + (g + 1) * 1,000,003, as evolve_search.py:399-407 does per
  // generation.
  if (gen) seed += ((uint64_t)gen[0] + 1ull) * 1000003ull;
  const int k = offsets[code + 1] - offsets[code];
  if (k <= 0 || k > kMaxK) return;
  const uint32_t* R = rows + (size_t)offsets[code] * 3 * kRowWords;
  for (int e = threadIdx.x; e < k * 8; e += blockDim.x) {
    const int i = e >> 3, v = e & 7;
    RowV w = {{0, 0, 0, 0, 0, 0}};
    for (int b = 0; b < 3; ++b)
      if ((v >> b) & 1) xor_row(w, ConstRows{R}(3 * i + b));
#pragma unroll
    for (int q = 0; q < 6; ++q) bp[e * kRowWords + q] = w.w[q];
  }
  __syncthreads();
  const int tmax = tmax_arr ? tmax_arr[code] : tmax_s;
  const bool use_pid = (3 * k <= 63);
  const int m = 3 * k;
  int tz = 0;
  uint32_t it = 0;
  const uint64_t stride = (uint64_t)bpc * blockDim.x;
  const uint64_t wid0 = 1ull + (uint64_t)part * blockDim.x + threadIdx.x;
  // wid * C for the stream's two multipliers, advanced by stride * C per
  // sample (products modulo 2^64, so the same values as multiplying afresh).
  uint64_t wc = wid0 * 0xBF58476D1CE4E5B9ull, wf = wid0 * 0xA5A3564E27F8861Full;
  const uint64_t dwc = stride * 0xBF58476D1CE4E5B9ull, dwf = stride * 0xA5A3564E27F8861Full;
  for (uint64_t wid = wid0; wid <= sample_count;
       wid += stride, ++it, wc += dwc, wf += dwf) {
    if ((it & 31u) == 0 &&
        __hip_atomic_load(&stopf[code], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      break;
    RowV v = {{0, 0, 0, 0, 0, 0}};
    if (use_pid) {
      const uint64_t x = seed ^ (set_id * 0x9E3779B97F4A7C15ull) ^ wc;
      uint64_t pid = mod_mersenne(splitmix64(x), m) + 1ull;
      for (int i = 0; i < k; ++i, pid >>= 3) {
        const uint32_t* src = bp + ((i << 3) + (int)(pid & 7ull)) * kRowWords;
#pragma unroll
        for (int q = 0; q < 6; ++q) v.w[q] ^= src[q];
      }
    } else {
      const uint64_t xf = seed ^ (set_id * 0xD6E8FEB86659FD93ull) ^ wf;
      const int forced = (int)(splitmix64(xf) % (uint64_t)k);
      for (int i = 0; i < k; ++i) {
        const uint64_t x = seed ^ (set_id * 0x9E3779B97F4A7C15ull) ^ wc ^
                           ((uint64_t)(i + 1) * 0x94D049BB133111EBull);
        int c = (int)(splitmix64(x) & 7ull);
        if (i == forced && c == 0) c = 1;
        const uint32_t* src = bp + ((i << 3) + c) * kRowWords;
#pragma unroll
        for (int q = 0; q < 6; ++q) v.w[q] ^= src[q];
      }
    }
    const int zeros = kNpts - (int)weight(v);
    if (zeros > tmax) {
      atomicMax(&zmax[code], zeros);
      __hip_atomic_store(&stopf[code], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    tz = max(tz, zeros);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tz = max(tz, __shfl_xor(tz, off, kWave));
  if (lane_id() == 0) red[threadIdx.x / kWave] = tz;
  __syncthreads();
  if (threadIdx.x == 0) {
    int r = red[0];
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) r = max(r, red[w]);
    atomicMax(&zmax[code], r);
  }
}

__global__ void md_sampled_init_kernel(int32_t n, int32_t* zmax, uint32_t* stopf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { zmax[i] = 0; stopf[i] = 0u; }
}

__global__ void md_sampled_finalize_kernel(const int32_t* __restrict__ offsets, int32_t n,
                                           const int32_t* __restrict__ zmax,
                                           const int32_t* __restrict__ tmax_arr, int32_t tmax_s,
                                           int32_t* __restrict__ out, uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = offsets[i + 1] - offsets[i];
  if (k < 0 || k > kMaxK) { atomicOr(&ctrl[2], kErrBadK); out[i] = 0; return; }
  if (k == 0) { out[i] = 0; return; }
  const int tmax = tmax_arr ? tmax_arr[i] : tmax_s;
  out[i] = min(zmax[i], tmax + 1);
}

// ------------------------------------------------------------- launchers --

hipError_t launch_md_rows(const int32_t* d_idx, int64_t n_idx, const uint32_t* d_mplanes,
                          uint32_t* d_rows, uint32_t* d_ctrl, hipStream_t s) {
  if (n_idx <= 0) return hipSuccess;
  const int64_t blocks = (n_idx + 255) / 256;
  hipLaunchKernelGGL(md_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_idx, n_idx,
                     d_mplanes, d_rows, d_ctrl);
  return hipGetLastError();
}

hipError_t launch_md_fused(const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets,
                           int32_t n, const uint32_t* d_mplanes, const MdParams* d_params,
                           const int32_t* d_tmax, int32_t tmax, int32_t* d_out, const MdAux& ax,
                           uint32_t* d_order, uint32_t* d_fused_count, int kfused, int kmax,
                           uint32_t* d_ctrl, hipStream_t s, hipEvent_t ev_start,
                           hipEvent_t ev_stop) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(md_order_kernel, dim3(1), dim3(kOrderThreads), 0, s, d_offsets, n, n_idx,
                     d_tmax, tmax, d_out, ax.big_count, ax.big_list, d_order, d_fused_count,
                     kfused, kmax, ax.walked, d_ctrl);
  // About kFusedCodesPerWave codes per wave, at most kFusedMaxWaves waves.
  const int64_t waves = std::min<int64_t>(((int64_t)n + kFusedCodesPerWave - 1) /
                                          kFusedCodesPerWave, kFusedMaxWaves);
  const unsigned blocks = (unsigned)((waves + kFusedWaves - 1) / kFusedWaves);
  hipExtLaunchKernelGGL(md_fused_kernel, dim3(blocks), dim3(kFusedWaves * kWave), 0, s, ev_start,
                        ev_stop, 0, d_idx, d_offsets, (int32_t)waves, d_order, d_fused_count,
                        d_mplanes, d_params, d_tmax, tmax, d_out, ax.walked, d_ctrl);
  return hipGetLastError();
}

hipError_t launch_md_prepl(int pass, const int32_t* d_idx, int64_t n_idx,
                           const uint32_t* d_mplanes, uint32_t* d_rows, const int32_t* d_offsets,
                           int32_t n, const MdParams* d_params, MdPlan* d_plan, uint32_t* d_order,
                           uint32_t* d_roff, uint32_t* d_claims, uint64_t* d_state,
                           int32_t* d_out, const int32_t* d_tmax, int32_t tmax, const MdAux& ax,
                           uint32_t* d_ctrl, hipStream_t s) {
  hipLaunchKernelGGL(md_prepl_kernel, dim3(1 + kPrepRowBlocks), dim3(kPlanThreads), 0, s, d_idx,
                     n_idx, d_mplanes, d_rows, d_offsets, n, d_params, d_plan, d_order, d_roff,
                     d_claims, d_state, d_out, d_tmax, tmax, ax, (int32_t)pass, d_ctrl);
  return hipGetLastError();
}

int md_main_blocks_per_cu(bool thresholded) {
  int nb = 0;
  hipError_t e = thresholded
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, md_main_thr_kernel, 256, 0)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, md_main_exact_kernel, 256, 0);
  if (e != hipSuccess || nb <= 0) nb = 4;
  return nb;
}

hipError_t launch_md_main(bool thresholded, int grid, const uint32_t* d_rows,
                          const MdPlan* d_plan, const uint32_t* d_order, const uint32_t* d_roff,
                          uint32_t* d_claims, uint64_t* d_state, const int32_t* d_tmax,
                          int32_t tmax, int32_t* d_out, const MdParams* d_params, const MdAux& ax,
                          int pass, hipStream_t s, hipEvent_t ev_start, hipEvent_t ev_stop) {
  // With events, the launch records them from its own dispatch (start and
  // end of the kernel), without marker packets of their own on the stream.
  const int32_t l1 = (int32_t)pass;
  if (thresholded)
    hipExtLaunchKernelGGL(md_main_thr_kernel, dim3(grid), dim3(256), 0, s, ev_start, ev_stop, 0,
                          d_rows, d_plan, d_order, d_roff, d_claims, d_state, d_tmax, tmax, d_out,
                          d_params, ax, l1);
  else
    hipExtLaunchKernelGGL(md_main_exact_kernel, dim3(grid), dim3(256), 0, s, ev_start, ev_stop, 0,
                          d_rows, d_plan, d_order, d_roff, d_state, d_tmax, tmax, d_out, d_params,
                          ax, l1);
  return hipGetLastError();
}

hipError_t launch_md_single(bool thresholded, const SingleRows& rows, int w, int L,
                            int check_log2, int thr, int tmax, uint32_t tag, uint32_t* d_state,
                            uint64_t* h_result, hipStream_t s) {
  const uint64_t grid = code_chunks(w, L);
  if (grid == 0 || w > kMaxKExact - 1) return hipErrorInvalidValue;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  const uint32_t cmask = (1u << check_log2) - 1u;
  if (thresholded)
    hipLaunchKernelGGL(md_single_kernel<true>, dim3((unsigned)grid), dim3(64), 0, s, rows, w, L,
                       cmask, thr, tmax, tag, d_state, h_result);
  else
    hipLaunchKernelGGL(md_single_kernel<false>, dim3((unsigned)grid), dim3(64), 0, s, rows, w, L,
                       cmask, thr, tmax, tag, d_state, h_result);
  return hipGetLastError();
}

hipError_t launch_md_sampled(const uint32_t* d_rows, const int32_t* d_offsets, int32_t n,
                             const int32_t* d_tmax, int32_t tmax, int64_t sample_count,
                             uint64_t sample_seed, uint32_t set_period, const uint32_t* d_gen,
                             int32_t* d_zmax, uint32_t* d_stop, int32_t* d_out, uint32_t* d_ctrl,
                             hipStream_t s) {
  hipLaunchKernelGGL(md_sampled_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, d_zmax,
                     d_stop);
  // Workgroups per code: about 16 samples per thread, at most 64 groups.
  uint64_t bpc = (uint64_t)sample_count / (kSampledBlock * 16ull);
  if (bpc < 1) bpc = 1;
  if (bpc > 64) bpc = 64;
  while (bpc > 1 && (uint64_t)n * bpc > (1ull << 30)) bpc >>= 1;
  hipLaunchKernelGGL(md_sampled_kernel, dim3((unsigned)((uint64_t)n * bpc)), dim3(kSampledBlock),
                     0, s, d_rows, d_offsets, n, d_tmax, tmax, (uint64_t)sample_count,
                     sample_seed, set_period, d_gen, (uint32_t)bpc, d_zmax, d_stop);
  hipLaunchKernelGGL(md_sampled_finalize_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                     d_offsets, n, d_zmax, d_tmax, tmax, d_out, d_ctrl);
  return hipGetLastError();
}

}  // namespace gt
