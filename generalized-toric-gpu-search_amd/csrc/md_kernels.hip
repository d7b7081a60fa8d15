// Exact / thresholded / sampled minimum distance of GF(8) generalized toric
// codes on gfx950.
//
// Replaces eval_min_distance_batch_bp (kernel.py:97-203, kernel_cuda_bp.py:146-238)
// and eval_min_distance / eval_min_distance_single_bp (kernel.py:63-94,
// kernel_cuda_bp.py:99-127).  Work decomposition (see gt_common.h):
//
//   md_prep     workgroup 0: codes bucketed by walk dimension, largest first
//               (MdPlan + order[]), reset of the per-code minimum and of the
//               chunk counters; workgroups 1..: one wave per code, the bit
//               rows x^b * (g_i + g_{k-1}) of the subcode vanishing at the
//               torus point (1,1) (gt_common.h), which has the same maximum
//               number of zeros as C(S) with 1/8 of the codewords.  The
//               sampled path keeps the full rows x^b * g_i (md_rows).
//   md_main     persistent: each wave takes chunk ids (a static snake deal
//               in exact mode, claims from 16 counters in thresholded mode),
//               decodes one against the plan in LDS (bucket, piece, code),
//               walks 64 lanes x 2^L Gray steps, and folds its wave minimum
//               into the code's minimum with one atomicMin.
//   md_finalize minimum weight -> max_zeros, clamped at tmax + 1.
//
// Inner step (per projective codeword, all VALU, rows in SGPRs):
//   6 v_xor_b32 (row add), 2 v_or3_b32 (plane OR), 2 v_bcnt_u32_b32,
//   and one v_min3_u32 per two codewords.
#include <cstddef>

#include "gt_launch.h"
#include "gt_walk.h"

namespace gt {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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

// Exact-walk rows of one code (one wave): lane 3i + b (i < k - 1) writes
// x^b * (g_i + g_{k-1}) at row 3i + b of the code.  Lanes i < k check their
// index.  Codes with k < 2 have no walk; k > 21 and offsets outside
// [0, n_idx] are flagged by the plan (and no row is written for them).
__device__ __forceinline__ void md_rows_code(int code, int32_t n, int64_t n_idx,
                                             const int32_t* __restrict__ idx,
                                             const int32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ mplanes,
                                             uint32_t* __restrict__ rows, uint32_t* ctrl) {
  if (code >= n) return;
  const uint32_t lane = lane_id();
  const int o = offsets[code];
  const int k = offsets[code + 1] - o;
  if (k < 1 || k > kMaxKExact || o < 0 || (int64_t)o + k > n_idx) return;
  const uint32_t i = lane / 3u, bpow = lane - 3u * i;
  int g = 0;
  bool bad = false;
  if (i < (uint32_t)k) {
    g = idx[o + (int)i];
    bad = g < 0 || g >= kNpts;
  }
  if (__ballot(bad) != 0ull) {
    if (lane == 0) atomicOr(&ctrl[2], kErrBadIndex);
    return;
  }
  if (k < 2 || i + 1 >= (uint32_t)k) return;
  const int gl = idx[o + k - 1];
  const uint4 a0 = *reinterpret_cast<const uint4*>(mplanes + g * kRowWords);
  const uint2 a1 = *reinterpret_cast<const uint2*>(mplanes + g * kRowWords + 4);
  const uint4 b0 = *reinterpret_cast<const uint4*>(mplanes + gl * kRowWords);
  const uint2 b1 = *reinterpret_cast<const uint2*>(mplanes + gl * kRowWords + 4);
  uint32_t w[6] = {a0.x ^ b0.x, a0.y ^ b0.y, a0.z ^ b0.z, a0.w ^ b0.w, a1.x ^ b1.x, a1.y ^ b1.y};
  if (bpow >= 1u) mul_x(w);
  if (bpow >= 2u) mul_x(w);
  uint4* out = reinterpret_cast<uint4*>(rows + ((size_t)o * 3 + lane) * kRowWords);
  out[0] = make_uint4(w[0], w[1], w[2], w[3]);
  out[1] = make_uint4(w[4], w[5], 0u, 0u);
}

__global__ __launch_bounds__(256) void md_rows_kernel(const int32_t* __restrict__ idx,
                                                      int64_t n_idx,
                                                      const uint32_t* __restrict__ mplanes,
                                                      uint32_t* __restrict__ rows,
                                                      uint32_t* ctrl) {
  md_rows_one((int64_t)blockIdx.x * blockDim.x + threadIdx.x, idx, n_idx, mplanes, rows, ctrl);
}

// Largest-first plan (MdPlan).  One workgroup of 16 waves takes the codes in
// batches of 4096: their sizes are loaded together (one memory latency per
// batch) and their bucket ids kept in LDS, and the buckets present anywhere
// in the batch are collected once, so the per-wave ballot loops visit those
// buckets only (a mixed batch uses a handful of the 21).  Pass 1 sizes the
// buckets, pass 2 writes each code's position: bucket base + codes of that
// bucket in earlier rounds + in earlier waves of this round + lanes below it
// (mbcnt of the bucket's ballot).  The order inside a bucket is batch order,
// so the plan is deterministic.
constexpr int kPlanThreads = 1024;
constexpr int kPlanRounds = 4;
constexpr int kPlanBatch = kPlanThreads * kPlanRounds;

__device__ __forceinline__ void md_plan_load_batch(const int32_t* __restrict__ offsets, int32_t n,
                                                   int64_t n_idx, int base, int8_t* jb,
                                                   uint32_t* best, bool first_pass, uint32_t& err,
                                                   uint32_t& present) {
  const int tid = threadIdx.x;
  int kk[kPlanRounds];
  bool inside[kPlanRounds];
#pragma unroll
  for (int r = 0; r < kPlanRounds; ++r) {
    const int i = base + r * kPlanThreads + tid;
    const int o0 = (i < n) ? offsets[i] : 0, o1 = (i < n) ? offsets[i + 1] : 0;
    kk[r] = o1 - o0;
    inside[r] = o0 >= 0 && (int64_t)o1 <= n_idx;   // rows and indices of the code in range
  }
#pragma unroll
  for (int r = 0; r < kPlanRounds; ++r) {
    const int i = base + r * kPlanThreads + tid;
    int j = -1;
    if (i < n) {
      const int k = kk[r];
      if (k < 0 || k > kMaxKExact || !inside[r]) err |= kErrBadK;
      else if (k > 1) j = kMaxKExact - walk_dim(k);
      if (first_pass) best[i] = 64u;
    }
    jb[r * kPlanThreads + tid] = (int8_t)j;
    if (j >= 0) present |= 1u << j;
  }
}

__device__ __forceinline__ void md_plan_body(const int32_t* __restrict__ offsets, int32_t n,
                                             int64_t n_idx, const MdParams* __restrict__ p,
                                             MdPlan* __restrict__ plan,
                                             uint32_t* __restrict__ order,
                                             uint32_t* __restrict__ claims,
                                             uint32_t* __restrict__ best, uint32_t* ctrl) {
  constexpr int kB = kMaxKExact;
  constexpr int kWaves = kPlanThreads / kWave;
  __shared__ uint32_t wc[kWaves][kB];
  __shared__ uint32_t run[kB];
  __shared__ uint64_t ccs[kB];       // chunks per code of bucket b, one thread each
  __shared__ int8_t jb[kPlanBatch];  // bucket id per code of the batch, -1 for none
  __shared__ uint32_t present_s;     // buckets holding at least one code
  const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
  uint32_t err = 0;
  if (tid < kClaimShards) claims[tid * kClaimStride] = 0u;
  if (tid < kB) {
    const int w = kMaxKExact - tid;   // walk dimension of bucket tid (k = w + 1)
    ccs[tid] = code_chunks(w, p->L[min(w + 1, kMaxKExact)]);
  }
  if (tid == 0) present_s = 0u;
  __syncthreads();

  // Pass 1: lane b of each wave accumulates the size of bucket b.
  uint32_t mine = 0;
  for (int base = 0; base < n; base += kPlanBatch) {
    uint32_t pres = 0;
    md_plan_load_batch(offsets, n, n_idx, base, jb, best, true, err, pres);
    if (pres) atomicOr(&present_s, pres);
    __syncthreads();
    const uint32_t present = __builtin_amdgcn_readfirstlane(present_s);
    for (int r = 0; r < kPlanRounds; ++r) {
      const int j = jb[r * kPlanThreads + tid];
      for (uint32_t rem = present; rem != 0; rem &= rem - 1) {
        const int b = __builtin_ctz(rem);
        const uint32_t c = (uint32_t)__popcll(__ballot(j == b));
        mine += (lane == b) ? c : 0u;
      }
    }
    __syncthreads();   // jb is rewritten by the next batch
  }
  if (lane < kB) wc[wave][lane] = mine;
  __syncthreads();
  if (tid < kB) {
    uint32_t t = 0;
    for (int w = 0; w < kWaves; ++w) t += wc[w][tid];
    run[tid] = t;
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t chunks = 0;
    uint32_t pos = 0;
    for (int b = 0; b < kB; ++b) {
      const uint32_t t = run[b];
      const uint64_t cc = ccs[b];
      plan->cb[b] = (uint32_t)chunks;
      plan->pb[b] = pos;
      plan->cc[b] = (uint32_t)cc;
      run[b] = pos;
      if (t > 0 && cc >= (1ull << 31)) chunks = 1ull << 31;
      chunks += (uint64_t)t * cc;
      if (chunks > (1ull << 31)) chunks = 1ull << 31;
      pos += t;
    }
    const bool overflow = chunks >= (1ull << 31);
    plan->cb[kB] = overflow ? 0u : (uint32_t)chunks;
    ctrl[0] = 0u;
    ctrl[3] = overflow ? 0xFFFFFFFFu : (uint32_t)chunks;
    if (overflow) atomicOr(&ctrl[2], kErrBadK);
  }
  __syncthreads();
  const uint32_t present = __builtin_amdgcn_readfirstlane(present_s);

  // Pass 2: positions.  A batch of at most 4096 codes is still in jb.
  for (int base = 0; base < n; base += kPlanBatch) {
    if (n > kPlanBatch) {
      uint32_t unused = 0;
      md_plan_load_batch(offsets, n, n_idx, base, jb, best, false, err, unused);
      __syncthreads();
    }
    for (int r = 0; r < kPlanRounds; ++r) {
      const int i = base + r * kPlanThreads + tid;
      const int j = jb[r * kPlanThreads + tid];
      uint32_t rank = 0, cnt = 0;
      for (uint32_t rem = present; rem != 0; rem &= rem - 1) {
        const int b = __builtin_ctz(rem);
        const uint64_t m = __ballot(j == b);
        const uint32_t below =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        rank = (j == b) ? below : rank;
        cnt = (lane == b) ? (uint32_t)__popcll(m) : cnt;
      }
      if (lane < kB) wc[wave][lane] = cnt;
      __syncthreads();
      if (j >= 0) {
        uint32_t off = run[j] + rank;
        for (int w = 0; w < wave; ++w) off += wc[w][j];
        order[off] = (uint32_t)i;
      }
      __syncthreads();
      if (tid < kB) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += wc[w][tid];
        run[tid] += t;
      }
      __syncthreads();
    }
    if (n > kPlanBatch) __syncthreads();   // jb is rewritten by the next batch
  }
  if (err) atomicOr(&ctrl[2], err);
}

// Plan and bit rows in one launch: workgroup 0 builds the plan, workgroups
// 1.. write the rows, one wave per code (the two are independent).
__global__ __launch_bounds__(kPlanThreads) void md_prep_kernel(
    const int32_t* __restrict__ idx, int64_t n_idx, const uint32_t* __restrict__ mplanes,
    uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets, int32_t n,
    const MdParams* __restrict__ p, MdPlan* __restrict__ plan, uint32_t* __restrict__ order,
    uint32_t* __restrict__ claims, uint32_t* __restrict__ best, uint32_t* ctrl) {
  if (blockIdx.x == 0) {
    md_plan_body(offsets, n, n_idx, p, plan, order, claims, best, ctrl);
    return;
  }
  md_rows_code((int)((blockIdx.x - 1) * (kPlanThreads / kWave) + threadIdx.x / kWave), n, n_idx,
               idx, offsets, mplanes, rows, ctrl);
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
// row offset are the only global loads before the rows.
struct MdShared {
  uint32_t plan[sizeof(MdPlan) / 4];
  uint8_t L[kMaxKExact + 1];
  uint8_t check_log2[kMaxKExact + 1];
};

__device__ __forceinline__ void md_load_shared(MdShared& sh, const MdPlan* __restrict__ plan,
                                               const MdParams* __restrict__ p) {
  static_assert(sizeof(MdPlan) / 4 <= 256, "one plan word per thread");
  if (threadIdx.x < sizeof(MdPlan) / 4)
    sh.plan[threadIdx.x] = reinterpret_cast<const uint32_t*>(plan)[threadIdx.x];
  if (threadIdx.x < kMaxKExact + 1) {
    sh.L[threadIdx.x] = p->L[threadIdx.x];
    sh.check_log2[threadIdx.x] = p->check_log2[threadIdx.x];
  }
  __syncthreads();
}

struct MdChunk {
  int code;
  int w;            // walk dimension
  int L;
  uint32_t cmask;   // early-exit check mask
  int j;            // Gray piece, or tiny_top(w) for the tiny chunk
  uint32_t c;       // chunk index inside the piece
};

// Wave-uniform decode of chunk id g < total.
__device__ __forceinline__ MdChunk md_decode(const MdShared& sh, uint32_t g,
                                             const uint32_t* __restrict__ order) {
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
  r.w = kMaxKExact - lo;
  const int kq = min(r.w + 1, kMaxKExact);
  r.L = __builtin_amdgcn_readfirstlane(sh.L[kq]);
  r.cmask = (1u << __builtin_amdgcn_readfirstlane(sh.check_log2[kq])) - 1u;
  uint32_t rel = g - b0;
  const int jt = tiny_top(r.w);
  int j = r.w - 1;
  for (; j > jt; --j) {
    const uint32_t span = (uint32_t)piece_chunks(j, r.L) * nb;
    if (rel < span) break;
    rel -= span;
  }
  r.j = j;
  r.c = rel / nb;
  r.code = (int)order[__builtin_amdgcn_readfirstlane(pb[lo]) + (rel - r.c * nb)];
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

// Exact mode: a static deal, no claim atomics.  W waves; in round r wave i
// walks chunk r W + i (r even) or r W + W - 1 - i (r odd).  Sizes fall along
// the ids, so each round hands out chunks of about one size and the snake
// order evens out what the rounds leave.  (Chunk claims from shared
// counters cost about 90 ns per claim on one counter, serialised at the
// device coherence point: 66 us of a 4096-code pass.)  A wave keeps the
// running minimum of the code it walks and folds it (one atomicMin) only when
// it moves to another code or leaves.
__global__ __launch_bounds__(256) void md_main_exact_kernel(
    const uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets,
    const MdPlan* __restrict__ plan, const uint32_t* __restrict__ order, uint32_t* best_w,
    const MdParams* __restrict__ p) {
  __shared__ MdShared sh;
  md_load_shared(sh, plan, p);
  const uint32_t lane = lane_id();
  const uint32_t total = __builtin_amdgcn_readfirstlane(sh.plan[offsetof(MdPlan, cb) / 4 + kMaxKExact]);
  const uint32_t W = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid =
      __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
  int pcode = -1;
  uint32_t pmin = 64u;
  for (uint32_t r = 0;; ++r) {
    const uint64_t g64 = (uint64_t)r * W + ((r & 1u) ? (W - 1u - wid) : wid);
    if (g64 >= total) break;
    const MdChunk ck = md_decode(sh, (uint32_t)g64, order);
    const uint32_t* R = rows + (size_t)offsets[ck.code] * 3 * kRowWords;
    const uint32_t b = md_walk<false>(ck, R, lane, -1, nullptr);
    if (ck.code != pcode) {
      if (pcode >= 0 && lane == 0) atomicMin(&best_w[pcode], pmin);
      pcode = ck.code;
      pmin = b;
    } else {
      pmin = min(pmin, b);
    }
  }
  if (pcode >= 0 && lane == 0) atomicMin(&best_w[pcode], pmin);
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
    const uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets,
    const MdPlan* __restrict__ plan, const uint32_t* __restrict__ order,
    uint32_t* claims, uint32_t* best_w, const int32_t* __restrict__ tmax_arr, int32_t tmax_s,
    const MdParams* __restrict__ p) {
  __shared__ MdShared sh;
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
    int code = 0;
    bool fold = false;
    uint32_t b = 64u;
    if (have) {
      const MdChunk ck = md_decode(sh, g, order);
      code = ck.code;
      const uint32_t* R = rows + (size_t)offsets[code] * 3 * kRowWords;
      const int thr = 48 - (tmax_arr ? tmax_arr[code] : tmax_s);
      const uint32_t bw = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(&best_w[code], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if ((int)bw > thr) {
        b = md_walk<true>(ck, R, lane, thr, &best_w[code]);
        fold = true;
      }
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
      if (fold) atomicMin(&best_w[code], b);
      next = atomicAdd(&claims[shard * kClaimStride], 1u);
    }
    g = __builtin_amdgcn_readfirstlane(next) * kClaimShards + shard;
  }
}

__global__ __launch_bounds__(256) void md_finalize_kernel(const int32_t* __restrict__ offsets,
                                                          int32_t n,
                                                          const uint32_t* __restrict__ best,
                                                          const int32_t* __restrict__ tmax_arr,
                                                          int32_t tmax_s,
                                                          int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = offsets[i + 1] - offsets[i];
  if (k <= 0 || k > kMaxKExact) { out[i] = 0; return; }
  const int tmax = tmax_arr ? tmax_arr[i] : tmax_s;
  // k = 1: a monomial has weight 49 (no walk was run).
  const int b = (k == 1) ? 49 : (int)best[i];
  out[i] = (b <= 48 - tmax) ? tmax + 1 : 49 - b;
}

template <bool THR>
__global__ __launch_bounds__(64) void md_single_kernel(const uint32_t* __restrict__ R, int w, int L,
                                                       uint32_t check_mask, int thr,
                                                       uint32_t* best_out) {
  const uint32_t lane = lane_id();
  if (THR && (int)__hip_atomic_load(best_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= thr)
    return;
  uint32_t b = run_chunk<THR>(ConstRows{R}, w, L, blockIdx.x, lane, thr, check_mask, best_out);
  b = wave_min(b);
  if (lane == 0) atomicMin(best_out, b);
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

hipError_t launch_md_prep(const int32_t* d_idx, int64_t n_idx, const uint32_t* d_mplanes,
                          uint32_t* d_rows, const int32_t* d_offsets, int32_t n,
                          const MdParams* d_params, MdPlan* d_plan, uint32_t* d_order,
                          uint32_t* d_claims, uint32_t* d_best, uint32_t* d_ctrl,
                          hipStream_t s) {
  constexpr int kCodesPerBlock = kPlanThreads / kWave;   // one wave per code
  const int64_t row_blocks = ((int64_t)std::max(n, 0) + kCodesPerBlock - 1) / kCodesPerBlock;
  hipLaunchKernelGGL(md_prep_kernel, dim3((unsigned)(1 + row_blocks)), dim3(kPlanThreads), 0, s,
                     d_idx, n_idx, d_mplanes, d_rows, d_offsets, n, d_params, d_plan, d_order,
                     d_claims, d_best, d_ctrl);
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
                          const int32_t* d_offsets, const MdPlan* d_plan,
                          const uint32_t* d_order, uint32_t* d_claims, uint32_t* d_best,
                          const int32_t* d_tmax, int32_t tmax, const MdParams* d_params,
                          hipStream_t s) {
  if (thresholded)
    hipLaunchKernelGGL(md_main_thr_kernel, dim3(grid), dim3(256), 0, s, d_rows, d_offsets,
                       d_plan, d_order, d_claims, d_best, d_tmax, tmax, d_params);
  else
    hipLaunchKernelGGL(md_main_exact_kernel, dim3(grid), dim3(256), 0, s, d_rows, d_offsets,
                       d_plan, d_order, d_best, d_params);
  return hipGetLastError();
}

hipError_t launch_md_finalize(const int32_t* d_offsets, int32_t n, const uint32_t* d_best,
                              const int32_t* d_tmax, int32_t tmax, int32_t* d_out,
                              hipStream_t s) {
  hipLaunchKernelGGL(md_finalize_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_offsets, n,
                     d_best, d_tmax, tmax, d_out);
  return hipGetLastError();
}

hipError_t launch_md_single(bool thresholded, const uint32_t* d_rows, int w, int L,
                            int check_log2, int thr, uint32_t* d_best, hipStream_t s) {
  const uint64_t grid = code_chunks(w, L);
  if (grid == 0) return hipSuccess;
  if (grid >= (1ull << 31)) return hipErrorInvalidValue;
  const uint32_t cmask = (1u << check_log2) - 1u;
  if (thresholded)
    hipLaunchKernelGGL(md_single_kernel<true>, dim3((unsigned)grid), dim3(64), 0, s, d_rows, w,
                       L, cmask, thr, d_best);
  else
    hipLaunchKernelGGL(md_single_kernel<false>, dim3((unsigned)grid), dim3(64), 0, s, d_rows, w,
                       L, cmask, thr, d_best);
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
