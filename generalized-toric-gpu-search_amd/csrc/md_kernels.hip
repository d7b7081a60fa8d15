// Exact / thresholded / sampled minimum distance of GF(8) generalized toric
// codes on gfx950.
//
// Replaces eval_min_distance_batch_bp (kernel.py:97-203, kernel_cuda_bp.py:146-238)
// and eval_min_distance / eval_min_distance_single_bp (kernel.py:63-94,
// kernel_cuda_bp.py:99-127).  Work decomposition (see gt_common.h):
//
//   md_prep     workgroup 0: codes bucketed by walk dimension, largest first
//               (MdPlan, order[], roff[]), claim counters reset;
//               workgroups 1..: one wave per code, its state word reset, the
//               result of a code without a walk (k <= 1) written, and the bit
//               rows x^b * (g_i + g_{k-1}) of the subcode vanishing at the
//               torus point (1,1) (gt_common.h), which has the same maximum
//               number of zeros as C(S) with 1/8 of the codewords.  The
//               sampled path keeps the full rows x^b * g_i (md_rows).
//   md_main     persistent.  Exact mode: two static deals, one share of each
//               per wave: the codes of walk dimension <= 4 (no Gray piece),
//               whole, in equal counts; the other codes' work units (16-step
//               Gray blocks of their pieces and their tiny parts, in plan
//               order) in equal runs.  A code inside one run is written by
//               its wave, a code split over several runs is folded into its
//               state word by each and written by the last.
//               Thresholded mode: chunk ids claimed from 16 counters,
//               decoded against the plan in LDS (bucket, piece, code),
//               64 lanes x 2^L Gray steps each, every chunk folded at once
//               (other waves stop on the folded value).  Results are
//               max_zeros clamped at tmax + 1; codes without a walk
//               (k <= 1) get theirs from md_prep.
//
// Inner step (per projective codeword, all VALU; the block's rows are
// wave-uniform, loaded with scalar loads and copied to VGPRs):
//   6 v_xor_b32 (row add, VOP2 with two VGPR operands), 2 v_or3_b32 (plane
//   OR), 2 v_bcnt_u32_b32, and one v_min3_u32 per two codewords.
#include <cstddef>

#include <hip/hip_ext.h>

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
// Result of a code whose minimum weight is b (max_zeros clamped at tmax + 1).
__device__ __forceinline__ int32_t md_result(int b, int tmax) {
  return (b <= 48 - tmax) ? tmax + 1 : 49 - b;
}

// The wave of a code also resets its state word and writes the result of a
// code without a walk.
__device__ __forceinline__ void md_rows_code(int code, int32_t n, int64_t n_idx,
                                             const int32_t* __restrict__ idx,
                                             const int32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ mplanes,
                                             uint32_t* __restrict__ rows, uint64_t* state,
                                             int32_t* out, const int32_t* tmax_arr,
                                             int32_t tmax_s, uint32_t* ctrl) {
  if (code >= n) return;
  const uint32_t lane = lane_id();
  const int o = offsets[code];
  const int k = offsets[code + 1] - o;
  const bool bad_k = k < 0 || k > kMaxKExact || o < 0 || (int64_t)o + k > n_idx;
  if (lane == 0) {
    state[code] = 64ull;   // minimum weight 64, no chunk finished
    // No walk: k = 0 (and bad input) gives 0, k = 1 a monomial (weight 49).
    if (bad_k || k <= 1)
      out[code] = (!bad_k && k == 1) ? md_result(49, tmax_arr ? tmax_arr[code] : tmax_s) : 0;
  }
  if (k < 1 || bad_k) return;
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
  uint4* row = reinterpret_cast<uint4*>(rows + ((size_t)o * 3 + lane) * kRowWords);
  row[0] = make_uint4(w[0], w[1], w[2], w[3]);
  row[1] = make_uint4(w[4], w[5], 0u, 0u);
}

__global__ __launch_bounds__(256) void md_rows_kernel(const int32_t* __restrict__ idx,
                                                      int64_t n_idx,
                                                      const uint32_t* __restrict__ mplanes,
                                                      uint32_t* __restrict__ rows,
                                                      uint32_t* ctrl) {
  md_rows_one((int64_t)blockIdx.x * blockDim.x + threadIdx.x, idx, n_idx, mplanes, rows, ctrl);
}

// Largest-first plan (MdPlan, order[] = code index and roff[] = first row
// per position).  One workgroup of 256 threads takes the codes in batches of
// 4096, thread t the codes base + 256 r + t, r < 16, so every global access
// of a round is one contiguous 1 KB run (a single CU issues all of them: 64
// separate lines per instruction cost 4 us of a 10 us plan).  The offsets of
// a batch are loaded together (one memory latency) and kept in registers
// with the sizes packed five bits each.  Every thread counts its codes per
// bucket in its own LDS column (hist[b][thread]: the threads of a wave hit
// consecutive words, so the adds never conflict).  Scan threads (bucket,
// segment of 22 columns) sum their segments; one wave turns the bucket
// totals into bucket bases, chunk and unit offsets; the scan threads then
// rewrite every column entry as that thread's first position in the bucket;
// pass 2 takes each code's position with a returning add on its own column
// entry.  For a call of at most 4096 codes the code indices and first rows
// are staged in LDS by position (the rows in the column words, free by then)
// and copied out in contiguous runs; larger calls store them where they
// belong.
// Workgroups of 256 threads and at most 72 VGPRs: with md_main's five
// 88-VGPR waves on every SIMD, a wave of each still fits, so the prep of a
// batch issued on another stream runs beside a walk instead of after it.
// The column counts and the staging area are dynamic LDS (kPlanDynLds bytes
// per workgroup, given at launch): the compiler sizes registers for the
// occupancy that static LDS allows, and 29 KB of static LDS would let it
// spend more than 72.
// (Per-wave ballots over the buckets present cost 16 rounds x buckets per
// pass at this width: 10 us more per 4096 codes.)
constexpr int kPlanThreads = 256;
constexpr int kPlanRounds = 16;
constexpr int kPlanBatch = kPlanThreads * kPlanRounds;
constexpr int kPlanSegs = 12;                                             // segments per bucket
constexpr int kPlanSegLen = (kPlanThreads + kPlanSegs - 1) / kPlanSegs;   // columns per segment
constexpr int kPlanPack = 6;                                              // sizes per word
static_assert(kMaxKExact * kPlanSegs <= kPlanThreads, "one scan thread per bucket segment");
static_assert(kMaxKExact < 32, "a size fits five bits");
constexpr size_t kPlanHistWords = (size_t)kMaxKExact * kPlanThreads;
constexpr size_t kPlanDynLds = kPlanHistWords * 4 + ((size_t)kPlanBatch + 2) * 2;   // + spare slot
static_assert(kPlanBatch <= 65536, "staged code indices fit 16 bits");
static_assert(kPlanBatch + 1 <= kPlanHistWords, "staged first rows fit the column words");

// This thread's codes base + r * kPlanThreads + tid of one batch: first
// index o[r] and size k(r), 0 when the code has no walk (its bucket is
// otherwise kMaxKExact - walk_dim(k) = kMaxKExact + 1 - k, k >= 2).
struct PlanCodes {
  int32_t o[kPlanRounds];
  uint32_t kp[(kPlanRounds + kPlanPack - 1) / kPlanPack];
  __device__ __forceinline__ uint32_t k(int r) const {
    return (kp[r / kPlanPack] >> (5 * (r % kPlanPack))) & 31u;
  }
  __device__ __forceinline__ int bucket(int r) const { return kMaxKExact + 1 - (int)k(r); }
  // The bucket, or 0 (never a code's: k <= kMaxKExact) for no walk.
  __device__ __forceinline__ int slot(int r) const { return k(r) != 0u ? bucket(r) : 0; }
};

__device__ __forceinline__ void md_plan_load_batch(const int32_t* __restrict__ offsets, int32_t n,
                                                   int64_t n_idx, int base, uint32_t& err,
                                                   PlanCodes& pc) {
  const int tid = (int)threadIdx.x;
  const uint32_t lane = (uint32_t)tid % kWave;
  // offsets[i] for every code i of the thread (and for i = n: the end of
  // code n - 1); a code's end is the next lane's value, and for lane 63 the
  // value lane r of the wave loads as `nx` in round r.
  const int wnext = base + (tid - (int)lane) + kWave;   // lane 63's next code in round 0
  const int inx = wnext + (int)lane * kPlanThreads;
  const int32_t nx = (lane < (uint32_t)kPlanRounds && inx <= n) ? offsets[inx] : 0;
#pragma unroll
  for (int r = 0; r < kPlanRounds; ++r) {
    const int i = base + r * kPlanThreads + tid;
    pc.o[r] = (i <= n) ? offsets[i] : 0;
  }
#pragma unroll
  for (int q = 0; q < (kPlanRounds + kPlanPack - 1) / kPlanPack; ++q) pc.kp[q] = 0u;
#pragma unroll
  for (int r = 0; r < kPlanRounds; ++r) {
    const int i = base + r * kPlanThreads + tid;
    const int32_t down = __shfl_down(pc.o[r], 1, kWave);
    const int32_t last = __builtin_amdgcn_readlane(nx, r);
    const int32_t o1 = (lane == kWave - 1) ? last : down;
    if (i < n) {
      const int k = o1 - pc.o[r];
      // Rows and indices of the code in range.
      const bool bad = k < 0 || k > kMaxKExact || pc.o[r] < 0 || (int64_t)o1 > n_idx;
      if (bad) err |= kErrBadK;
      else if (k > 1) pc.kp[r / kPlanPack] |= (uint32_t)k << (5 * (r % kPlanPack));
    }
  }
}

__device__ __forceinline__ void md_plan_body(const int32_t* __restrict__ offsets, int32_t n,
                                             int64_t n_idx, const MdParams* __restrict__ p,
                                             MdPlan* __restrict__ plan,
                                             uint32_t* __restrict__ order,
                                             uint32_t* __restrict__ roff,
                                             uint32_t* __restrict__ claims, uint32_t* ctrl) {
  constexpr int kB = kMaxKExact;
  __shared__ uint32_t cnt[kB + 1];   // codes per bucket, then bucket bases; [kB]: positions
  __shared__ uint64_t sxc[kB], sxu[kB];
  __shared__ uint32_t segsum[kB][kPlanSegs];
  extern __shared__ uint32_t plan_lds[];
  // Per bucket and thread: its codes, then its next position.
  uint32_t(*hist)[kPlanThreads] = reinterpret_cast<uint32_t(*)[kPlanThreads]>(plan_lds);
  // Code index by position (calls of at most 4096 codes).
  uint16_t* stage = reinterpret_cast<uint16_t*>(plan_lds + kPlanHistWords);
  const int tid = threadIdx.x, wave = tid / kWave;
  const uint32_t lane = (uint32_t)(tid % kWave);
  uint32_t err = 0;
  // Wave 0's lane b reads the chunks per code of bucket b (k = kB + 1 - b;
  // bucket 0, always empty, takes k = kB's) now, with the sizes.
  const uint32_t ccb = (tid < kB) ? p->chunks[min(kB + 1 - tid, kB)] : 0u;
  if (tid < kClaimShards) claims[tid * kClaimStride] = 0u;
#pragma unroll
  for (int b = 0; b < kB; ++b) hist[b][tid] = 0u;   // own column: no barrier before own adds
  PlanCodes pc;

  // Pass 1: this thread's codes per bucket (non-returning LDS adds).  The
  // last batch (the only one when n <= 4096) stays in registers.
  for (int base = 0; base < n; base += kPlanBatch) {
    md_plan_load_batch(offsets, n, n_idx, base, err, pc);
#pragma unroll
    for (int r = 0; r < kPlanRounds; ++r)   // no walk: adds 0 to bucket 0 (never a code's)
      atomicAdd(&hist[pc.slot(r)][tid], pc.k(r) != 0u ? 1u : 0u);
  }
  __syncthreads();

  // Segment sums: scan thread (sb, sg) takes columns [c0, c0 + kPlanSegLen).
  const int sb = tid / kPlanSegs, sg = tid % kPlanSegs, c0 = sg * kPlanSegLen;
  if (sb < kB) {
    uint32_t v = 0;
#pragma unroll
    for (int c = 0; c < kPlanSegLen; ++c)
      if (c0 + c < kPlanThreads) v += hist[sb][c0 + c];
    segsum[sb][sg] = v;
  }
  __syncthreads();

  // Bucket bases, chunk and unit offsets (wave 0, lane b = bucket b).
  if (wave == 0) {
    const int b = (int)lane;
    // Chunks (saturated at 2^31) and work units per code of bucket b (walk
    // dimension kB - b).
    const uint64_t cc = ccb;
    const uint64_t cu = (b < kB) ? code_units(kB - b) : 0ull;
    uint32_t t = 0;
    if (b < kB) {
#pragma unroll
      for (int g = 0; g < kPlanSegs; ++g) t += segsum[b][g];
      cnt[b] = t;
    }
    // Chunks of the bucket, clamped at 2^31 (a plan of 2^31 chunks or more
    // is refused), so the sum of 21 cannot wrap.
    const uint64_t xc = (t == 0) ? 0ull : (cc >= (1ull << 31)) ? (1ull << 31)
                                                               : min((uint64_t)t * cc, 1ull << 31);
    // Units: exact whenever the chunk total is below 2^31 (at most 2^31
    // chunks of at most 2^23 blocks each).
    const uint64_t xu = (uint64_t)t * cu;
    if (b < kB) {
      sxc[b] = xc;
      sxu[b] = xu;
    }
    // Exclusive sums: the wave's own LDS stores are visible to its later
    // loads; the loads below b are independent of each other.
    uint64_t ec = 0, eu = 0;
    uint32_t ep = 0;
#pragma unroll
    for (int q = 0; q < kB; ++q) {   // branch-free: all 63 loads issue back to back
      const uint64_t sel = 0ull - (uint64_t)(q < b);
      ec += sxc[q] & sel;
      eu += sxu[q] & sel;
      ep += cnt[q] & (uint32_t)sel;
    }
    // A plan of 2^31 chunks or more is refused; its units are zeroed so the
    // exact walk does nothing either (the call reports the error).
    const uint64_t all_chunks = readlane64(ec + xc, kB - 1);
    const bool refused = all_chunks >= (1ull << 31);
    if (b < kB) {
      plan->ub[b] = refused ? 0ull : eu;
      plan->cb[b] = (uint32_t)min(ec, 1ull << 31);
      plan->pb[b] = ep;
      plan->cc[b] = (uint32_t)cc;
    }
    // Every lane has read cnt[] above (the loads are ordered before these
    // stores within the wave).
    if (b < kB) cnt[b] = ep;
    if (b == kB - 1) {
      const uint64_t chunks = ec + xc;
      const bool overflow = chunks >= (1ull << 31);
      plan->cb[kB] = overflow ? 0u : (uint32_t)chunks;
      plan->ub[kB] = overflow ? 0ull : eu + xu;
      plan->pb[kB] = ep + t;
      cnt[kB] = ep + t;
      ctrl[0] = 0u;
      ctrl[3] = overflow ? 0xFFFFFFFFu : (uint32_t)chunks;
      if (overflow) atomicOr(&ctrl[2], kErrBadK);
    }
  }
  __syncthreads();

  // Each column entry becomes that thread's first position in the bucket.
  if (sb < kB) {
    uint32_t at = cnt[sb];
#pragma unroll
    for (int g = 0; g < kPlanSegs; ++g) at += (g < sg) ? segsum[sb][g] : 0u;
    // Two halves of loads in flight (registers: the batch stays live).
    constexpr int kHalf = (kPlanSegLen + 1) / 2;
#pragma unroll
    for (int h = 0; h < kPlanSegLen; h += kHalf) {
      uint32_t v[kHalf];
#pragma unroll
      for (int c = 0; c < kHalf; ++c)
        v[c] = (h + c < kPlanSegLen && c0 + h + c < kPlanThreads) ? hist[sb][c0 + h + c] : 0u;
#pragma unroll
      for (int c = 0; c < kHalf; ++c) {
        if (h + c < kPlanSegLen && c0 + h + c < kPlanThreads) hist[sb][c0 + h + c] = at;
        at += v[c];
      }
    }
  }
  __syncthreads();

  // Pass 2: positions (returning adds on the own column).  One batch: still
  // in registers; more: each loaded again.
  const bool staged = n <= kPlanBatch;
  for (int base = staged ? -1 : 0;;) {
    if (base >= 0) {
      if (base >= n) break;
      uint32_t unused = 0;
      md_plan_load_batch(offsets, n, n_idx, base, unused, pc);
    }
    const int first = ((base >= 0) ? base : 0) + tid;
    if (staged) {
      // Branch-free: every returning add issues before the first result is
      // used; a code without a walk writes the spare slot kPlanBatch.
      uint32_t pos[kPlanRounds];
#pragma unroll
      for (int r = 0; r < kPlanRounds; ++r)
        pos[r] = atomicAdd(&hist[pc.slot(r)][tid], pc.k(r) != 0u ? 1u : 0u);
      uint32_t pp[kPlanRounds / 2];   // positions (<= 4096) two to a word across the barrier
#pragma unroll
      for (int r = 0; r < kPlanRounds; ++r) {
        pos[r] = pc.k(r) != 0u ? pos[r] : (uint32_t)kPlanBatch;
        stage[pos[r]] = (uint16_t)(first + r * kPlanThreads);
        if (r % 2 == 0) pp[r / 2] = pos[r];
        else pp[r / 2] |= pos[r] << 16;
      }
      // The columns are done with: their words take each code's first row
      // by position.
      __syncthreads();
#pragma unroll
      for (int r = 0; r < kPlanRounds; ++r)
        plan_lds[(pp[r / 2] >> (16 * (r % 2))) & 0xFFFFu] = (uint32_t)pc.o[r];
    } else {
#pragma unroll
      for (int r = 0; r < kPlanRounds; ++r) {
        if (pc.k(r) != 0u) {
          const uint32_t pos = atomicAdd(&hist[pc.bucket(r)][tid], 1u);
          order[pos] = (uint32_t)(first + r * kPlanThreads);
          roff[pos] = (uint32_t)pc.o[r];
        }
      }
    }
    if (base < 0) break;
    base += kPlanBatch;
  }
  if (staged) {
    __syncthreads();
    // Contiguous runs out of LDS.
    const uint32_t npos = cnt[kB];
#pragma unroll
    for (int r = 0; r < kPlanRounds; ++r) {
      const uint32_t q = (uint32_t)(r * kPlanThreads + tid);
      if (q < npos) {
        order[q] = stage[q];
        roff[q] = plan_lds[q];
      }
    }
  }
  if (err) atomicOr(&ctrl[2], err);
}

// Plan and bit rows in one launch: workgroup 0 builds the plan, workgroups
// 1.. write the rows, one wave per code (the two are independent).
__global__ __launch_bounds__(kPlanThreads, 7) void md_prep_kernel(
    const int32_t* __restrict__ idx, int64_t n_idx, const uint32_t* __restrict__ mplanes,
    uint32_t* __restrict__ rows, const int32_t* __restrict__ offsets, int32_t n,
    const MdParams* __restrict__ p, MdPlan* __restrict__ plan, uint32_t* __restrict__ order,
    uint32_t* __restrict__ roff, uint32_t* __restrict__ claims, uint64_t* __restrict__ state,
    int32_t* __restrict__ out, const int32_t* __restrict__ tmax_arr, int32_t tmax_s,
    uint32_t* ctrl) {
  if (blockIdx.x == 0) {
    md_plan_body(offsets, n, n_idx, p, plan, order, roff, claims, ctrl);
    return;
  }
  md_rows_code((int)((blockIdx.x - 1) * (kPlanThreads / kWave) + threadIdx.x / kWave), n, n_idx,
               idx, offsets, mplanes, rows, state, out, tmax_arr, tmax_s, ctrl);
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
  uint32_t nch;     // chunks of the code
  uint32_t roff;    // first row of the code (3 rows per index)
};

// Wave-uniform decode of chunk id g < total.
__device__ __forceinline__ MdChunk md_decode(const MdShared& sh, uint32_t g,
                                             const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ roff) {
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
  const uint32_t pos = __builtin_amdgcn_readfirstlane(pb[lo]) + (rel - r.c * nb);
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
// the final minimum there and writes the code's result.
__device__ __forceinline__ void md_fold(uint64_t* state, uint32_t b, uint32_t n, uint32_t nch,
                                        int tmax, int32_t* out_code) {
  uint32_t* st = reinterpret_cast<uint32_t*>(state);
  const uint32_t prev = __hip_atomic_fetch_min(&st[0], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);
  const uint32_t done =
      __hip_atomic_fetch_add(&st[1], n + (prev >> 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + n;
  if (done == nch) {
    const int bw = (int)__hip_atomic_load(&st[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *out_code = md_result(bw, tmax);
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
    const MdParams* __restrict__ p) {
  constexpr int kSmallB = kMaxKExact - kSmallW;   // first bucket of small codes
  __shared__ MdShared sh;
  md_load_shared(sh, plan, p);
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
      const int code = (int)order[pos];
      const uint32_t* R = rows + (size_t)roff[pos] * 3 * kRowWords;
      const uint32_t b = wave_min(tiny_chunk(ConstRows{R}, tiny_top(kMaxKExact - lo), lane));
      if (lane == 0) out[code] = md_result((int)b, tmax_arr ? tmax_arr[code] : tmax_s);
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
  uint64_t cu = code_units(kMaxKExact - lo);
  const uint64_t rel = u0 - uniform64(ub[lo]);
  uint32_t qc = (uint32_t)udiv64(rel, cu);
  uint64_t off = rel - (uint64_t)qc * cu;
  uint64_t u = u0;
  for (;;) {
    const int w = kMaxKExact - lo;
    const uint32_t pos = pbu(lo) + qc;
    const int code = (int)order[pos];
    const uint32_t* R = rows + (size_t)roff[pos] * 3 * kRowWords;
    const uint64_t a = u - off;           // first unit of the code
    const uint64_t e = min(a + cu, u1);   // end of this wave's part of it
    const uint32_t b = wave_min(walk_units(ConstRows{R}, w, off, e - a, lane));
    const int tmax = tmax_arr ? tmax_arr[code] : tmax_s;
    if (a >= u0 && a + cu <= u1) {
      if (lane == 0) out[code] = md_result((int)b, tmax);
    } else {
      const uint64_t first = (a < u0) ? deal.wave_of(a) : wid;
      const uint64_t last = (a + cu > u1) ? deal.wave_of(a + cu - 1) : wid;
      if (lane == 0) md_fold(&state[code], b, 1u, (uint32_t)(last - first + 1), tmax, &out[code]);
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
      cu = code_units(kMaxKExact - lo);
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
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ roff, uint32_t* claims, uint64_t* state, const int32_t* __restrict__ tmax_arr, int32_t tmax_s,
    int32_t* __restrict__ out, const MdParams* __restrict__ p) {
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
    int code = 0, tmax = 0;
    bool fold = false;
    uint32_t b = 64u, nch = 0;
    if (have) {
      const MdChunk ck = md_decode(sh, g, order, roff);
      code = ck.code;
      nch = ck.nch;
      fold = true;   // walked or skipped, the chunk is counted
      const uint32_t* R = rows + (size_t)ck.roff * 3 * kRowWords;
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
      if (fold) md_fold(&state[code], b, 1u, nch, tmax, &out[code]);
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

hipError_t launch_md_prep(const int32_t* d_idx, int64_t n_idx, const uint32_t* d_mplanes,
                          uint32_t* d_rows, const int32_t* d_offsets, int32_t n,
                          const MdParams* d_params, MdPlan* d_plan, uint32_t* d_order,
                          uint32_t* d_roff, uint32_t* d_claims, uint64_t* d_state, int32_t* d_out,
                          const int32_t* d_tmax, int32_t tmax, uint32_t* d_ctrl, hipStream_t s) {
  constexpr int kCodesPerBlock = kPlanThreads / kWave;   // one wave per code
  const int64_t row_blocks = ((int64_t)std::max(n, 0) + kCodesPerBlock - 1) / kCodesPerBlock;
  hipLaunchKernelGGL(md_prep_kernel, dim3((unsigned)(1 + row_blocks)), dim3(kPlanThreads),
                     (unsigned)kPlanDynLds, s,
                     d_idx, n_idx, d_mplanes, d_rows, d_offsets, n, d_params, d_plan, d_order,
                     d_roff, d_claims, d_state, d_out, d_tmax, tmax, d_ctrl);
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
                          int32_t tmax, int32_t* d_out, const MdParams* d_params, hipStream_t s,
                          hipEvent_t ev_start, hipEvent_t ev_stop) {
  // With events, the launch records them from its own dispatch (start and
  // end of the kernel), without marker packets of their own on the stream.
  if (thresholded)
    hipExtLaunchKernelGGL(md_main_thr_kernel, dim3(grid), dim3(256), 0, s, ev_start, ev_stop, 0,
                          d_rows, d_plan, d_order, d_roff, d_claims, d_state, d_tmax, tmax, d_out,
                          d_params);
  else
    hipExtLaunchKernelGGL(md_main_exact_kernel, dim3(grid), dim3(256), 0, s, ev_start, ev_stop, 0,
                          d_rows, d_plan, d_order, d_roff, d_state, d_tmax, tmax, d_out, d_params);
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
