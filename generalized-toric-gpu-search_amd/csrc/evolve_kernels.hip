// Evolutionary search, one generation on the device (the population stays in
// HBM between generations; a captured graph replays the whole generation).
//
// Plays the role of the host loop of run_evolution (evolve_search.py:322-539)
// with exact scores: propose -> canonical forms -> first-occurrence dedup
// against every class evaluated so far -> choose -> canonical forms of the
// chosen -> min distance -> survivors.  This file holds the steps that are
// specific to the search; the others are the canonical, hash-set and
// min-distance kernels.
#include "gt_bfs.h"
#include "gt_evolve.h"

namespace gt {

namespace {

constexpr uint64_t kFull49 = (1ull << kNpts) - 1ull;

// Position of the t-th (0-based, ascending) set bit of m; t < popcount(m).
__device__ __forceinline__ int nth_bit(uint64_t m, int t) {
  for (int s = 0; s < t; ++s) m &= m - 1;
  return __builtin_ctzll(m);
}

// Uniform random k-subset of the 49 points (Floyd's algorithm): for j =
// 49 - k .. 48 draw r in [0, j]; add r, or j when r is already in.
__device__ __forceinline__ uint64_t random_set(uint64_t seed, uint32_t g, uint32_t i, int k) {
  uint64_t m = 0;
  for (int j = kNpts - k, t = 0; j < kNpts; ++j, ++t) {
    const uint32_t r = evo_below(evo_rand(seed, g, i, (uint32_t)t), (uint32_t)(j + 1));
    m |= ((m >> r) & 1ull) ? (1ull << j) : (1ull << r);
  }
  return m;
}

// Order-reversing map of fp32 scores to uint32: larger score -> smaller key.
__device__ __forceinline__ uint32_t score_desc_key(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t asc = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ~asc;
}

GT_HD int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

}  // namespace

__global__ __launch_bounds__(256) void evo_propose_kernel(
    const uint64_t* __restrict__ pop_m, const int32_t* __restrict__ pop_d,
    const int32_t* __restrict__ pop_age, int32_t pop, int32_t n_prop, uint64_t seed,
    const uint32_t* __restrict__ state, uint64_t* __restrict__ props) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_prop) return;
  const uint32_t g = state[0];
  int w = 0;
  int64_t best = 0;
  for (int t = 0; t < 4; ++t) {
    const int c = (int)evo_below(evo_rand(seed, g, (uint32_t)i, (uint32_t)t), (uint32_t)pop);
    const int64_t sc = (int64_t)pop_d[c] * 1000 - pop_age[c];
    if (t == 0 || sc > best) {
      best = sc;
      w = c;
    }
  }
  const uint64_t m = pop_m[w];
  const int k = __popcll(m);
  const uint64_t comp = ~m & kFull49;
  // Sizes are checked on the host (2 <= k <= 47), so every draw below has a
  // nonempty range.
  const int p1 = nth_bit(m, (int)evo_below(evo_rand(seed, g, i, 4), (uint32_t)k));
  const int q1 = nth_bit(comp, (int)evo_below(evo_rand(seed, g, i, 5), (uint32_t)(kNpts - k)));
  const int p2 = nth_bit(m & ~(1ull << p1),
                         (int)evo_below(evo_rand(seed, g, i, 6), (uint32_t)(k - 1)));
  const int q2 = nth_bit(comp & ~(1ull << q1),
                         (int)evo_below(evo_rand(seed, g, i, 7), (uint32_t)(kNpts - 1 - k)));
  uint64_t child = m ^ (1ull << p1) ^ (1ull << q1);
  if (evo_rand(seed, g, i, 8) >> 63) child ^= (1ull << p2) ^ (1ull << q2);
  props[i] = child;
}

// Exclusive scan of v over a workgroup of 1024 threads (lanes, then waves);
// wsum: 17 words of LDS.  Returns the exclusive prefix; *total gets the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
  uint32_t inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; ++w) {
      const uint32_t u = wsum[w];
      wsum[w] = run;
      run += u;
    }
    wsum[16] = run;
  }
  __syncthreads();
  const uint32_t r = wsum[wave] + inc - v;
  *total = wsum[16];
  __syncthreads();   // wsum is reused by the next scan
  return r;
}

// With scores and more new proposals than pop: the pop smallest keys
// (descending score, then proposal order) are found by a radix select over
// the 32-bit score keys (three passes of 11, 11 and 10 bits with an LDS
// histogram give the pop-th smallest key T and how many keys equal to T to
// take), and the chosen sets are written in proposal order.
__global__ __launch_bounds__(1024) void evo_select_kernel(
    const uint64_t* __restrict__ props, const uint32_t* __restrict__ keep,
    const uint32_t* __restrict__ count, const float* __restrict__ score, int32_t pop,
    int32_t n_prop, int32_t k, uint64_t seed, const uint32_t* __restrict__ state,
    uint64_t* __restrict__ chosen) {
  extern __shared__ uint32_t key[];        // cnt score keys
  __shared__ uint32_t hist[2048];
  __shared__ uint32_t wsum[17];
  __shared__ uint32_t sel_bin, sel_below;
  const uint32_t g = state[0];
  const int cnt = min((int)*count, n_prop);
  const int real = min(cnt, pop);
  if (score != nullptr && cnt > pop) {
    for (int j = threadIdx.x; j < cnt; j += blockDim.x) key[j] = score_desc_key(score[keep[j]]);
    uint32_t prefix = 0, pmask = 0, need = (uint32_t)pop;
    const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
    for (int pass = 0; pass < 3; ++pass) {
      const int sh = shifts[pass];
      const uint32_t bm = (1u << widths[pass]) - 1u;
      for (int b = threadIdx.x; b < 2048; b += blockDim.x) hist[b] = 0u;
      __syncthreads();
      for (int j = threadIdx.x; j < cnt; j += blockDim.x) {
        const uint32_t x = key[j];
        if ((x & pmask) == prefix) atomicAdd(&hist[(x >> sh) & bm], 1u);
      }
      __syncthreads();
      const uint32_t h0 = hist[2 * threadIdx.x], h1 = hist[2 * threadIdx.x + 1];
      uint32_t tot;
      const uint32_t ex = block_excl_scan(h0 + h1, wsum, &tot);
      if (ex < need && need <= ex + h0) {
        sel_bin = 2u * threadIdx.x;
        sel_below = ex;
      } else if (ex + h0 < need && need <= ex + h0 + h1) {
        sel_bin = 2u * threadIdx.x + 1u;
        sel_below = ex + h0;
      }
      __syncthreads();
      need -= sel_below;
      prefix |= sel_bin << sh;
      pmask |= bm << sh;
      __syncthreads();   // sel_* are rewritten by the next pass
    }
    // Keys < prefix are taken, and the first `need` keys equal to it.
    const int per = (cnt + (int)blockDim.x - 1) / (int)blockDim.x;   // <= 8
    const int i0 = (int)threadIdx.x * per;
    uint32_t eq = 0;
    for (int r = 0; r < per && i0 + r < cnt; ++r) eq += (key[i0 + r] == prefix) ? 1u : 0u;
    uint32_t tot;
    uint32_t eq_before = block_excl_scan(eq, wsum, &tot);
    uint32_t bits = 0, ns = 0;
    for (int r = 0; r < per && i0 + r < cnt; ++r) {
      const uint32_t x = key[i0 + r];
      bool take = x < prefix;
      if (x == prefix) take = (eq_before++ < need);
      if (take) {
        bits |= 1u << r;
        ++ns;
      }
    }
    uint32_t pos = block_excl_scan(ns, wsum, &tot);
    for (int r = 0; r < per; ++r)
      if ((bits >> r) & 1u) chosen[pos++] = props[keep[i0 + r]];
  } else {
    for (int j = threadIdx.x; j < real; j += blockDim.x) chosen[j] = props[keep[j]];
  }
  for (int j = real + threadIdx.x; j < pop; j += blockDim.x)
    chosen[j] = random_set(seed, g, (uint32_t)(n_prop + j), k);
}

// First-occurrence dedup of one generation's canonical masks in one
// workgroup: a table of slot -> smallest index in LDS (linear probing, keys
// compared through pcm), then "first occurrence and not in the evaluated set"
// (read-only probe of the device hash set), then ordered compaction.
__global__ __launch_bounds__(1024) void evo_dedup_kernel(
    const uint64_t* __restrict__ pcm, int32_t n, int32_t slots,
    const uint64_t* __restrict__ seen, uint64_t seen_mask, uint32_t* __restrict__ keep,
    uint32_t* __restrict__ count) {
  extern __shared__ uint32_t slot[];
  __shared__ uint32_t wbase[17];
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  const uint32_t smask = (uint32_t)slots - 1u;
  for (int h = threadIdx.x; h < slots; h += blockDim.x) slot[h] = kEmpty;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint64_t key = pcm[i];
    uint32_t h = (uint32_t)mix64(key) & smask;
    for (int probe = 0; probe < slots; ++probe) {
      uint32_t cur = slot[h];
      if (cur == kEmpty) {
        cur = atomicCAS(&slot[h], kEmpty, (uint32_t)i);
        if (cur == kEmpty) break;
      }
      if (pcm[cur] == key) {
        atomicMin(&slot[h], (uint32_t)i);
        break;
      }
      h = (h + 1u) & smask;
    }
  }
  __syncthreads();
  // Thread t owns items [t * per, t * per + per): contiguous, so the block
  // scan of the per-thread counts gives proposal order.
  const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;   // <= 8 (n <= 8192)
  const int i0 = (int)threadIdx.x * per;
  uint32_t bits = 0, cnt = 0;
  for (int r = 0; r < per; ++r) {
    const int i = i0 + r;
    if (i >= n) break;
    const uint64_t key = pcm[i];
    uint32_t h = (uint32_t)mix64(key) & smask;
    bool first = false;
    for (int probe = 0; probe < slots; ++probe) {
      const uint32_t cur = slot[h];
      if (cur == kEmpty) break;                  // not reached: every key was inserted
      if (pcm[cur] == key) {
        first = (cur == (uint32_t)i);
        break;
      }
      h = (h + 1u) & smask;
    }
    if (first) {
      uint64_t g = mix64(key) & seen_mask;
      for (uint64_t probe = 0; probe <= seen_mask; ++probe) {
        const uint64_t cur = seen[g];
        if (cur == key) {
          first = false;
          break;
        }
        if (cur == 0) break;
        g = (g + 1) & seen_mask;
      }
    }
    if (first) {
      bits |= 1u << r;
      ++cnt;
    }
  }
  // Exclusive scan of cnt over the workgroup (lanes, then waves).
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
  uint32_t inc = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(inc, o, 64);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wbase[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; ++w) {
      const uint32_t v = wbase[w];
      wbase[w] = run;
      run += v;
    }
    wbase[16] = run;
    *count = run;
  }
  __syncthreads();
  uint32_t pos = wbase[wave] + inc - cnt;
  for (int r = 0; r < per; ++r)
    if ((bits >> r) & 1u) keep[pos++] = (uint32_t)(i0 + r);
}

// Survivors by counting.  Parents are sorted by (distance desc, age asc,
// position) -- the initial population is put in that order here too -- and
// children have age 0, so within one distance the children come first in
// their own order and then the parents in theirs.  Destinations:
//   child j, distance d:  P(> d) + C(> d) + (children with d before j)
//   parent i, distance d: i + C(>= d)
// with P/C the parent/child counts by distance.  Writing only destinations
// below pop gives the best pop of both.  n_par = 0 sorts the initial
// population (the children) and logs it in slot 0.
__global__ __launch_bounds__(1024) void evo_survive_kernel(
    uint64_t* __restrict__ pop_m, int32_t* __restrict__ pop_d, int32_t* __restrict__ pop_age,
    const uint64_t* __restrict__ chosen, const uint64_t* __restrict__ chosen_cls,
    const int32_t* __restrict__ chosen_mz, const uint32_t* __restrict__ count, int32_t pop,
    int32_t n_par, int32_t max_gen, uint32_t* __restrict__ state, uint64_t* __restrict__ log_cls,
    int32_t* __restrict__ log_d, int32_t* __restrict__ log_best, int32_t* __restrict__ log_new) {
  constexpr int kD = kNpts + 1;                        // distances 0..49
  constexpr int kChunks = kEvoMaxPop / kWave;          // 64 children per chunk
  __shared__ uint16_t chunk_cnt[kChunks * kD];         // then: children of d in earlier chunks
  __shared__ uint32_t cc[kD], pc[kD], cgt[kD], pgt[kD];
  __shared__ int best_d;
  const uint32_t g = state[0];
  const bool init = n_par == 0;
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
  const int nwaves = (int)(blockDim.x >> 6);
  const int nchunks = (pop + kWave - 1) / kWave;
  for (int t = threadIdx.x; t < kChunks * kD; t += blockDim.x) chunk_cnt[t] = 0;
  if (threadIdx.x < kD) pc[threadIdx.x] = 0;
  __syncthreads();
  // Children: in-chunk stable rank by one ballot per distinct distance.
  int cd[4], crank[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int chunk = wave + r * nwaves;
    const int j = chunk * kWave + lane;
    const bool valid = chunk < nchunks && j < pop;
    cd[r] = valid ? min(max(kNpts - chosen_mz[j], 0), kNpts) : -1;
    crank[r] = 0;
    if (chunk < nchunks) {
      uint64_t todo = __ballot(valid);
      while (todo) {
        const int leader = __builtin_ctzll(todo);
        const int v = __shfl(cd[r], leader, 64);
        const uint64_t m = __ballot(cd[r] == v);
        if (cd[r] == v) crank[r] = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == leader) chunk_cnt[chunk * kD + v] = (uint16_t)__popcll(m);
        todo &= ~m;
      }
    }
  }
  int pd[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = (int)threadIdx.x + r * (int)blockDim.x;
    pd[r] = (i < n_par) ? min(max(pop_d[i], 0), kNpts) : -1;
    if (pd[r] >= 0) atomicAdd(&pc[pd[r]], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kD) {
    uint32_t run = 0;
    for (int c = 0; c < nchunks; ++c) {
      const uint32_t v = chunk_cnt[c * kD + threadIdx.x];
      chunk_cnt[c * kD + threadIdx.x] = (uint16_t)run;
      run += v;
    }
    cc[threadIdx.x] = run;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t rc = 0, rp = 0;
    int top = -1;
    for (int d = kNpts; d >= 0; --d) {
      cgt[d] = rc;
      pgt[d] = rp;
      rc += cc[d];
      rp += pc[d];
      if (top < 0 && cc[d] + pc[d] != 0) top = d;
    }
    best_d = top;
  }
  __syncthreads();
  // Read every source before any write: parents move within the arrays
  // being written.
  uint64_t vm[8];
  int vd[8], va[8], vpos[8];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    vpos[r] = -1;
    if (cd[r] >= 0) {
      const int chunk = wave + r * nwaves, j = chunk * kWave + lane, d = cd[r];
      vpos[r] = (int)(pgt[d] + cgt[d] + chunk_cnt[chunk * kD + d]) + crank[r];
      vm[r] = chosen[j];
      vd[r] = d;
      va[r] = 0;
    }
    vpos[4 + r] = -1;
    if (pd[r] >= 0) {
      const int i = (int)threadIdx.x + r * (int)blockDim.x, d = pd[r];
      vpos[4 + r] = i + (int)(cgt[d] + cc[d]);
      vm[4 + r] = pop_m[i];
      vd[4 + r] = d;
      va[4 + r] = pop_age[i] + 1;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    if (vpos[r] >= 0 && vpos[r] < pop) {
      pop_m[vpos[r]] = vm[r];
      pop_d[vpos[r]] = vd[r];
      pop_age[vpos[r]] = va[r];
    }
  }
  const uint32_t slot = init ? 0u : g + 1;
  if (slot <= (uint32_t)max_gen) {
    for (int j = threadIdx.x; j < pop; j += blockDim.x) {
      log_cls[(size_t)slot * pop + j] = chosen_cls[j];
      log_d[(size_t)slot * pop + j] = kNpts - chosen_mz[j];
    }
    if (threadIdx.x == 0) {
      log_best[slot] = best_d;
      log_new[slot] = init ? pop : (int)*count;
    }
  }
  if (threadIdx.x == 0) state[0] = slot;
}

// ------------------------------------------------------------- launchers --

hipError_t launch_evo_propose(const uint64_t* pop_m, const int32_t* pop_d, const int32_t* pop_age,
                              int32_t pop, int32_t n_prop, uint64_t seed, const uint32_t* state,
                              uint64_t* props, hipStream_t s) {
  if (n_prop <= 0) return hipSuccess;
  hipLaunchKernelGGL(evo_propose_kernel, dim3((n_prop + 255) / 256), dim3(256), 0, s, pop_m,
                     pop_d, pop_age, pop, n_prop, seed, state, props);
  return hipGetLastError();
}

hipError_t launch_evo_select(const uint64_t* props, const uint32_t* keep, const uint32_t* count,
                             const float* score, int32_t pop, int32_t n_prop, int32_t k,
                             uint64_t seed, const uint32_t* state, uint64_t* chosen,
                             hipStream_t s) {
  if (pop <= 0 || pop > kEvoMaxPop || n_prop <= 0 || n_prop > kEvoMaxProp)
    return hipErrorInvalidValue;
  const size_t lds = score ? (size_t)n_prop * 4 : 0;
  hipLaunchKernelGGL(evo_select_kernel, dim3(1), dim3(1024), lds, s, props, keep, count, score,
                     pop, n_prop, k, seed, state, chosen);
  return hipGetLastError();
}

hipError_t launch_evo_dedup(const uint64_t* pcm, int32_t n, const uint64_t* seen, uint64_t seen_cap,
                            uint32_t* keep, uint32_t* count, hipStream_t s) {
  if (n <= 0 || n > kEvoMaxProp) return hipErrorInvalidValue;
  const int slots = pow2_at_least(2 * n);
  hipLaunchKernelGGL(evo_dedup_kernel, dim3(1), dim3(1024), (size_t)slots * 4, s, pcm, n, slots,
                     seen, seen_cap - 1, keep, count);
  return hipGetLastError();
}

hipError_t launch_evo_survive(uint64_t* pop_m, int32_t* pop_d, int32_t* pop_age,
                              const uint64_t* chosen, const uint64_t* chosen_cls,
                              const int32_t* chosen_mz, const uint32_t* count, int32_t pop,
                              bool init, int32_t max_gen, uint32_t* state, uint64_t* log_cls,
                              int32_t* log_d, int32_t* log_best, int32_t* log_new,
                              hipStream_t s) {
  if (pop <= 0 || pop > kEvoMaxPop) return hipErrorInvalidValue;
  hipLaunchKernelGGL(evo_survive_kernel, dim3(1), dim3(1024), 0, s, pop_m, pop_d, pop_age,
                     chosen, chosen_cls, chosen_mz, count, pop, init ? 0 : pop, max_gen, state,
                     log_cls, log_d, log_best, log_new);
  return hipGetLastError();
}

}  // namespace gt
