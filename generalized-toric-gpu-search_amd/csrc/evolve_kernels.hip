// Evolutionary search, one generation on the device (the population stays in
// HBM between generations; a captured graph replays the whole generation).
//
// Plays the role of the host loop of run_evolution (evolve_search.py:322-539)
// with exact scores: propose -> canonical forms -> first-occurrence dedup
// against every class evaluated so far -> choose -> canonical forms of the
// chosen -> min distance -> survivors.  This file holds the steps that are
// specific to the search; the others are the canonical, hash-set and
// min-distance kernels.
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

// Ascending bitonic sort of n (a power of two) keys in LDS by one workgroup.
__device__ void bitonic_sort(uint64_t* a, int n) {
  for (int size = 2; size <= n; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
        const int lo = 2 * t - (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t x = a[lo], y = a[hi];
        if ((x > y) == up) {
          a[lo] = y;
          a[hi] = x;
        }
      }
      __syncthreads();
    }
  }
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

__global__ __launch_bounds__(1024) void evo_select_kernel(
    const uint64_t* __restrict__ props, const uint32_t* __restrict__ keep,
    const uint32_t* __restrict__ count, const float* __restrict__ score, int32_t pop,
    int32_t n_prop, int32_t k, uint64_t seed, const uint32_t* __restrict__ state,
    uint64_t* __restrict__ chosen) {
  extern __shared__ uint64_t keys[];
  const uint32_t g = state[0];
  const int cnt = min((int)*count, n_prop);
  const int real = min(cnt, pop);
  if (score != nullptr && cnt > pop) {
    // Key = (descending score, proposal order); the pop smallest keys win.
    const int n2 = pow2_at_least(cnt);
    for (int j = threadIdx.x; j < n2; j += blockDim.x)
      keys[j] = (j < cnt) ? (((uint64_t)score_desc_key(score[keep[j]]) << 32) | (uint32_t)j)
                          : ~0ull;
    __syncthreads();
    bitonic_sort(keys, n2);
    for (int j = threadIdx.x; j < pop; j += blockDim.x)
      chosen[j] = props[keep[(uint32_t)keys[j]]];
  } else {
    for (int j = threadIdx.x; j < real; j += blockDim.x) chosen[j] = props[keep[j]];
  }
  for (int j = real + threadIdx.x; j < pop; j += blockDim.x)
    chosen[j] = random_set(seed, g, (uint32_t)(n_prop + j), k);
}

__global__ __launch_bounds__(1024) void evo_survive_kernel(
    uint64_t* __restrict__ pop_m, int32_t* __restrict__ pop_d, int32_t* __restrict__ pop_age,
    const uint64_t* __restrict__ chosen, const uint64_t* __restrict__ chosen_cls,
    const int32_t* __restrict__ chosen_mz, const uint32_t* __restrict__ count, int32_t pop,
    int32_t max_gen, uint32_t* __restrict__ state, uint64_t* __restrict__ log_cls,
    int32_t* __restrict__ log_d, int32_t* __restrict__ log_best, int32_t* __restrict__ log_new) {
  extern __shared__ uint64_t keys[];
  const uint32_t g = state[0];
  const int n = 2 * pop;
  const int n2 = pow2_at_least(n);
  // Key = (64 - d, age, position): ascending order is best distance first,
  // then youngest, then parents before children in their own order.
  for (int j = threadIdx.x; j < n2; j += blockDim.x) {
    uint64_t key = ~0ull;
    if (j < n) {
      int d, age;
      if (j < pop) {
        d = pop_d[j];
        age = pop_age[j] + 1;
      } else {
        d = kNpts - chosen_mz[j - pop];
        age = 0;
      }
      key = ((uint64_t)(64 - d) << 56) | ((uint64_t)(uint32_t)age << 24) | (uint64_t)j;
    }
    keys[j] = key;
  }
  __syncthreads();
  bitonic_sort(keys, n2);
  // Gather every survivor before any write: parents are read from the
  // arrays being overwritten.  1024 threads x 4 covers kEvoMaxPop.
  uint64_t sm[4];
  int sd[4], sa[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = threadIdx.x + r * (int)blockDim.x;
    sm[r] = 0;
    sd[r] = sa[r] = 0;
    if (j < pop) {
      const uint64_t key = keys[j];
      const int src = (int)(key & 0xFFFFFFu);
      sm[r] = (src < pop) ? pop_m[src] : chosen[src - pop];
      sd[r] = 64 - (int)(key >> 56);
      sa[r] = (int)((key >> 24) & 0xFFFFFFFFu);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = threadIdx.x + r * (int)blockDim.x;
    if (j < pop) {
      pop_m[j] = sm[r];
      pop_d[j] = sd[r];
      pop_age[j] = sa[r];
    }
  }
  const uint32_t slot = g + 1;
  if (slot <= (uint32_t)max_gen) {
    for (int j = threadIdx.x; j < pop; j += blockDim.x) {
      log_cls[(size_t)slot * pop + j] = chosen_cls[j];
      log_d[(size_t)slot * pop + j] = kNpts - chosen_mz[j];
    }
    if (threadIdx.x == 0) {
      log_best[slot] = 64 - (int)(keys[0] >> 56);
      log_new[slot] = (int)*count;
    }
  }
  if (threadIdx.x == 0) state[0] = slot;
}

__global__ __launch_bounds__(1024) void evo_init_kernel(
    const uint64_t* __restrict__ cls, const int32_t* __restrict__ mz, int32_t pop,
    int32_t* __restrict__ pop_d, int32_t* __restrict__ pop_age, uint32_t* __restrict__ state,
    uint64_t* __restrict__ log_cls, int32_t* __restrict__ log_d, int32_t* __restrict__ log_best,
    int32_t* __restrict__ log_new) {
  __shared__ int best;
  if (threadIdx.x == 0) best = 0;
  __syncthreads();
  int local = 0;
  for (int j = threadIdx.x; j < pop; j += blockDim.x) {
    const int d = kNpts - mz[j];
    pop_d[j] = d;
    pop_age[j] = 0;
    log_cls[j] = cls[j];
    log_d[j] = d;
    local = max(local, d);
  }
  atomicMax(&best, local);
  __syncthreads();
  if (threadIdx.x == 0) {
    log_best[0] = best;
    log_new[0] = pop;
    state[0] = 0;
  }
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
  const size_t lds = score ? (size_t)pow2_at_least(n_prop) * 8 : 0;
  hipLaunchKernelGGL(evo_select_kernel, dim3(1), dim3(1024), lds, s, props, keep, count, score,
                     pop, n_prop, k, seed, state, chosen);
  return hipGetLastError();
}

hipError_t launch_evo_survive(uint64_t* pop_m, int32_t* pop_d, int32_t* pop_age,
                              const uint64_t* chosen, const uint64_t* chosen_cls,
                              const int32_t* chosen_mz, const uint32_t* count, int32_t pop,
                              int32_t max_gen, uint32_t* state, uint64_t* log_cls,
                              int32_t* log_d, int32_t* log_best, int32_t* log_new,
                              hipStream_t s) {
  if (pop <= 0 || pop > kEvoMaxPop) return hipErrorInvalidValue;
  const size_t lds = (size_t)pow2_at_least(2 * pop) * 8;
  hipLaunchKernelGGL(evo_survive_kernel, dim3(1), dim3(1024), lds, s, pop_m, pop_d, pop_age,
                     chosen, chosen_cls, chosen_mz, count, pop, max_gen, state, log_cls, log_d,
                     log_best, log_new);
  return hipGetLastError();
}

hipError_t launch_evo_init(const uint64_t* cls, const int32_t* mz, int32_t pop, int32_t* pop_d,
                           int32_t* pop_age, uint32_t* state, uint64_t* log_cls, int32_t* log_d,
                           int32_t* log_best, int32_t* log_new, hipStream_t s) {
  if (pop <= 0 || pop > kEvoMaxPop) return hipErrorInvalidValue;
  hipLaunchKernelGGL(evo_init_kernel, dim3(1), dim3(1024), 0, s, cls, mz, pop, pop_d, pop_age,
                     state, log_cls, log_d, log_best, log_new);
  return hipGetLastError();
}

}  // namespace gt
