// Continuously batched min-distance serving: one persistent kernel.
//
// Requests (k <= kServeMaxK exponent points and a threshold tmax) arrive in a
// ring in fine-grained host memory.  Wave 0 of workgroup 0 is the dispatcher:
// it polls the ring's publish counter, builds the bit rows of each new request
// into the request's slot (slot = request index mod slots), appends the
// request's chunk range to the global chunk sequence and publishes the new
// chunk total.  Every other wave is a worker: it claims the next global chunk
// id with one atomic, waits until that id is published, finds the request by
// binary search over the in-flight window, walks the chunk (the Gray walk of
// the batch kernel, rows read through coherent loads), and folds the wave
// minimum into the slot.  The wave that finishes the last chunk of a request
// writes its result and sequence number to host memory.
//
// Requests that arrive while others run are admitted as soon as a slot and
// the host-side work budget allow; their chunks queue behind the chunks
// already published, so work of different requests overlaps on the device.
//
// Window invariant (kept by the host): request r is published only when
// r - (oldest request whose result the host has not consumed) < slots.  Hence
// slot r % slots is free, and ord_start entries of the last `slots`
// dispatched requests cannot be overwritten while any of them is in flight
// (the ring holds 4 * slots entries).
//
// Exit: every wave leaves when the host sets `shutdown`, when the dispatcher
// has been idle for idle_ticks of the device wall clock, or after life_ticks;
// workers wait on nothing but the published chunk total and the stop flag.
#include "gt_serve.h"
#include "gt_walk.h"

namespace gt {

template <class T>
__device__ __forceinline__ T ld_acq_sys(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ __forceinline__ T ld_rlx(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void dispatcher(ServeDev* d, ServeHost* h, const uint32_t* __restrict__ mplanes,
                           uint32_t lane) {
  const uint64_t t_start = wall_clock64();
  uint64_t t_idle = t_start;
  uint64_t disp = 0;          // requests dispatched so far
  uint64_t total = 0;         // chunks published so far
  const uint32_t slots = d->slots;
  const uint64_t ord_mask = 4ull * slots - 1ull;
  for (;;) {
    const uint64_t now = wall_clock64();
    const bool stop = __builtin_amdgcn_readfirstlane(ld_acq_sys(&h->shutdown)) != 0u ||
                      now - t_idle > d->idle_ticks || now - t_start > d->life_ticks;
    if (stop) {
      if (lane == 0) {
        __hip_atomic_store(&d->shutdown, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&h->exited, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      return;
    }
    const uint64_t pub = ld_acq_sys(&h->published);
    const uint64_t pubu = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(pub >> 32)) << 32) |
                          __builtin_amdgcn_readfirstlane((uint32_t)pub);
    if (pubu <= disp) {
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    t_idle = now;
    const uint64_t n_new = min(pubu - disp, (uint64_t)kWave);
    const uint64_t r = disp + lane;
    const uint32_t slot = (uint32_t)(r & (slots - 1));
    uint32_t nch = 0;
    int k = 0;
    if (lane < n_new) {
      const ServeReq* q = &h->ring[slot];
      k = q->k;
      nch = (k >= 1 && k <= kServeMaxK) ? (uint32_t)code_chunks(k, d->L[k]) : 0u;
      uint32_t* rows = d->rows + (size_t)slot * kServeRowWords;
      for (int i = 0; i < k && i < kServeMaxK; ++i) {
        int g = q->idx[i];
        if (g < 0 || g >= kNpts) g = 0;
        uint32_t w[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) w[t] = mplanes[g * kRowWords + t];
        for (int b = 0; b < 3; ++b) {
          uint32_t* o = rows + (3 * i + b) * kRowWords;
#pragma unroll
          for (int t = 0; t < 6; ++t) o[t] = w[t];
          mul_x(w);
        }
      }
      d->s_k[slot] = (nch != 0u) ? k : 0;
      d->s_tmax[slot] = q->tmax;
      d->s_best[slot] = 64u;
      d->s_done[slot] = 0u;
      d->s_nch[slot] = nch;
      d->s_seq[slot] = r + 1;
    }
    // Exclusive prefix of chunk counts across the new requests.
    uint32_t incl = nch;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t o = (uint32_t)__shfl_up((int)incl, off, kWave);
      if (lane >= (uint32_t)off) incl += o;
    }
    const uint32_t sum = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
    if (lane < n_new) {
      d->ord_start[r & ord_mask] = total + (incl - nch);
      if (nch == 0u) {   // k = 0 (result 0) or unsupported k (result -1): done now
        h->results[slot].result = (k == 0) ? 0 : -1;
        __hip_atomic_store(&h->results[slot].seq, r + 1, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    __threadfence();
    disp += n_new;
    total += sum;
    if (lane == 0) {
      __hip_atomic_store(&d->dispatched, disp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&d->avail, total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&h->consumed, disp, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
         __builtin_amdgcn_readfirstlane((uint32_t)v);
}

// Claims as in md_main_kernel: one lane-0 region per iteration (it completes
// the previous chunk and claims the next id), read back with readfirstlane
// while every lane is active; all other conditions are made wave-uniform.
__device__ void worker(ServeDev* d, ServeHost* h, uint32_t lane) {
  const uint64_t slots = d->slots;
  const uint64_t ord_mask = 4ull * slots - 1ull;
  uint64_t first = 0;
  if (lane == 0) first = atomicAdd((unsigned long long*)&d->next_chunk, 1ull);
  uint64_t g = uniform64(first);
  for (;;) {
    // Wait until chunk g is published, or the server stops.
    bool stop = false;
    for (;;) {
      const uint64_t av = uniform64(
          __hip_atomic_load(&d->avail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
      if (g < av) break;
      if (__builtin_amdgcn_readfirstlane(
              __hip_atomic_load(&d->shutdown, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT))) {
        stop = true;
        break;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    if (stop) return;
    // Request of chunk g: last r in the in-flight window with ord_start <= g.
    const uint64_t disp = uniform64(
        __hip_atomic_load(&d->dispatched, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT));
    uint64_t lo = (disp > slots) ? disp - slots : 0, hi = disp - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (uniform64(ld_rlx(&d->ord_start[mid & ord_mask])) <= g) lo = mid; else hi = mid - 1;
    }
    const uint32_t slot = (uint32_t)(lo & (slots - 1));
    const int k = (int)__builtin_amdgcn_readfirstlane((uint32_t)ld_rlx(&d->s_k[slot]));
    const int tmax = (int)__builtin_amdgcn_readfirstlane((uint32_t)ld_rlx(&d->s_tmax[slot]));
    const uint32_t nch = __builtin_amdgcn_readfirstlane(ld_rlx(&d->s_nch[slot]));
    const uint64_t seq = uniform64(ld_rlx(&d->s_seq[slot]));
    const uint64_t local = g - uniform64(ld_rlx(&d->ord_start[lo & ord_mask]));
    const int thr = 48 - tmax;
    uint32_t* best_g = &d->s_best[slot];
    const bool skip = (int)__builtin_amdgcn_readfirstlane(ld_rlx(best_g)) <= thr;
    uint32_t b = 64u;
    if (!skip) {
      const int L = d->L[k];
      const uint32_t cmask = (1u << d->check_log2[k]) - 1u;
      const CoherentRows rows{d->rows + (size_t)slot * kServeRowWords};
      b = run_chunk<true>(rows, k, L, local, lane, thr, cmask, best_g);
      b = wave_min(b);
    }
    uint64_t next = 0;
    if (lane == 0) {
      if (!skip) atomicMin(best_g, b);
      __threadfence();
      const uint32_t done = atomicAdd(&d->s_done[slot], 1u) + 1u;
      if (done == nch) {
        __threadfence();
        const int bw = (int)ld_rlx(best_g);
        h->results[slot].result = (bw <= thr) ? tmax + 1 : 49 - bw;
        __hip_atomic_store(&h->results[slot].seq, seq, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
      next = atomicAdd((unsigned long long*)&d->next_chunk, 1ull);
    }
    g = uniform64(next);
  }
}

__global__ __launch_bounds__(256) void serve_kernel(ServeDev* d, ServeHost* h,
                                                    const uint32_t* __restrict__ mplanes) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  if (blockIdx.x == 0 && wave == 0)
    dispatcher(d, h, mplanes, lane);
  else
    worker(d, h, lane);
}

__global__ void serve_init_kernel(ServeDev* d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    d->next_chunk = 0;
    d->avail = 0;
    d->dispatched = 0;
    d->shutdown = 0;
  }
}

int serve_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, serve_kernel, 256, 0) != hipSuccess ||
      nb <= 0)
    nb = 2;
  return nb;
}

hipError_t launch_serve(int grid, ServeDev* d, ServeHost* h, const uint32_t* mplanes,
                        hipStream_t s) {
  hipLaunchKernelGGL(serve_init_kernel, dim3(1), dim3(64), 0, s, d);
  hipLaunchKernelGGL(serve_kernel, dim3(grid), dim3(256), 0, s, d, h, mplanes);
  return hipGetLastError();
}

}  // namespace gt
