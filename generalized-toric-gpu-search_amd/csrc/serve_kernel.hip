// Continuously batched min-distance serving: one persistent kernel.
//
// Requests (k <= kServeMaxK exponent points and a threshold tmax) arrive in a
// ring in fine-grained host memory.  Wave 0 of workgroup 0 is the dispatcher:
// it polls the ring's publish counter, copies each new request's exponent
// indices (6 bits each, one 16-byte store) and state into the request's slot
// (slot = request index mod slots), appends the request's chunk range to the
// global chunk sequence and publishes the new chunk total.  Every other wave
// is a worker: it claims the next global chunk id with one atomic, waits
// until that id is published, finds the request by a 64-ary search over the
// in-flight window and folds the wave minimum of its chunk into the slot.
// A request of k <= 8 points is one chunk: the wave runs the request's
// whole search (the anchored subcodes of its first level, rows in the
// wave's LDS, lower levels when the first does not certify: fused_search,
// gt_anchor.h).  A larger request's chunks cover up to three phases
// (serve_phases, gt_serve.h): its first level (4 by default), the level
// below it, and level 2 (certified for k >= 9 by max_zeros >= k - 1;
// gt_common.h).  A wave builds the rows of its chunk's level in its LDS area
// and walks the chunk (the Gray walk of the batch kernel).  A chunk of a
// later phase first waits until every chunk of the earlier phases is
// counted (those chunks have lower ids, so running workers already hold
// them: the wait ends; it is also bounded), then is skipped when the
// request's minimum so far certifies it: a first-level maximum of 7m - 7 or
// more zeros (gt_common.h; codewords of later phases in the minimum only
// raise it, and z* <= max(first-level maximum, 7(m - 1)) still holds), or,
// before the last phase, a middle level that ran whole with full rank and
// reached its own 7m' - 7.  The middle phase runs only when max(z, k - 1)
// certifies its level beforehand or z meets its found certificate; the last
// phase is exact.  Building the rows in the workers keeps the dispatcher,
// one wave for the whole device, to a few stores per request.  The wave
// that finishes the last chunk of a request writes its result and sequence
// number to host memory.
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
// Every chunk published before the stop is still walked (a worker holding a
// published id walks it before it looks at the flag), so every dispatched
// request gets its result.  Requests the host published but the dispatcher
// had not taken when it stopped (`consumed` < `published`) are dispatched by
// the next launch, which the host starts at request `consumed`.
#include "gt_anchor.h"
#include "gt_serve.h"

namespace gt {

template <class T>
__device__ __forceinline__ T ld_rlx(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void dispatcher(ServeDev* d, ServeHost* h, uint32_t lane) {
  const uint64_t t_start = wall_clock64();
  uint64_t t_idle = t_start;
  uint64_t disp = d->first_req;   // requests dispatched so far (numbered across launches)
  const uint64_t disp0 = disp;
  uint64_t total = 0;         // chunks published so far
  uint64_t passes = 0, busy = 0;
  const uint32_t slots = d->slots;
  const uint64_t ord_mask = 4ull * slots - 1ull;
  for (;;) {
    const uint64_t now = wall_clock64();
    // Relaxed polls (an acquire per poll would invalidate this XCD's caches
    // every time); one system-scope acquire fence once new requests are seen.
    const bool stop = __builtin_amdgcn_readfirstlane(
                          __hip_atomic_load(&h->shutdown, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_SYSTEM)) != 0u ||
                      now - t_idle > d->idle_ticks || now - t_start > d->life_ticks;
    if (stop) {
      if (lane == 0) {
        atomicAdd((unsigned long long*)&d->stats[kStDispPasses], passes);
        atomicAdd((unsigned long long*)&d->stats[kStDispRequests], disp - disp0);
        atomicAdd((unsigned long long*)&d->stats[kStDispTicks], busy);
        atomicAdd((unsigned long long*)&d->stats[kStLifeTicks], now - t_start);
        __hip_atomic_store(&d->shutdown, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&h->exited, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      return;
    }
    const uint64_t pub =
        __hip_atomic_load(&h->published, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const uint64_t pubu = uniform64(pub);
    if (pubu <= disp) {
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    t_idle = now;
    const uint64_t n_new = min(pubu - disp, (uint64_t)kWave);
    const uint64_t r = disp + lane;
    const uint32_t slot = (uint32_t)(r & (slots - 1));
    uint32_t nch = 0;
    int k = 0;
    int32_t imm = 0;   // immediate result of a request without chunks
    if (lane < n_new) {
      // The descriptor (48 bytes) in three 16-byte loads: one round trip to
      // host memory.
      static_assert(sizeof(ServeReq) == 3 * sizeof(uint4), "ServeReq layout");
      const uint4* qv = reinterpret_cast<const uint4*>(&h->ring[slot]);
      const uint4 v0 = qv[0], v1 = qv[1], v2 = qv[2];
      // Bits 0..7: k; bits 8..15: the request's walk length L (0: the
      // server's per-k default).
      k = (int)(v0.x & 0xFFu);
      const int Lq = (int)((v0.x >> 8) & 0xFFu);
      const int32_t q_tmax = (int32_t)v0.y;
      // k <= 8: one chunk, the whole search in one wave; larger k: the
      // chunks of its phases.
      int L = 0;
      uint32_t pc[3] = {0u, 0u, 0u};
      if (k > kFusedMaxK && k <= kServeMaxK) {
        L = (Lq != 0) ? Lq : (int)d->L[k];
        int pm[3];
        serve_phases(d->level, k, L, pm, pc);
      }
      nch = (k < 2 || k > kServeMaxK) ? 0u
            : (k <= kFusedMaxK)       ? 1u
                                      : pc[0] + pc[1] + pc[2];
      const uint32_t ib[8] = {v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
      uint64_t lo = 0, hi = 0;
#pragma unroll
      for (int i = 0; i < kServeMaxK; ++i) {
        uint32_t g = (ib[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        if (i >= k || g >= (uint32_t)kNpts) g = 0;   // the host checked the indices
        const int bit = 6 * i;
        if (bit + 6 <= 64) {
          lo |= (uint64_t)g << bit;
        } else if (bit >= 64) {
          hi |= (uint64_t)g << (bit - 64);
        } else {
          lo |= (uint64_t)g << bit;
          hi |= (uint64_t)g >> (64 - bit);
        }
      }
      d->s_idx[slot] = make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi,
                                  (uint32_t)(hi >> 32));
      d->s_k[slot] = (nch != 0u) ? (k | (L << 8)) : 0;
      d->s_tmax[slot] = q_tmax;
      d->s_state[slot] = 64ull;
      d->s_nch[slot] = nch;
      d->s_ph[slot] = (uint64_t)pc[0] | ((uint64_t)pc[1] << 32);
      d->s_flags[slot] = 0u;
      d->s_seq[slot] = r + 1;
      // k <= 1 needs no walk: k = 0 gives 0, k = 1 (a monomial, weight 49)
      // gives 0 or, below tmax = -1, the clamp tmax + 1; unsupported k: -1.
      if (nch == 0u) imm = (k == 0) ? 0 : (k == 1) ? ((49 <= 48 - q_tmax) ? q_tmax + 1 : 0) : -1;
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
      d->ord_start[r & ord_mask] = ord_entry(r, total + (incl - nch));
      if (nch == 0u)   // k <= 1 or unsupported k: done now
        __hip_atomic_store(&h->results[slot].word, serve_res_word(r, imm),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    disp += n_new;
    total += sum;
    // One release per pass: it orders every row, state and ord_start store of
    // this pass (and the relaxed `dispatched`) before the new chunk total.
    if (lane == 0) {
      __hip_atomic_store(&d->dispatched, disp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&d->avail, total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&h->consumed, disp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    ++passes;
    busy += wall_clock64() - now;
  }
}

// Rows of the request a wave is walking, staged in that wave's LDS area.
// Claims as in md_main_thr_kernel: one lane-0 region per iteration (it completes
// the previous chunk and claims the next id), read back with readfirstlane
// while every lane is active; all other conditions are made wave-uniform.
//
// Memory ordering.  A worker acquires `avail` once it holds a chunk id that
// its last observed value does not cover; that acquire makes the rows, slot
// state and ord_start entries of every request published up to that value
// visible to plain loads.  The request of chunk g is the last r in the
// in-flight window whose ord_start entry carries r's tag and starts at or
// before g; entries of requests this wave cannot see yet carry other tags
// and start after g anyway.  The request's rows are copied into LDS once per
// visit, and the Gray walk reads them there.
__device__ void worker(ServeDev* d, ServeHost* h, const uint32_t* __restrict__ mplanes,
                       uint32_t lane, uint32_t* lrows) {
  const uint64_t slots = d->slots;
  const uint64_t ord_mask = 4ull * slots - 1ull;
  const uint64_t start_mask = (1ull << kOrdStartBits) - 1ull;
  uint64_t av_known = 0;
  uint32_t staged_slot = 0xFFFFFFFFu;
  uint64_t staged_seq = 0;
  int staged_m = 0;
  uint64_t st_chunks = 0, st_skipped = 0, st_staged = 0;
  uint64_t st_wait = 0, st_lookup = 0, st_walk = 0, st_fold = 0;
  uint64_t first = 0;
  if (lane == 0) first = atomicAdd((unsigned long long*)&d->next_chunk, 1ull);
  uint64_t g = uniform64(first);
  for (;;) {
    // Wait until chunk g is published, or the server stops.
    uint64_t t0 = wall_clock64();
    bool stop = false;
    if (g >= av_known) {
      // Spin on relaxed loads; one acquire fence once the chunk is published
      // (an acquire per poll would invalidate this XCD's caches every time).
      for (;;) {
        const uint64_t av = uniform64(ld_rlx(&d->avail));
        if (g < av) { av_known = av; break; }
        if (__builtin_amdgcn_readfirstlane(ld_rlx(&d->shutdown))) { stop = true; break; }
        __builtin_amdgcn_s_sleep(8);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    uint64_t t1 = wall_clock64();
    st_wait += t1 - t0;
    if (stop) break;
    // Request of chunk g: the last r in the window whose entry carries r's tag
    // and starts at or before g (true for the oldest in-window request and
    // monotone over the window).  64-ary search: each level probes 64
    // entries with one load, one per lane, and keeps the last true one; two
    // dependent loads for a window of 4096 requests instead of twelve.
    const uint64_t disp = uniform64(ld_rlx(&d->dispatched));
    const uint64_t first_req = d->first_req;   // requests before it have no entries here
    uint64_t lo = (disp > first_req + slots) ? disp - slots : first_req, hi = disp - 1;
    uint64_t step = 1;
    while (step * kWave < hi - lo + 1) step *= kWave;
    uint64_t e_lo = 0;
    for (;;) {
      const uint64_t r = lo + (uint64_t)lane * step;
      uint64_t e = 0;
      bool ok = false;
      if (r <= hi) {
        e = d->ord_start[r & ord_mask];
        ok = (e >> kOrdStartBits) == ord_tag(r) && (e & start_mask) <= g;
      }
      const uint64_t bal = __ballot(ok);
      const int last = (bal != 0) ? 63 - __builtin_clzll(bal) : 0;
      lo += (uint64_t)last * step;
      e_lo = readlane64(e, last);
      if (step == 1) break;
      hi = (lo + step - 1 < hi) ? lo + step - 1 : hi;
      step /= kWave;
    }
    const uint32_t slot = (uint32_t)(lo & (slots - 1));
    const int kl = (int)__builtin_amdgcn_readfirstlane((uint32_t)d->s_k[slot]);
    const int k = kl & 0xFF, L = kl >> 8;
    const int tmax = (int)__builtin_amdgcn_readfirstlane((uint32_t)d->s_tmax[slot]);
    const uint32_t nch = __builtin_amdgcn_readfirstlane(d->s_nch[slot]);
    const uint64_t seq = uniform64(d->s_seq[slot]);
    const uint64_t local = g - (e_lo & start_mask);
    const int thr = 48 - tmax;
    uint64_t* state = &d->s_state[slot];
    // Low half of the slot state: the running minimum weight.
    const uint32_t* best_g = reinterpret_cast<const uint32_t*>(state);
    const bool skip = (int)__builtin_amdgcn_readfirstlane(ld_rlx(best_g)) <= thr;
    uint32_t b = 64u;
    uint64_t t2 = t1;
    bool walked_chunk = !skip;   // (a later phase's chunk may be skipped below)
    if (!skip) {
      // Lane i < k: the request's point i.
      const uint4 iv = d->s_idx[slot];
      const int gi = (int)unpack_idx6(iv, (lane < (uint32_t)k) ? lane : 0u);
      if (k <= kFusedMaxK) {
        // The whole search of a small request (rows built in the LDS area).
        t2 = wall_clock64();
        uint64_t walked = 0;
        b = fused_search(gi, k, tmax, lane, mplanes, d->level, lrows, walked);
        staged_slot = 0xFFFFFFFFu;
        ++st_staged;
      } else {
        // The chunk's phase and level.
        const uint64_t ph = uniform64(d->s_ph[slot]);
        const uint32_t c0 = (uint32_t)ph, c1 = (uint32_t)(ph >> 32);
        const int phase = (local < c0) ? 0 : (local < (uint64_t)c0 + c1) ? 1 : 2;
        const uint64_t lp = local - ((phase >= 1) ? c0 : 0u) - ((phase == 2) ? c1 : 0u);
        const int m0 = (int)d->level[k], m1 = mid_level(d->level, k);
        const int m = (phase == 0) ? m0 : (phase == 1) ? m1 : second_level(k);
        uint32_t* flags = &d->s_flags[slot];
        bool walk = true;
        if (phase > 0) {
          // Wait until the earlier phases' chunks are all counted.
          const uint32_t need = (phase == 1) ? c0 : c0 + c1;
          const uint32_t* cnt = reinterpret_cast<const uint32_t*>(state) + 1;
          bool ready = false;
          for (uint32_t it = 0; it < (1u << 22); ++it) {
            if (__builtin_amdgcn_readfirstlane(ld_rlx(cnt)) >= need) { ready = true; break; }
            __builtin_amdgcn_s_sleep(2);
          }
          uint32_t f = __builtin_amdgcn_readfirstlane(ld_rlx(flags));
          if (!ready) {
            if (lane == 0) atomicOr(flags, kPhTimeout);
            f |= kPhTimeout;
          }
          const int z = kNpts - (int)__builtin_amdgcn_readfirstlane(ld_rlx(best_g));
          const bool first_ok = !(f & kPhNoRank0) && z >= found_cert(m0);
          if (phase == 1) {
            if (ready && (z > tmax || first_ok)) {
              walk = false;   // certified
            } else if (!(level_cert(m1) <= k - 1 || (!(f & kPhNoRank0) && found_cert(m1) <= z))) {
              if (lane == 0) atomicOr(flags, kPhMidOff);   // the middle level would not certify
              walk = false;
            }
          } else {
            const bool mid_ok = c1 > 0u && !(f & (kPhNoRank1 | kPhMidOff)) && z >= found_cert(m1);
            if (!(f & kPhTimeout) && (z > tmax || first_ok || mid_ok)) walk = false;
          }
          walked_chunk = walk;
        }
        if (walk) {
          const int w = k - m;
          if (slot != staged_slot || seq != staged_seq || m != staged_m) {
            // Rows of the phase's level.  Without a direction of full rank:
            // at level 2 every index is the same point, the request spans
            // the zero codeword, and zero rows give its 49 zeros; at a
            // higher level the phase is flagged and its chunks walk nothing.
            wave_lds_sync();
            const bool ok = md_rows_switch(m, gi, k, lane, mplanes, lrows);
            if (!ok && phase == 2)
              for (int r = (int)lane; r < 3 * w; r += kWave)
#pragma unroll
                for (int t = 0; t < 6; ++t) lrows[r * kRowWords + t] = 0u;
            if (!ok && phase < 2 && lane == 0)
              atomicOr(flags, (phase == 0) ? kPhNoRank0 : kPhNoRank1);
            wave_lds_sync();
            staged_slot = slot;
            staged_seq = seq;
            staged_m = (ok || phase == 2) ? m : -1;   // (a failed level is built again)
            ++st_staged;
          }
          if (staged_m == m) {
            t2 = wall_clock64();
            // Subcode s of the phase's level, chunk lp of it (or all of it).
            const uint64_t cc = serve_sub_chunks(w, L);
            const uint32_t sub = (uint32_t)(lp / cc);
            const LdsRows rows{lrows + (size_t)sub * 3 * w * kRowWords};
            if (w <= kServeWholeW) {
              b = walk_all(rows, w, lane);
            } else {
              const uint32_t cmask = (1u << d->check_log2[k]) - 1u;
              b = run_chunk<true>(rows, w, L, lp - (uint64_t)sub * cc, lane, thr, cmask, best_g);
            }
            b = wave_min(b);
          }
        }
      }
      if (walked_chunk) ++st_chunks;
      else ++st_skipped;
    } else {
      t2 = wall_clock64();
      ++st_skipped;
    }
    const uint64_t t3 = wall_clock64();
    st_lookup += t2 - t1;
    st_walk += t3 - t2;
    uint64_t next = 0;
    if (lane == 0) {
      // Fold the minimum with a returning atomic, then count the chunk with
      // an operand computed from the returned value (1 + (prev >> 31) == 1:
      // weights are at most 64), and an s_waitcnt 0 between them: the count
      // cannot issue before the minimum has been performed at the device
      // coherence point (an agent-scope release would instead write back the
      // whole L2).  The wave whose count completes the request therefore reads
      // every wave's minimum there.
      uint32_t* st = reinterpret_cast<uint32_t*>(state);
      uint32_t prev = 0u;
      if (!skip) prev = __hip_atomic_fetch_min(&st[0], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      const uint32_t done = __hip_atomic_fetch_add(&st[1], 1u + (prev >> 31), __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT) + 1u;
      if (done == nch) {
        const int bw = (int)ld_rlx(&st[0]);
        __hip_atomic_store(&h->results[slot].word,
                           serve_res_word(seq - 1, (bw <= thr) ? tmax + 1 : 49 - bw),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      next = atomicAdd((unsigned long long*)&d->next_chunk, 1ull);
    }
    g = uniform64(next);
    st_fold += wall_clock64() - t3;
  }
  if (lane == 0) {
    atomicAdd((unsigned long long*)&d->stats[kStChunks], st_chunks);
    atomicAdd((unsigned long long*)&d->stats[kStSkipped], st_skipped);
    atomicAdd((unsigned long long*)&d->stats[kStWaitTicks], st_wait);
    atomicAdd((unsigned long long*)&d->stats[kStLookupTicks], st_lookup);
    atomicAdd((unsigned long long*)&d->stats[kStWalkTicks], st_walk);
    atomicAdd((unsigned long long*)&d->stats[kStFoldTicks], st_fold);
    atomicAdd((unsigned long long*)&d->stats[kStStaged], st_staged);
    atomicAdd((unsigned long long*)&d->stats[kStWorkers], 1ull);
  }
}

__global__ __launch_bounds__(256) void serve_kernel(ServeDev* d, ServeHost* h,
                                                    const uint32_t* __restrict__ mplanes) {
  __shared__ uint32_t lrows[256 / kWave][kServeRowWords];
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x / kWave;
  if (blockIdx.x == 0 && wave == 0)
    dispatcher(d, h, lane);
  else
    worker(d, h, mplanes, lane, lrows[wave]);
}

// Resets the counters and clears the ord_start ring (a cleared entry carries
// tag 0, which no request has).
__global__ void serve_init_kernel(ServeDev* d, uint64_t first_req) {
  const uint64_t n = 4ull * d->slots;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    d->ord_start[i] = 0ull;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    d->next_chunk = 0;
    d->avail = 0;
    d->dispatched = first_req;
    d->first_req = first_req;
    d->shutdown = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < kServeStats) d->stats[threadIdx.x] = 0ull;
}

int serve_blocks_per_cu() {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, serve_kernel, 256, 0) != hipSuccess ||
      nb <= 0)
    nb = 2;
  return nb;
}

hipError_t launch_serve(int grid, ServeDev* d, ServeHost* h, const uint32_t* mplanes,
                        uint64_t first_req, hipStream_t s) {
  hipLaunchKernelGGL(serve_init_kernel, dim3(64), dim3(256), 0, s, d, first_req);
  hipLaunchKernelGGL(serve_kernel, dim3(grid), dim3(256), 0, s, d, h, mplanes);
  return hipGetLastError();
}

}  // namespace gt
