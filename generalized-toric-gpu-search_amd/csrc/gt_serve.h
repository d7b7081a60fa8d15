// Shared host/device layout of the persistent serving kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"
#include "gt_launch.h"

namespace gt {

// The exact limit of the batch path (kernel.py:246-247); 21 six-bit indices
// fit one uint4 in a slot.
constexpr int kServeMaxK = kMaxKExact;
// Bit rows of one request's walk in a worker's LDS: a request of k <= 8
// points is searched whole by one wave (at most kFusedRows = 45 rows); a
// larger one is walked chunk by chunk in up to three phases (serve_phases),
// at most the 3 x 3 (k - 3) rows of level 3's three subcodes.
constexpr int kServeRowWords = 3 * 3 * (kServeMaxK - 3) * kRowWords;
static_assert(3 * 3 * (kServeMaxK - 3) >= 45, "the rows of a whole small request fit");

// Phases of a request of k > kFusedMaxK points: its first level m[0], the
// level below it m[1] (0: none, mid_level), and level 2 (second_level),
// with their chunks at walk length L (c[1] = 0 without a middle level, and
// only the first phase when the first level is already level 2).  A later
// phase's chunks are walked only when the earlier phases did not certify
// the request (serve_kernel.hip).
// A subcode of walk dimension at most kServeWholeW (at most 299,593
// codewords, about 90 us for one wave) is one chunk, walked whole; larger
// ones are chunked as in the batch walk (a chunk also pays a claim, a
// lookup, the rows and a fold, and a later phase's chunk waits for the
// earlier phases: fewer chunks, fewer waiting workers).  On the 20,000
// request k = 3..10 stream: 6 -> 7 took the burst capacity from 2.95 to
// 3.3-3.4 M requests/s and p99 at 0.7x from 0.065 to 0.057-0.059 ms; 8 made
// a k = 12 request's first phase three 2.4 M-codeword single-wave chunks.
constexpr int kServeWholeW = 7;
__host__ __device__ __forceinline__ uint32_t serve_sub_chunks(int w, int L) {
  return (w <= kServeWholeW) ? 1u : (uint32_t)code_chunks(w, L);
}
__host__ __device__ __forceinline__ void serve_phases(const uint8_t* level, int k, int L,
                                                      int m[3], uint32_t c[3]) {
  m[0] = (int)level[k];
  m[1] = mid_level(level, k);
  m[2] = second_level(k);
  c[0] = (uint32_t)level_nsub(m[0]) * serve_sub_chunks(k - m[0], L);
  c[1] = (m[1] > 0) ? (uint32_t)level_nsub(m[1]) * serve_sub_chunks(k - m[1], L) : 0u;
  c[2] = (m[0] > m[2]) ? serve_sub_chunks(k - m[2], L) : 0u;
}
// Bits of a request's phase flags.
enum : uint32_t {
  kPhNoRank0 = 1u,    // the first level had no direction of full rank
  kPhNoRank1 = 2u,    // the middle level had none
  kPhMidOff = 4u,     // the middle phase was not certified beforehand: skipped
  kPhTimeout = 8u,    // a phase wait ran out: the last phase is walked
};

// One request descriptor (fine-grained host memory, written by the host):
// 48 bytes, read by the dispatcher with three 16-byte loads.
struct ServeReq {
  int32_t k;            // bits 0..7 k, bits 8..15 walk length L for this request (0: default)
  int32_t tmax;
  uint64_t tag;
  uint8_t idx[32];      // exponent indices 0..48, entries k.. unused
};

// Slot copy of a request's indices: index i in bits 6i .. 6i+5 of 128 bits.
__host__ __device__ __forceinline__ uint32_t unpack_idx6(const uint4& v, uint32_t i) {
  const uint64_t lo = ((uint64_t)v.y << 32) | v.x, hi = ((uint64_t)v.w << 32) | v.z;
  const uint32_t bit = 6u * i;
  uint64_t x;
  if (bit + 6u <= 64u) x = lo >> bit;
  else if (bit >= 64u) x = hi >> (bit - 64u);
  else x = (lo >> bit) | (hi << (64u - bit));
  return (uint32_t)(x & 63u);
}

// One result (fine-grained host memory, written by the device with one 64-bit
// store): bits 0..31 the result, bits 32..63 (request index + 1) mod 2^32, so
// the host sees a result and the request it belongs to at once.
struct ServeRes {
  uint64_t word;
};
__host__ __device__ __forceinline__ uint64_t serve_res_word(uint64_t r, int32_t result) {
  return ((r + 1ull) << 32) | (uint32_t)result;
}

// Host-side control block (fine-grained host memory).
struct ServeHost {
  uint64_t published;   // host -> device: requests written to the ring
  uint64_t consumed;    // device -> host: requests taken by the dispatcher
  uint32_t shutdown;    // host -> device
  uint32_t exited;      // device -> host: the kernel has stopped
  ServeReq* ring;       // `slots` descriptors, request r in ring[r % slots]
  ServeRes* results;    // `slots` results, request r in results[r % slots]
};

// Counters the kernel accumulates (per wave, added once when the wave exits);
// times are device wall-clock ticks.
enum ServeStat : int {
  kStDispPasses = 0,   // dispatcher passes that took at least one request
  kStDispRequests,     // requests dispatched
  kStDispTicks,        // dispatcher time inside those passes
  kStChunks,           // chunks walked
  kStSkipped,          // chunks skipped (their request had stopped, or was certified)
  kStWaitTicks,        // worker time waiting for a chunk to be published
  kStLookupTicks,      // worker time finding the request, reading its state, staging rows
  kStWalkTicks,        // worker time in the Gray walk
  kStFoldTicks,        // worker time folding, counting and claiming
  kStStaged,           // row stagings into LDS
  kStLifeTicks,        // dispatcher lifetime
  kStWorkers,          // worker waves
  kServeStats
};

// Device-side state (device memory).
struct ServeDev {
  uint64_t next_chunk;  // worker claim counter
  uint64_t avail;       // chunks published by the dispatcher
  uint64_t dispatched;  // requests dispatched
  uint32_t shutdown;
  uint32_t slots;       // power of two; in-flight requests <= slots
  uint64_t idle_ticks;  // dispatcher idle limit (device wall clock ticks)
  uint64_t life_ticks;  // kernel lifetime limit
  uint64_t first_req;   // first request index this launch dispatches (set by the init kernel)
  uint8_t L[kServeMaxK + 1];
  uint8_t check_log2[kServeMaxK + 1];
  uint8_t level[kServeMaxK + 1];   // first level of a request (its search or first phase)
  int32_t* s_k;         // per slot: k | (L << 8), 0 when the request has no chunks
  int32_t* s_tmax;
  uint64_t* s_state;    // per slot: bits 0..31 minimum weight, 32..63 chunks finished
  uint32_t* s_nch;
  uint64_t* s_seq;
  uint64_t* ord_start;  // 4 * slots tagged entries (ord_entry), request r at r % (4 * slots)
  uint4* s_idx;         // per slot: the request's exponent indices, 6 bits each (unpack_idx6)
  uint64_t* s_ph;       // per slot: chunks of the first phase | chunks of the middle phase << 32
  uint32_t* s_flags;    // per slot: phase flags (kPh*)
  uint64_t stats[kServeStats];
};

// ord_start entry of request r: its first global chunk id in bits 0..39 and
// the tag (r mod 2^23) | 2^23 in bits 40..63.  The tag is never 0 (cleared
// entries never match) and differs between requests that share a ring entry
// (4 * slots <= 2^22), so a reader that sees an entry of another request
// (older, or not yet visible to it) can tell.
constexpr uint64_t kOrdStartBits = 40;
__host__ __device__ __forceinline__ uint64_t ord_tag(uint64_t r) {
  return (r & 0x7FFFFFull) | 0x800000ull;
}
__host__ __device__ __forceinline__ uint64_t ord_entry(uint64_t r, uint64_t start) {
  return (ord_tag(r) << kOrdStartBits) | start;
}

int serve_blocks_per_cu();
// first_req: the request index the new kernel starts dispatching at (the
// host keeps its numbering when a kernel that stopped itself left published
// requests undispatched).
hipError_t launch_serve(int grid, ServeDev* d, ServeHost* h, const uint32_t* mplanes,
                        uint64_t first_req, hipStream_t s);

}  // namespace gt
