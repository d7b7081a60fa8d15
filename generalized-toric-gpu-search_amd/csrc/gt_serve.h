// Shared host/device layout of the persistent serving kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"
#include "gt_launch.h"

namespace gt {

constexpr int kServeMaxK = 16;
constexpr int kServeRowWords = 3 * kServeMaxK * kRowWords;   // bit rows of one slot

// One request descriptor (fine-grained host memory, written by the host).
struct ServeReq {
  int32_t k;
  int32_t tmax;
  int32_t idx[kServeMaxK];
  uint64_t tag;
};

// One result (fine-grained host memory, written by the device): seq = request
// index + 1 once `result` is valid.
struct ServeRes {
  int32_t result;
  uint32_t pad;
  uint64_t seq;
};

// Host-side control block (fine-grained host memory).
struct ServeHost {
  uint64_t published;   // host -> device: requests written to the ring
  uint64_t consumed;    // device -> host: requests taken by the dispatcher
  uint32_t shutdown;    // host -> device
  uint32_t exited;      // device -> host: the kernel has stopped
  ServeReq* ring;       // `slots` descriptors, request r in ring[r % slots]
  ServeRes* results;    // `slots` results, request r in results[r % slots]
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
  uint8_t L[kServeMaxK + 1];
  uint8_t check_log2[kServeMaxK + 1];
  int32_t* s_k;
  int32_t* s_tmax;
  uint32_t* s_best;
  uint32_t* s_done;
  uint32_t* s_nch;
  uint64_t* s_seq;
  uint64_t* ord_start;  // 4 * slots entries, request r at r % (4 * slots)
  uint32_t* rows;       // slots * kServeRowWords
};

int serve_blocks_per_cu();
hipError_t launch_serve(int grid, ServeDev* d, ServeHost* h, const uint32_t* mplanes,
                        hipStream_t s);

}  // namespace gt
