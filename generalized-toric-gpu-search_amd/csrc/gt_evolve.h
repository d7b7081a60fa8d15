// One generation of the evolutionary search on the device: tournament and
// point-swap mutation, choice of the sets to evaluate, survivor selection.
// The canonical-form, hash-set and min-distance steps between them are the
// kernels of canon_kernels.hip, bfs_kernels.hip and md_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"

namespace gt {

// Largest population (survivor sort of 2 * pop keys in 64 KB of LDS) and
// proposal count (score sort of n_prop keys).
constexpr int kEvoMaxPop = 4096;
constexpr int kEvoMaxProp = 8192;

// state[0] = completed generations (the generation index g of the next one).
// Random draws are a stateless function of (seed, g, i, j): counter-based,
// so a generation reads no generator state and replays of one graph advance
// by state[0] alone.
GT_HD uint64_t evo_rand(uint64_t seed, uint32_t g, uint32_t i, uint32_t j) {
  uint64_t z = seed ^ ((uint64_t)g << 40) ^ ((uint64_t)i << 8) ^ (uint64_t)j;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Integer in [0, m) from the high 32 bits of r (multiply-shift).
GT_HD uint32_t evo_below(uint64_t r, uint32_t m) {
  return (uint32_t)(((r >> 32) * (uint64_t)m) >> 32);
}

// Proposals: n_prop children, each a 1- or 2-point swap of a size-4
// tournament winner (highest d * 1000 - age, first of equals).
hipError_t launch_evo_propose(const uint64_t* pop_m, const int32_t* pop_d, const int32_t* pop_age,
                              int32_t pop, int32_t n_prop, uint64_t seed, const uint32_t* state,
                              uint64_t* props, hipStream_t s);
// Chosen sets: props[keep[j]] for the first min(*count, pop) new proposals in
// proposal order or, with scores and more than pop new proposals, the pop
// best-scored ones (ties: proposal order), written in proposal order; random
// k-sets fill the rest.
hipError_t launch_evo_select(const uint64_t* props, const uint32_t* keep, const uint32_t* count,
                             const float* score, int32_t pop, int32_t n_prop, int32_t k,
                             uint64_t seed, const uint32_t* state, uint64_t* chosen,
                             hipStream_t s);
// Proposals that are the first of their class in this generation and not in
// the evaluated set (device hash set `seen`, read only), in proposal order,
// into keep[0 .. *count) (one workgroup; LDS table of 2 n slots).
hipError_t launch_evo_dedup(const uint64_t* pcm, int32_t n, const uint64_t* seen, uint64_t seen_cap,
                            uint32_t* keep, uint32_t* count, hipStream_t s);
// Survivors: the best pop of parents (age + 1) and children (age 0) by
// distance, then age, then position (parents first); logs the children's
// classes and distances in slot g + 1 and advances state[0].  init: no
// parents; the children (the initial population) are sorted into the
// population and logged in slot 0.
hipError_t launch_evo_survive(uint64_t* pop_m, int32_t* pop_d, int32_t* pop_age,
                              const uint64_t* chosen, const uint64_t* chosen_cls,
                              const int32_t* chosen_mz, const uint32_t* count, int32_t pop,
                              bool init, int32_t max_gen, uint32_t* state, uint64_t* log_cls,
                              int32_t* log_d, int32_t* log_best, int32_t* log_new,
                              hipStream_t s);

}  // namespace gt
