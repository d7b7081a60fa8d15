// Layout of the surrogate MLP (shared by its kernels and the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "gt_common.h"

namespace gt {

constexpr int kMlpIn = 128;    // 49 mask bits, |S|/16, 48 difference bins, padding
constexpr int kMlpHid = 128;
constexpr int kMlpTile = 32;   // rows per workgroup
// Flat fp32 parameter vector: W1[hid][in], b1, W2[hid][hid], b2, w3[hid], b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid;
constexpr int kMlpParams = kOffB3 + 1;
constexpr int kMlpPartialStride = (kMlpParams + 1 + 3) & ~3;   // params + loss slot
// A training step's gradient comes from at most kMlpMaxParts workgroups, each
// summing a run of consecutive tiles (so the reduction reads at most this many
// partials).
constexpr int kMlpMaxParts = 64;

struct MlpBufs {
  float* theta = nullptr;   // kMlpParams
  float* grads = nullptr;   // kMlpParams + 1 (last = loss)
  float* m1 = nullptr;
  float* m2 = nullptr;
  uint16_t* w1b = nullptr;
  uint16_t* w2b = nullptr;
  uint16_t* w3b = nullptr;
};

// With d_gate, a launch does nothing while *d_gate <= gate_min (read on the
// device: the evolutionary search scores its proposals only when more new
// ones than the population survive deduplication).
hipError_t launch_mlp_features(const uint64_t* d_masks, int32_t n, uint16_t* d_feat, hipStream_t s,
                               const uint32_t* d_gate = nullptr, uint32_t gate_min = 0);
hipError_t launch_mlp_predict(const MlpBufs& b, const uint16_t* d_feat, int32_t n, float* d_out,
                              hipStream_t s, const uint32_t* d_gate = nullptr,
                              uint32_t gate_min = 0);
// Partials (gradient workgroups) of a step on n rows.
int mlp_grad_parts(int32_t n);
hipError_t launch_mlp_grads(const MlpBufs& b, const uint16_t* d_feat, const float* d_target,
                            int32_t n, float* d_partial, hipStream_t s);
// Gradient of one batch and the Adam step (lr, step >= 1) in two kernels;
// the batch loss (before the step) goes to grads[kMlpParams] and, if
// d_loss != NULL, to *d_loss.
hipError_t launch_mlp_train_step(const MlpBufs& b, const uint16_t* d_feat, const float* d_target,
                                 int32_t n, float* d_partial, float lr, int step, float* d_loss,
                                 hipStream_t s);
hipError_t launch_mlp_cast(const MlpBufs& b, hipStream_t s);
// Targets (d - mean) / std of n integer distances (one workgroup).
hipError_t launch_mlp_standardize(const int32_t* d_dist, int32_t n, float* d_target,
                                  hipStream_t s);
hipError_t launch_bf16_gemm(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
                            hipStream_t s);

// Accessors for other translation units (the evolutionary search reads the
// ranker's parameters inside its captured generation).
}  // namespace gt
struct gt_mlp;
namespace gt {
const MlpBufs* mlp_buffers(gt_mlp* m);
hipStream_t mlp_stream(gt_mlp* m);
int mlp_device(const gt_mlp* m);
// Train on n device-resident (mask, distance) pairs, enqueued on stream s
// with no host synchronisation: features once, standardised targets, then
// `epochs` passes of Adam steps over consecutive batches in order.
int mlp_train_device(gt_mlp* m, const uint64_t* d_masks, const int32_t* d_dist, int32_t n,
                     int32_t batch, float lr, int32_t epochs, hipStream_t s);

}  // namespace gt
