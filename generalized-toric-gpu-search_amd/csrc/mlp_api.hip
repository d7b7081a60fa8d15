// C ABI of the surrogate MLP (declared in include/gtoric.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/gtoric.h"
#include "gt_launch.h"
#include "gt_mlp.h"

using namespace gt;

namespace {

int mfail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(buf);
}

#define ML_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return mfail("%s: %s", #call, hipGetErrorString(e_));    \
  } while (0)

#define ML_OK(call)         \
  do {                      \
    int r_ = (call);        \
    if (r_ != 0) return r_; \
  } while (0)

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    ML_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    cap = std::max<size_t>(bytes, 256);
    return 0;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace

struct gt_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  MlpBufs b;
  std::vector<void*> fixed;
  Buf masks, feat, target, out, partial, losses;
  int step = 0;

  int features(const uint64_t* d_masks, int32_t n, hipStream_t s) {
    ML_OK(feat.ensure((size_t)n * kMlpIn * 2));
    ML_HIP(launch_mlp_features(d_masks, n, feat.as<uint16_t>(), s));
    return 0;
  }
  int grads(const uint64_t* d_masks, const float* d_target, int32_t n, hipStream_t s) {
    ML_OK(features(d_masks, n, s));
    ML_OK(partial.ensure((size_t)mlp_grad_parts(n) * kMlpPartialStride * 4));
    ML_HIP(launch_mlp_grads(b, feat.as<uint16_t>(), d_target, n, partial.as<float>(), s));
    return 0;
  }
  int train_step(const uint64_t* d_masks, const float* d_target, int32_t n, float lr,
                 float* d_loss, hipStream_t s) {
    ML_OK(features(d_masks, n, s));
    return train_step_feat(feat.as<uint16_t>(), d_target, n, lr, d_loss, s);
  }
  // One Adam step on rows whose features are already in d_feat.
  int train_step_feat(const uint16_t* d_feat, const float* d_target, int32_t n, float lr,
                      float* d_loss, hipStream_t s) {
    ML_OK(partial.ensure((size_t)mlp_grad_parts(n) * kMlpPartialStride * 4));
    ML_HIP(launch_mlp_train_step(b, d_feat, d_target, n, partial.as<float>(), lr, ++step, d_loss,
                                 s));
    return 0;
  }
  int upload(const uint64_t* masks_h, const float* target_h, int32_t n) {
    ML_OK(masks.ensure((size_t)n * 8));
    ML_HIP(hipMemcpyAsync(masks.p, masks_h, (size_t)n * 8, hipMemcpyHostToDevice, stream));
    if (target_h) {
      ML_OK(target.ensure((size_t)n * 4));
      ML_HIP(hipMemcpyAsync(target.p, target_h, (size_t)n * 4, hipMemcpyHostToDevice, stream));
    }
    return 0;
  }
};

namespace gt {
int mlp_train_device(gt_mlp* m, const uint64_t* d_masks, const int32_t* d_dist, int32_t n,
                     int32_t batch, float lr, int32_t epochs, hipStream_t s) {
  if (!m || n <= 0 || batch <= 0 || epochs < 0) return mfail("bad training arguments");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_OK(m->features(d_masks, n, s));
  ML_OK(m->target.ensure((size_t)n * 4));
  ML_HIP(launch_mlp_standardize(d_dist, n, m->target.as<float>(), s));
  for (int e = 0; e < epochs; ++e)
    for (int r0 = 0; r0 < n; r0 += batch) {
      const int rows = std::min(batch, n - r0);
      ML_OK(m->train_step_feat(m->feat.as<uint16_t>() + (size_t)r0 * kMlpIn,
                               m->target.as<float>() + r0, rows, lr, nullptr, s));
    }
  return 0;
}

const MlpBufs* mlp_buffers(gt_mlp* m) { return &m->b; }
hipStream_t mlp_stream(gt_mlp* m) { return m->stream; }
int mlp_device(const gt_mlp* m) { return m->device; }
}  // namespace gt

extern "C" {

int gt_mlp_num_params(int32_t* n) {
  if (!n) return mfail("null argument");
  *n = kMlpParams;
  return 0;
}

int gt_mlp_create(int32_t device, uint64_t seed, gt_mlp** out) {
  if (!out) return mfail("null output");
  *out = nullptr;
  gt_mlp* m = new gt_mlp();
  m->device = device;
  auto bail = [&](int r) {
    for (void* p : m->fixed) (void)hipFree(p);
    for (Buf* x : {&m->masks, &m->feat, &m->target, &m->out, &m->partial, &m->losses})
      if (x->p) (void)hipFree(x->p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return r;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(mfail("hipSetDevice(%d) failed", device));
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(mfail("hipStreamCreate failed"));
  auto alloc = [&](void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return false;
    m->fixed.push_back(*p);
    return true;
  };
  if (!alloc((void**)&m->b.theta, kMlpParams * 4) ||
      !alloc((void**)&m->b.grads, (kMlpParams + 1) * 4) ||
      !alloc((void**)&m->b.m1, kMlpParams * 4) || !alloc((void**)&m->b.m2, kMlpParams * 4) ||
      !alloc((void**)&m->b.w1b, kMlpHid * kMlpIn * 2) ||
      !alloc((void**)&m->b.w2b, kMlpHid * kMlpHid * 2) || !alloc((void**)&m->b.w3b, kMlpHid * 2))
    return bail(mfail("hipMalloc of MLP parameters failed"));
  // He-uniform initialisation from a seeded generator; all biases 0 (targets
  // are standardised by the caller).
  std::vector<float> th(kMlpParams, 0.f);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<float> u(-1.f, 1.f);
  const float a1 = std::sqrt(6.f / 98.f), a2 = std::sqrt(6.f / kMlpHid), a3 = std::sqrt(3.f / kMlpHid);
  for (int o = 0; o < kMlpHid; ++o)
    for (int i = 0; i < kMlpIn; ++i) th[kOffW1 + o * kMlpIn + i] = (i < 98) ? a1 * u(rng) : 0.f;
  for (int p = 0; p < kMlpHid * kMlpHid; ++p) th[kOffW2 + p] = a2 * u(rng);
  for (int p = 0; p < kMlpHid; ++p) th[kOffW3 + p] = a3 * u(rng);
  std::vector<float> zero(kMlpParams, 0.f);
  if (hipMemcpy(m->b.theta, th.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->b.m1, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->b.m2, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      launch_mlp_cast(m->b, m->stream) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess)
    return bail(mfail("MLP initialisation failed"));
  *out = m;
  return 0;
}

int gt_mlp_destroy(gt_mlp* m) {
  if (!m) return 0;
  (void)hipSetDevice(m->device);
  (void)hipStreamSynchronize(m->stream);
  for (void* p : m->fixed) (void)hipFree(p);
  for (Buf* x : {&m->masks, &m->feat, &m->target, &m->out, &m->partial, &m->losses})
    if (x->p) (void)hipFree(x->p);
  (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int gt_mlp_get_params(gt_mlp* m, float* theta) {
  if (!m || !theta) return mfail("null argument");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_HIP(hipMemcpyAsync(theta, m->b.theta, kMlpParams * 4, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_mlp_set_params(gt_mlp* m, const float* theta) {
  if (!m || !theta) return mfail("null argument");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_HIP(hipMemcpyAsync(m->b.theta, theta, kMlpParams * 4, hipMemcpyHostToDevice, m->stream));
  ML_HIP(launch_mlp_cast(m->b, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_mlp_features(gt_mlp* m, const uint64_t* masks, int32_t n, uint16_t* feat) {
  if (!m || !masks || !feat || n < 0) return mfail("bad argument");
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_OK(m->upload(masks, nullptr, n));
  ML_OK(m->features(m->masks.as<uint64_t>(), n, m->stream));
  ML_HIP(hipMemcpyAsync(feat, m->feat.p, (size_t)n * kMlpIn * 2, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_mlp_grads(gt_mlp* m, const uint64_t* masks, const float* target, int32_t n, float* grads,
                 float* loss) {
  if (!m || !masks || !target || n <= 0) return mfail("bad argument");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_OK(m->upload(masks, target, n));
  ML_OK(m->grads(m->masks.as<uint64_t>(), m->target.as<float>(), n, m->stream));
  std::vector<float> g(kMlpParams + 1);
  ML_HIP(hipMemcpyAsync(g.data(), m->b.grads, g.size() * 4, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  if (grads) std::memcpy(grads, g.data(), kMlpParams * 4);
  if (loss) *loss = g[kMlpParams];
  return 0;
}

int gt_mlp_train_step_dev(gt_mlp* m, const uint64_t* d_masks, const float* d_target, int32_t n,
                          float lr, void* stream) {
  if (!m || !d_masks || !d_target || n <= 0) return mfail("bad argument");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : m->stream;
  ML_OK(m->train_step(d_masks, d_target, n, lr, nullptr, s));
  return 0;
}

int gt_mlp_train_step(gt_mlp* m, const uint64_t* masks, const float* target, int32_t n, float lr,
                      float* loss) {
  if (!m || !masks || !target || n <= 0) return mfail("bad argument");
  {
    std::lock_guard<std::mutex> lk(m->mu);
    ML_HIP(hipSetDevice(m->device));
    ML_OK(m->upload(masks, target, n));
  }
  ML_OK(gt_mlp_train_step_dev(m, m->masks.as<uint64_t>(), m->target.as<float>(), n, lr, nullptr));
  std::lock_guard<std::mutex> lk(m->mu);
  if (loss) ML_HIP(hipMemcpyAsync(loss, m->b.grads + kMlpParams, 4, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_mlp_train_batches(gt_mlp* m, const uint64_t* masks, const float* target, int32_t n,
                         int32_t batch, float lr, float* losses) {
  if (!m || !masks || !target || n <= 0 || batch <= 0) return mfail("bad argument");
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  ML_OK(m->upload(masks, target, n));
  const int nb = (n + batch - 1) / batch;
  ML_OK(m->losses.ensure((size_t)nb * 4));
  // Features of all rows in one launch, then the steps back to back; each
  // writes its loss to its own slot.
  ML_OK(m->features(m->masks.as<uint64_t>(), n, m->stream));
  for (int b = 0; b < nb; ++b) {
    const int r0 = b * batch, rows = std::min(batch, n - r0);
    ML_OK(m->train_step_feat(m->feat.as<uint16_t>() + (size_t)r0 * kMlpIn,
                             m->target.as<float>() + r0, rows, lr, m->losses.as<float>() + b,
                             m->stream));
  }
  if (losses)
    ML_HIP(hipMemcpyAsync(losses, m->losses.p, (size_t)nb * 4, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_mlp_predict_dev(gt_mlp* m, const uint64_t* d_masks, int32_t n, float* d_out, void* stream) {
  if (!m || !d_masks || !d_out || n < 0) return mfail("bad argument");
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipSetDevice(m->device));
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : m->stream;
  ML_OK(m->features(d_masks, n, s));
  ML_HIP(launch_mlp_predict(m->b, m->feat.as<uint16_t>(), n, d_out, s));
  return 0;
}

int gt_mlp_predict(gt_mlp* m, const uint64_t* masks, int32_t n, float* out) {
  if (!m || !masks || !out || n < 0) return mfail("bad argument");
  if (n == 0) return 0;
  {
    std::lock_guard<std::mutex> lk(m->mu);
    ML_HIP(hipSetDevice(m->device));
    ML_OK(m->upload(masks, nullptr, n));
    ML_OK(m->out.ensure((size_t)n * 4));
  }
  ML_OK(gt_mlp_predict_dev(m, m->masks.as<uint64_t>(), n, m->out.as<float>(), nullptr));
  std::lock_guard<std::mutex> lk(m->mu);
  ML_HIP(hipMemcpyAsync(out, m->out.p, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
  ML_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int gt_bf16_gemm(int32_t device, const uint16_t* A, const uint16_t* B, float* C, int32_t M,
                 int32_t N, int32_t K) {
  if (!A || !B || !C) return mfail("null argument");
  if (M <= 0 || N <= 0 || K <= 0 || M % 16 || N % 16 || K % 16)
    return mfail("M, N, K must be positive multiples of 16");
  ML_HIP(hipSetDevice(device));
  void *dA = nullptr, *dB = nullptr, *dC = nullptr;
  hipStream_t s = nullptr;
  auto cleanup = [&]() {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dC) (void)hipFree(dC);
    if (s) (void)hipStreamDestroy(s);
  };
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&dA, (size_t)M * K * 2) != hipSuccess || hipMalloc(&dB, (size_t)K * N * 2) != hipSuccess ||
      hipMalloc(&dC, (size_t)M * N * 4) != hipSuccess ||
      hipMemcpyAsync(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice, s) != hipSuccess ||
      launch_bf16_gemm((const uint16_t*)dA, (const uint16_t*)dB, (float*)dC, M, N, K, s) != hipSuccess ||
      hipMemcpyAsync(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    return mfail("bf16 GEMM failed");
  }
  cleanup();
  return 0;
}

}  // extern "C"
