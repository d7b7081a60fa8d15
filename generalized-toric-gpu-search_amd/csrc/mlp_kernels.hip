// Surrogate MLP that ranks exponent sets S before exact evaluation.
//
// Network: x (128 inputs: 49 mask bits, |S|/16, 48 difference-multiset bins,
// zero padding) -> 128 ReLU -> 128 ReLU -> 1.  Weights are kept in fp32
// (Adam state) with bf16 copies that feed the matrix cores; every product
// accumulates in fp32 (v_mfma_f32_16x16x16_bf16).  Activations and
// back-propagated errors are rounded to bf16 where they become MFMA operands;
// the CPU reference (oracle/mlp_ref.py) applies the same roundings.
//
// Training step on n rows: each workgroup walks a run of consecutive 32-row
// tiles, with the forward pass, the loss derivative and the backward pass of
// one tile at a time in LDS, and accumulates the tiles' weight gradients in
// registers (the dW1 and dW2 MFMA accumulators of its waves persist across
// tiles); it writes one gradient partial.  A second kernel sums the partials
// in a fixed order (deterministic), applies Adam and refreshes the bf16
// copies.
#include "gt_mlp.h"

namespace gt {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                       // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C[M][N] = A[M][K] * B[K][N] over 16x16 output tiles shared by the waves of
// the workgroup; A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn]
// (bf16, LDS or global).  MFMA 16x16x16 lane layout: lane l supplies
// A(l % 16, 4 * (l / 16) + e) and B(4 * (l / 16) + e, l % 16), e = 0..3, and
// receives C(4 * (l / 16) + e, l % 16).  epi(m, n, value) consumes each result.
// kAk / kBk: A / B is contiguous along k (sak / sbk = 1, rows 8-byte
// aligned), so a lane's four values are one 8-byte load.  kK > 0: K is that
// compile-time constant and the k loop is unrolled, so the fragment loads of
// all k-steps are in flight before the MFMAs wait.
template <bool kAk, bool kBk, int kK, class Epi>
__device__ __forceinline__ void wg_gemm(const uint16_t* A, int sam, int sak, const uint16_t* B,
                                        int sbk, int sbn, int M, int N, int K, Epi epi) {
  const int nw = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int tn = N / 16, tiles = (M / 16) * tn;
  const int r = lane & 15, q = lane >> 4;
  for (int t = wave; t < tiles; t += nw) {
    const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
    f4v acc = {0.f, 0.f, 0.f, 0.f};
    auto step = [&](int k0) {
      s4v a, b;
      if constexpr (kAk) {
        a = *reinterpret_cast<const s4v*>(A + (m0 + r) * sam + k0 + 4 * q);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = (short)A[(m0 + r) * sam + (k0 + 4 * q + e) * sak];
      }
      if constexpr (kBk) {
        b = *reinterpret_cast<const s4v*>(B + (n0 + r) * sbn + k0 + 4 * q);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = (short)B[(k0 + 4 * q + e) * sbk + (n0 + r) * sbn];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    };
    if constexpr (kK > 0) {
#pragma unroll
      for (int k0 = 0; k0 < kK; k0 += 16) step(k0);
    } else {
      for (int k0 = 0; k0 < K; k0 += 16) step(k0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) epi(m0 + 4 * q + e, n0 + r, acc[e]);
  }
}

// ------------------------------------------------------------- features --

// Features of a set: mask bits, |S| / 16, and for each nonzero torus
// difference d = (du, dv) (index 1..48) the number of ordered pairs (s, t) of
// S with s - t = d, divided by |S|.
//
// One wave per set.  The pair count for d is popcount(S & (S + d)), S + d
// being the mask translated by d: the seven 7-bit rows rotate by du rows
// (a rotation of the 49-bit mask by 7 du) and each row's bits by dv.  Lane l
// writes features l and 64 + l.
constexpr uint64_t kF49 = (1ull << kNpts) - 1ull;

GT_HD uint64_t rep7(uint64_t row) {   // the 7-bit pattern `row` in all seven rows
  uint64_t r = 0;
  for (int u = 0; u < 7; ++u) r |= row << (7 * u);
  return r;
}

__device__ __forceinline__ uint32_t pairs_at(uint64_t m, int d) {
  const int du = d / 7, dv = d % 7;
  uint64_t t = du ? (((m << (7 * du)) | (m >> (kNpts - 7 * du))) & kF49) : m;
  if (dv) {
    const uint64_t low = rep7((1ull << (7 - dv)) - 1ull);   // columns 0 .. 6 - dv
    t = ((t & low) << dv) | ((t & ~low & kF49) >> (7 - dv));
  }
  return (uint32_t)__popcll(m & t);
}

__global__ __launch_bounds__(256) void mlp_features_kernel(const uint64_t* __restrict__ masks,
                                                           int32_t n,
                                                           uint16_t* __restrict__ feat,
                                                           const uint32_t* __restrict__ gate,
                                                           uint32_t gate_min) {
  const int lane = (int)(threadIdx.x & 63u);
  const int set = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (set >= n || (gate != nullptr && *gate <= gate_min)) return;
  const uint64_t m = masks[set] & kF49;
  const int k = __popcll(m);
  const float invk = (k > 0) ? 1.0f / (float)k : 0.0f;
  float lo;
  if (lane < kNpts) lo = ((m >> lane) & 1ull) ? 1.0f : 0.0f;
  else if (lane == kNpts) lo = (float)k / 16.0f;
  else lo = (float)pairs_at(m, lane - kNpts) * invk;           // d = 1 .. 14
  const int d2 = lane + 64 - kNpts;                              // d = 15 .. 78
  const float hi = (d2 < kNpts) ? (float)pairs_at(m, d2) * invk : 0.0f;
  uint16_t* x = feat + (size_t)set * kMlpIn;
  x[lane] = f2bf(lo);
  x[64 + lane] = f2bf(hi);
}

// ------------------------------------------------------------ train step --

struct MlpView {
  const float* theta;      // fp32 master parameters
  const uint16_t* w1b;     // [128][128] bf16
  const uint16_t* w2b;     // [128][128] bf16
  const uint16_t* w3b;     // [128] bf16
};

// Row stride (bf16 elements) of the 128-wide matrices held in LDS.  Rows 272
// bytes apart shift by 4 banks, so a 16-lane group reading 16 rows (the MFMA
// fragment loads, 8 bytes per lane) or 4 rows of 16 columns (the transposed
// reads of the weight-gradient products) hits distinct banks; with 256-byte
// rows every such group read one bank.
constexpr int kLd = 136;

// Forward pass of one tile into LDS: X (loaded), H1, H2 (bf16, row stride
// kLd), y (fp32).  W.w1b / W.w2b have row stride kWs (128 in global memory,
// kLd when staged in LDS).
template <int kWs>
__device__ __forceinline__ void tile_forward(const MlpView& W, const uint16_t* __restrict__ feat,
                                             int row0, int n, uint16_t* X, uint16_t* H1,
                                             uint16_t* H2, float* y) {
  // Four features per 8-byte load (global rows are 256 bytes).
  for (int e = threadIdx.x; e < kMlpTile * kMlpIn / 4; e += blockDim.x) {
    const int m = e / (kMlpIn / 4), i4 = e % (kMlpIn / 4);
    uint2 v = make_uint2(0u, 0u);
    if (row0 + m < n) v = *reinterpret_cast<const uint2*>(feat + (size_t)(row0 + m) * kMlpIn + 4 * i4);
    *reinterpret_cast<uint2*>(X + m * kLd + 4 * i4) = v;
  }
  __syncthreads();
  const float* b1 = W.theta + kOffB1;
  const float* b2 = W.theta + kOffB2;
  wg_gemm<true, true, kMlpIn>(X, kLd, 1, W.w1b, 1, kWs, kMlpTile, kMlpHid, kMlpIn,
          [&](int m, int h, float acc) { H1[m * kLd + h] = f2bf(fmaxf(acc + b1[h], 0.f)); });
  __syncthreads();
  wg_gemm<true, true, kMlpHid>(H1, kLd, 1, W.w2b, 1, kWs, kMlpTile, kMlpHid, kMlpHid,
          [&](int m, int h, float acc) { H2[m * kLd + h] = f2bf(fmaxf(acc + b2[h], 0.f)); });
  __syncthreads();
  {
    // y[m] = b3 + sum_h w3[h] H2[m][h]: eight lanes per row, 16 terms each,
    // then a butterfly over the eight (every lane ends with the same sum).
    static_assert(kMlpTile * 8 == 256 && kMlpHid == 8 * 16, "256 threads: 32 rows x 8 parts");
    const int m = (int)threadIdx.x >> 3, part = (int)threadIdx.x & 7;
    float s = 0.f;
#pragma unroll
    for (int h = part * 16; h < part * 16 + 16; ++h)
      s += bf2f(W.w3b[h]) * bf2f(H2[m * kLd + h]);
    s += __shfl_xor(s, 1, 8);
    s += __shfl_xor(s, 2, 8);
    s += __shfl_xor(s, 4, 8);
    if (part == 0) y[m] = W.theta[kOffB3] + s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void mlp_predict_kernel(MlpView W,
                                                          const uint16_t* __restrict__ feat,
                                                          int32_t n, float* __restrict__ out,
                                                          const uint32_t* __restrict__ gate,
                                                          uint32_t gate_min) {
  __shared__ __attribute__((aligned(16))) uint16_t X[kMlpTile * kLd], H1[kMlpTile * kLd],
      H2[kMlpTile * kLd];
  __shared__ float y[kMlpTile];
  if (gate != nullptr && *gate <= gate_min) return;   // uniform over the workgroup
  const int row0 = blockIdx.x * kMlpTile;
  tile_forward<kMlpIn>(W, feat, row0, n, X, H1, H2, y);
  if (threadIdx.x < kMlpTile && row0 + (int)threadIdx.x < n) out[row0 + threadIdx.x] = y[threadIdx.x];
}

// dC[o][i] += sum_m A[m][o] B[m][i] over one tile (m < kMlpTile rows), for
// the 16 output tiles of this wave (o, i tiles t = wave + 4 j, column-major
// over 8 x 8 tiles): A and B are row-major bf16 in LDS with row strides sa
// and sb; acc[j] holds the wave's j-th 16x16 tile (MFMA layout).
template <int kSa, int kSb>
__device__ __forceinline__ void acc_tn_gemm(const uint16_t* A, const uint16_t* B, f4v* acc) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int t = wave + 4 * j;
    const int o0 = (t / 8) * 16, i0 = (t % 8) * 16;
#pragma unroll
    for (int k0 = 0; k0 < kMlpTile; k0 += 16) {
      s4v a, b;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[e] = (short)A[(k0 + 4 * q + e) * kSa + o0 + r];
        b[e] = (short)B[(k0 + 4 * q + e) * kSb + i0 + r];
      }
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
    }
  }
}

template <int kCols>
__device__ __forceinline__ void store_acc(const f4v* acc, float* P) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int t = wave + 4 * j;
    const int o0 = (t / 8) * 16, i0 = (t % 8) * 16;
#pragma unroll
    for (int e = 0; e < 4; ++e) P[(o0 + 4 * q + e) * kCols + i0 + r] = acc[j][e];
  }
}

// Loss 0.5 * mean((y - t)^2).  Workgroup b takes tiles b * per .. b * per +
// per - 1 (in order) and writes their summed gradient to partial[b][0 ..
// kMlpParams) and their summed squared error to partial[b][kMlpParams].
__global__ __launch_bounds__(256) void mlp_grad_kernel(MlpView W,
                                                       const uint16_t* __restrict__ feat,
                                                       const float* __restrict__ target,
                                                       int32_t n, int32_t per,
                                                       float* __restrict__ partial) {
  static_assert(kMlpIn == 128 && kMlpHid == 128, "8 x 8 output tiles per weight matrix");
  __shared__ __attribute__((aligned(16))) uint16_t X[kMlpTile * kLd], H1[kMlpTile * kLd],
      H2[kMlpTile * kLd];
  __shared__ __attribute__((aligned(16))) uint16_t D2[kMlpTile * kLd], D1[kMlpTile * kLd];
  __shared__ __attribute__((aligned(16))) uint16_t W1s[kMlpHid * kLd], W2s[kMlpHid * kLd];
  __shared__ float y[kMlpTile], dy[kMlpTile], sq[kMlpTile];
  // The bf16 weights stay in LDS for all of this workgroup's tiles: every
  // MFMA fragment of the forward and backward products is then an LDS read.
  static_assert(kMlpHid * kMlpIn == kMlpHid * kMlpHid, "W1 and W2 copied together");
  for (int e = threadIdx.x; e < kMlpHid * kMlpIn / 8; e += blockDim.x) {
    const int row = e / (kMlpIn / 8), c8 = e % (kMlpIn / 8);
    *reinterpret_cast<uint4*>(W1s + row * kLd + 8 * c8) = reinterpret_cast<const uint4*>(W.w1b)[e];
    *reinterpret_cast<uint4*>(W2s + row * kLd + 8 * c8) = reinterpret_cast<const uint4*>(W.w2b)[e];
  }
  // (tile_forward's first barrier orders these stores before the products.)
  MlpView WL = W;
  WL.w1b = W1s;
  WL.w2b = W2s;
  f4v g1[16], g2[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    g1[j] = f4v{0.f, 0.f, 0.f, 0.f};
    g2[j] = f4v{0.f, 0.f, 0.f, 0.f};
  }
  // Threads h < 128 keep the column sums of unit h; thread 0 the scalars.
  float gw3 = 0.f, gb2 = 0.f, gb1 = 0.f, gb3 = 0.f, sse = 0.f;
  const int h = threadIdx.x;
  const int tiles = (n + kMlpTile - 1) / kMlpTile;
  const int t_end = min(tiles, (int)(blockIdx.x + 1) * per);
  for (int tile = blockIdx.x * per; tile < t_end; ++tile) {
    const int row0 = tile * kMlpTile;
    tile_forward<kLd>(WL, feat, row0, n, X, H1, H2, y);
    if (threadIdx.x < kMlpTile) {
      const int m = threadIdx.x;
      const bool ok = row0 + m < n;
      const float diff = ok ? (y[m] - target[row0 + m]) : 0.f;
      dy[m] = diff / (float)n;
      sq[m] = diff * diff;
    }
    __syncthreads();
    // dH2 = dy * w3 * relu'(H2), rounded to bf16.
    for (int e = threadIdx.x; e < kMlpTile * kMlpHid; e += blockDim.x) {
      const int m = e / kMlpHid, hh = e % kMlpHid;
      const float g = (H2[m * kLd + hh] != 0) ? dy[m] * bf2f(W.w3b[hh]) : 0.f;
      D2[m * kLd + hh] = f2bf(g);
    }
    __syncthreads();
    if (h < kMlpHid) {
      for (int m = 0; m < kMlpTile; ++m) {
        gw3 += dy[m] * bf2f(H2[m * kLd + h]);
        gb2 += bf2f(D2[m * kLd + h]);
      }
    }
    if (threadIdx.x == 0) {
      for (int m = 0; m < kMlpTile; ++m) { gb3 += dy[m]; sse += sq[m]; }
    }
    // dH1 = (dH2 W2) * relu'(H1), rounded to bf16.
    wg_gemm<true, false, kMlpHid>(D2, kLd, 1, W2s, kLd, 1, kMlpTile, kMlpHid, kMlpHid,
            [&](int m, int hh, float acc) {
              D1[m * kLd + hh] = (H1[m * kLd + hh] != 0) ? f2bf(acc) : (uint16_t)0;
            });
    __syncthreads();
    // dW2[o][i] += sum_m dH2[m][o] H1[m][i];  dW1[o][i] += sum_m dH1[m][o] X[m][i].
    acc_tn_gemm<kLd, kLd>(D2, H1, g2);
    acc_tn_gemm<kLd, kLd>(D1, X, g1);
    if (h < kMlpHid)
      for (int m = 0; m < kMlpTile; ++m) gb1 += bf2f(D1[m * kLd + h]);
    __syncthreads();   // the next tile overwrites the LDS operands
  }
  float* P = partial + (size_t)blockIdx.x * kMlpPartialStride;
  store_acc<kMlpHid>(g2, P + kOffW2);
  store_acc<kMlpIn>(g1, P + kOffW1);
  if (h < kMlpHid) {
    P[kOffW3 + h] = gw3;
    P[kOffB2 + h] = gb2;
    P[kOffB1 + h] = gb1;
  }
  if (threadIdx.x == 0) {
    P[kOffB3] = gb3;
    P[kMlpParams] = sse;
  }
}

// grads[p] = sum over the partials in index order (deterministic); slot
// kMlpParams is the loss 0.5 * sum (y - t)^2 / n.
__global__ void mlp_reduce_kernel(const float* __restrict__ partial, int32_t parts, int32_t n,
                                  float* __restrict__ grads) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float s = 0.f;
  for (int t = 0; t < parts; ++t) s += partial[(size_t)t * kMlpPartialStride + p];
  grads[p] = (p == kMlpParams) ? 0.5f * s / (float)n : s;
}

// One Adam step fused with the reduction: grads[p] = sum over the partials in
// index order (the same sums as mlp_reduce_kernel), then the update of p and
// its bf16 copy.  Slot kMlpParams takes the loss, also copied to *loss_out.
__global__ void mlp_reduce_adam_kernel(const float* __restrict__ partial, int32_t tiles, int32_t n,
                                       float* __restrict__ theta, float* __restrict__ grads,
                                       float* __restrict__ m1, float* __restrict__ m2, float lr,
                                       float beta1, float beta2, float eps, float c1, float c2,
                                       uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b,
                                       uint16_t* __restrict__ w3b, float* __restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float g = 0.f;
  for (int t = 0; t < tiles; ++t) g += partial[(size_t)t * kMlpPartialStride + p];
  if (p == kMlpParams) {
    const float loss = 0.5f * g / (float)n;
    grads[p] = loss;
    if (loss_out) *loss_out = loss;
    return;
  }
  grads[p] = g;
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float w = theta[p] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  theta[p] = w;
  if (p < kOffB1) w1b[p - kOffW1] = f2bf(w);
  else if (p >= kOffW2 && p < kOffB2) w2b[p - kOffW2] = f2bf(w);
  else if (p >= kOffW3 && p < kOffB3) w3b[p - kOffW3] = f2bf(w);
}

// t[i] = (d[i] - mean) / std over the n targets (std 1 when all are equal),
// one workgroup; sums in double, in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void mlp_standardize_kernel(const int32_t* __restrict__ d,
                                                               int32_t n, float* __restrict__ t) {
  __shared__ double s1[1024], s2[1024];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = (double)d[i];
    a += x;
    b += x * x;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  const double mean = s1[0] / (double)n;
  const double var = s2[0] / (double)n - mean * mean;
  const double sd = (var > 1e-12) ? sqrt(var) : 1.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t[i] = (float)(((double)d[i] - mean) / sd);
}

__global__ void mlp_cast_kernel(const float* __restrict__ theta, uint16_t* __restrict__ w1b,
                                uint16_t* __restrict__ w2b, uint16_t* __restrict__ w3b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kMlpParams) return;
  const float w = theta[p];
  if (p < kOffB1) w1b[p - kOffW1] = f2bf(w);
  else if (p >= kOffW2 && p < kOffB2) w2b[p - kOffW2] = f2bf(w);
  else if (p >= kOffW3 && p < kOffB3) w3b[p - kOffW3] = f2bf(w);
}

// Plain bf16 GEMM through wg_gemm (row-major A[M][K], B[K][N] -> fp32 C[M][N]),
// one workgroup per 64 x 64 block of C; M, N, K multiples of 16.
__global__ __launch_bounds__(256) void bf16_gemm_kernel(const uint16_t* __restrict__ A,
                                                        const uint16_t* __restrict__ B,
                                                        float* __restrict__ C, int M, int N,
                                                        int K) {
  const int bm = blockIdx.y * 64, bn = blockIdx.x * 64;
  const int mm = min(64, M - bm), nn = min(64, N - bn);
  wg_gemm<true, false, 0>(A + (size_t)bm * K, K, 1, B + bn, N, 1, mm, nn, K,
          [&](int m, int n2, float acc) { C[(size_t)(bm + m) * N + bn + n2] = acc; });
}

// ------------------------------------------------------------- launchers --

static MlpView view(const MlpBufs& b) { return MlpView{b.theta, b.w1b, b.w2b, b.w3b}; }

hipError_t launch_mlp_features(const uint64_t* d_masks, int32_t n, uint16_t* d_feat,
                               hipStream_t s, const uint32_t* d_gate, uint32_t gate_min) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_features_kernel, dim3((n + 3) / 4), dim3(256), 0, s, d_masks, n, d_feat,
                     d_gate, gate_min);
  return hipGetLastError();
}

hipError_t launch_mlp_predict(const MlpBufs& b, const uint16_t* d_feat, int32_t n, float* d_out,
                              hipStream_t s, const uint32_t* d_gate, uint32_t gate_min) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_predict_kernel, dim3((n + kMlpTile - 1) / kMlpTile), dim3(256), 0, s,
                     view(b), d_feat, n, d_out, d_gate, gate_min);
  return hipGetLastError();
}

int mlp_grad_parts(int32_t n) {
  const int tiles = (n + kMlpTile - 1) / kMlpTile;
  const int per = (tiles + kMlpMaxParts - 1) / kMlpMaxParts;
  return (tiles + per - 1) / per;
}

hipError_t launch_mlp_grads(const MlpBufs& b, const uint16_t* d_feat, const float* d_target,
                            int32_t n, float* d_partial, hipStream_t s) {
  const int tiles = (n + kMlpTile - 1) / kMlpTile;
  const int per = (tiles + kMlpMaxParts - 1) / kMlpMaxParts, parts = mlp_grad_parts(n);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(parts), dim3(256), 0, s, view(b), d_feat, d_target, n,
                     per, d_partial);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s,
                     d_partial, parts, n, b.grads);
  return hipGetLastError();
}

hipError_t launch_mlp_train_step(const MlpBufs& b, const uint16_t* d_feat, const float* d_target,
                                 int32_t n, float* d_partial, float lr, int step, float* d_loss,
                                 hipStream_t s) {
  const int tiles = (n + kMlpTile - 1) / kMlpTile;
  const int per = (tiles + kMlpMaxParts - 1) / kMlpMaxParts, parts = mlp_grad_parts(n);
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(parts), dim3(256), 0, s, view(b), d_feat, d_target, n,
                     per, d_partial);
  hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s,
                     d_partial, parts, n, b.theta, b.grads, b.m1, b.m2, lr, beta1, beta2, eps, c1,
                     c2, b.w1b, b.w2b, b.w3b, d_loss);
  return hipGetLastError();
}

hipError_t launch_mlp_standardize(const int32_t* d_dist, int32_t n, float* d_target,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_standardize_kernel, dim3(1), dim3(1024), 0, s, d_dist, n, d_target);
  return hipGetLastError();
}

hipError_t launch_mlp_cast(const MlpBufs& b, hipStream_t s) {
  hipLaunchKernelGGL(mlp_cast_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, b.theta,
                     b.w1b, b.w2b, b.w3b);
  return hipGetLastError();
}

hipError_t launch_bf16_gemm(const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K,
                            hipStream_t s) {
  hipLaunchKernelGGL(bf16_gemm_kernel, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, A, B,
                     C, M, N, K);
  return hipGetLastError();
}

}  // namespace gt
