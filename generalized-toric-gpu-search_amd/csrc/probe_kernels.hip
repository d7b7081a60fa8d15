// Issue-rate probes: straight-line streams of the instruction forms the Gray
// walk issues, and the walk's own 16-step block, timed over the whole device.
// They measure what the instruction mix of the walk can attain on this part
// (the roofline's attainable ceiling), next to the nominal INT32 rate of one
// wave64 instruction per 4 clocks per SIMD.  Registers only: each thread keeps
// eight independent accumulators and stores them once at the end.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/gtoric.h"
#include "gt_launch.h"
#include "gt_walk.h"

namespace gt {

constexpr int kProbeChains = 8;     // independent accumulators per thread
constexpr int kProbeUnroll = 16;    // iterations unrolled per loop trip

enum ProbeForm : int {
  kPrXorVV = 0,   // v_xor_b32 v, v, v (VOP2, two VGPR operands)
  kPrXorSV,       // v_xor_b32 v, s, v (VOP2, SGPR operand)
  kPrOr3,         // v_or3_b32 (VOP3)
  kPrBcnt,        // v_bcnt_u32_b32 (VOP3)
  kPrMin3,        // v_min3_u32 (VOP3)
  kPrMovS,        // v_mov_b32 v, s
  kPrWalk,        // the walk's 16-step block (GT_STEP2 x 8, rows in VGPRs)
  kProbeForms
};

// Wave64 instructions of the probed form per thread per loop trip.
__host__ __device__ constexpr int probe_instr_per_trip(int form) {
  return form == kPrWalk ? 174 : kProbeChains * kProbeUnroll;
}

template <int FORM>
__global__ __launch_bounds__(256) void issue_probe_kernel(int trips, uint32_t seed, uint32_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[kProbeChains];
#pragma unroll
  for (int c = 0; c < kProbeChains; ++c) a[c] = seed * (t + 1) + 0x9E3779B9u * (c + 1);
  const uint32_t b = seed ^ (t * 0x85EBCA6Bu);
  const uint32_t sb = __builtin_amdgcn_readfirstlane(seed * 0xC2B2AE35u);
  if constexpr (FORM == kPrWalk) {
    // The walk block on register rows: same code as gray_chunk's inner loop.
    RowV v, s0, s1, s2, s3, d;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      v.w[i] = a[i];
      s0.w[i] = a[i] * 3u + b;
      s1.w[i] = a[i] * 5u ^ b;
      s2.w[i] = a[i] * 7u + 1u;
      s3.w[i] = a[i] ^ 0x5A5A5A5Au;
      d.w[i] = a[i] + b * 9u;
    }
    uint32_t best = 64u;
    for (int m = 0; m < trips; ++m) {
      GT_STEP2(s0, s1)
      GT_STEP2(s0, s2)
      GT_STEP2(s0, s1)
      GT_STEP2(s0, s3)
      GT_STEP2(s0, s1)
      GT_STEP2(s0, s2)
      GT_STEP2(s0, s1)
      GT_STEP2(s0, d)
      // The block's row for the next trip is copied like the walk's d.
#pragma unroll
      for (int i = 0; i < 6; ++i) asm volatile("v_mov_b32 %0, %1" : "=v"(d.w[i]) : "s"(sb ^ (uint32_t)m));
    }
    out[t] = best ^ v.w[0] ^ v.w[5];
    return;
  } else {
    for (int m = 0; m < trips; ++m) {
#pragma unroll
      for (int u = 0; u < kProbeUnroll; ++u) {
#pragma unroll
        for (int c = 0; c < kProbeChains; ++c) {
          if constexpr (FORM == kPrXorVV) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b));
          if constexpr (FORM == kPrXorSV) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(a[c]) : "s"(sb));
          if constexpr (FORM == kPrOr3) asm volatile("v_or3_b32 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "v"(a[(c + 1) % kProbeChains]));
          if constexpr (FORM == kPrBcnt) asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a[c]) : "v"(b));
          if constexpr (FORM == kPrMin3) asm volatile("v_min3_u32 %0, %0, %1, %2" : "+v"(a[c]) : "v"(b), "v"(a[(c + 1) % kProbeChains]));
          if constexpr (FORM == kPrMovS) asm volatile("v_mov_b32 %0, %1" : "=v"(a[c]) : "s"(sb));
        }
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < kProbeChains; ++c) r ^= a[c];
  out[t] = r;
}

}  // namespace gt

using namespace gt;

extern "C" int gt_issue_probe(int32_t device, int32_t form, int32_t trips, double* ms,
                              double* wave_instr) {
  if (form < 0 || form >= kProbeForms) return set_error("issue probe: unknown form");
  if (trips < 1 || trips > (1 << 20)) return set_error("issue probe: trips outside 1..2^20");
  if (!ms || !wave_instr) return set_error("issue probe: null output");
  if (hipSetDevice(device) != hipSuccess) return set_error("issue probe: hipSetDevice failed");
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return set_error("issue probe: CU count query failed");
  // Eight waves per SIMD: 8 workgroups of 4 waves per CU.
  const int grid = cu * 8, block = 256;
  uint32_t* out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  auto launch = [&](int tr) {
    switch (form) {
      case kPrXorVV: hipLaunchKernelGGL(issue_probe_kernel<kPrXorVV>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      case kPrXorSV: hipLaunchKernelGGL(issue_probe_kernel<kPrXorSV>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      case kPrOr3: hipLaunchKernelGGL(issue_probe_kernel<kPrOr3>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      case kPrBcnt: hipLaunchKernelGGL(issue_probe_kernel<kPrBcnt>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      case kPrMin3: hipLaunchKernelGGL(issue_probe_kernel<kPrMin3>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      case kPrMovS: hipLaunchKernelGGL(issue_probe_kernel<kPrMovS>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
      default: hipLaunchKernelGGL(issue_probe_kernel<kPrWalk>, dim3(grid), dim3(block), 0, s, tr, 7u, out); break;
    }
  };
  float t = 0.f;
  if (hipMalloc(&out, (size_t)grid * block * 4) != hipSuccess) return set_error("issue probe: hipMalloc failed");
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = set_error("issue probe: stream/event creation failed");
  } else {
    launch(trips);   // warm: clocks up, code loaded
    (void)hipEventRecord(e0, s);
    launch(trips);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess ||
        hipGetLastError() != hipSuccess)
      rc = set_error("issue probe: launch or timing failed");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(out);
  if (rc) return rc;
  *ms = t;
  *wave_instr = (double)grid * (block / kWave) * (double)trips * probe_instr_per_trip(form);
  return 0;
}
