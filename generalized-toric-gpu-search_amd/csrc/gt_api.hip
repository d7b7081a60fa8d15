// C ABI of libgtoric.so (declared in include/gtoric.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gtoric.h"
#include "gt_bfs.h"
#include "gt_evolve.h"
#include "gt_launch.h"
#include "gt_mlp.h"

using namespace gt;

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define GT_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) return fail("%s: %s", #call, hipGetErrorString(e_));    \
  } while (0)

#define GT_OK(call)            \
  do {                         \
    int r_ = (call);           \
    if (r_ != 0) return r_;    \
  } while (0)

// Reallocation state shared by a context's buffers: captured graphs hold raw
// buffer addresses, so growth is refused while capturing and counted so that
// a graph built before a reallocation refuses to launch.
struct BufState {
  uint64_t gen = 0;
  bool frozen = false;
};

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  BufState* st = nullptr;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (st && st->frozen)
      return fail("buffer growth to %zu bytes during graph capture; run the same call once "
                  "before capturing", bytes);
    // Only moving an existing allocation can invalidate a captured graph.
    if (st && p) ++st->gen;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return fail("hipFree: %s", hipGetErrorString(e));
    }
    size_t want = std::max<size_t>(bytes, 256);
    want = (want + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) { p = nullptr; return fail("hipMalloc(%zu): %s", want, hipGetErrorString(e)); }
    cap = want;
    return 0;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
  void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
};

// ------------------------------------------------------------- GF(8) ----
// Field GF(2)[x]/(x^3 + x + 1) with elements 0..7 in the polynomial basis.
int gf8_mul(int a, int b) {
  int r = 0;
  for (int i = 0; i < 3; ++i) {
    if (b & 1) r ^= a;
    a <<= 1;
    if (a & 8) a ^= 0xB;
    b >>= 1;
  }
  return r;
}
int gf8_pow(int a, int e) {
  int r = 1;
  for (int i = 0; i < e; ++i) r = gf8_mul(r, a);
  return r;
}

// Bit planes of the 49 evaluation rows: row i = (a, b) = divmod(i, 7)
// evaluated at torus point j = (x, y) = (j / 7 + 1, j % 7 + 1).
void build_mplanes(uint32_t out[kNpts * kRowWords]) {
  std::memset(out, 0, sizeof(uint32_t) * kNpts * kRowWords);
  for (int i = 0; i < kNpts; ++i) {
    const int a = i / 7, b = i % 7;
    uint64_t pl[3] = {0, 0, 0};
    for (int j = 0; j < kNpts; ++j) {
      const int x = j / 7 + 1, y = j % 7 + 1;
      const int v = gf8_mul(gf8_pow(x, a), gf8_pow(y, b));
      for (int p = 0; p < 3; ++p)
        if ((v >> p) & 1) pl[p] |= 1ull << j;
    }
    for (int p = 0; p < 3; ++p) {
      out[i * kRowWords + 2 * p] = (uint32_t)pl[p];
      out[i * kRowWords + 2 * p + 1] = (uint32_t)(pl[p] >> 32);
    }
  }
}

// Default Gray walk bits per k (a level-1 walk has dimension k - 1): big
// enough that per-chunk set-up is a few percent of the walk, small enough
// for a balanced tail; for k >= 19 large enough that a code has at most
// about 2^24 chunks (the chunk ids of one call must stay below 2^31).
int default_L(int k) {
  static const uint8_t tab[kMaxKExact + 1] = {4, 4, 4, 4, 4, 4, 4, 7, 8, 10, 11, 12,
                                              13, 14, 15, 16, 17, 18, 19, 21, 24, 27};
  return tab[k];
}

struct ProfRec {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

// Plan area of one exact search: claim counters (kClaimShards x 4 KB), then
// a 512-byte slot holding MdPlan and, in its last words, the final- and
// middle-pass list counts and md_fused's code count, then six arrays of 4
// bytes per code: order, row offsets, aux words, the final-pass list, the
// list of codes md_fused leaves and the middle-pass list.  The count of the
// list md_fused leaves lives apart (it must read 0 before a call's first
// kernel, and the plan area is reallocated as batches grow).
constexpr size_t kMdClaimBytes = (size_t)kClaimShards * kClaimStride * 4;
constexpr size_t kMdPlanAreaBytes = kMdClaimBytes + 512;
constexpr size_t kMdFbCountOff = kMdPlanAreaBytes - 4;
constexpr size_t kMdFusedCountOff = kMdPlanAreaBytes - 8;   // codes md_fused takes
constexpr size_t kMdMidCountOff = kMdPlanAreaBytes - 12;
constexpr size_t kMdPlanBytesPerCode = 24;
static_assert(sizeof(MdPlan) <= 500, "MdPlan and the three counts fit the slot");

// ---- scratch of the allocation-free exact pass (gt_md_batch_dev_noalloc) --
// One file-scope device array per GPU, held until the process ends: the plan
// area, the per-code state words, the bit rows of at most kNaMaxCodes codes
// with kNaMaxIdx indices in total, and the list count (19.5 MB).  Host code
// never passes the array's name to a kernel: it uses the address
// hipGetSymbolAddress returns for the current device (g_scratch_ptr), taken
// once per device by gt_open.
constexpr int32_t kNaMaxCodes = 16384;
constexpr int64_t kNaMaxIdx = 65536;
constexpr size_t kNaPlanOff = 0;
constexpr size_t kNaBestOff =
    kNaPlanOff + kMdPlanAreaBytes + (size_t)kNaMaxCodes * kMdPlanBytesPerCode;
constexpr size_t kNaRowsOff = kNaBestOff + (size_t)kNaMaxCodes * 8;
constexpr size_t kNaCountOff = kNaRowsOff + (size_t)kNaMaxIdx * kRowsPerIdx * kRowWords * 4;
constexpr size_t kNaBytes = kNaCountOff + 256;
static_assert(kNaBestOff % 256 == 0 && kNaRowsOff % 256 == 0, "aligned scratch regions");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

// Device address of `scratch` on `device` (current), looked up once per device.
static int scratch_address(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxDevices) return fail("device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    GT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = reinterpret_cast<uint8_t*>(p);
  }
  *out = g_scratch_ptr[device];
  return 0;
}

int gt::set_error(const char* msg) {
  g_err = msg;
  return -1;
}

struct gt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  int cu_count = 0;
  int clock_khz = 0;
  int lds_per_block = 0;
  std::string name;
  int md_grid[2] = {0, 0};
  uint8_t* scratch_dev = nullptr;   // this device's `scratch` array
  MdParams params{};
  int fused_k = kFusedMaxK;         // codes of k <= fused_k: one wave per code
  bool params_dirty = true;

  DevBuf mplanes, params_buf, ctrl;
  // Error slot the kernels of the current call write (ctrl[2] of this base):
  // word 0 for device-pointer calls, asynchronous searches and graphs
  // (collected by gt_sync / gt_evolve_sync), word kCtrlHostSlot for
  // host-pointer calls, which check their own slot before they return, so
  // an uncollected error of an earlier *_dev call is never reported by an
  // unrelated synchronous call.
  uint32_t* ctrl_p = nullptr;
  DevBuf idx, offs, tmax, rows, plan, best, out, zmax, stopf;  // min distance
  DevBuf mdcnt;   // count of the list md_fused leaves (zeroed at open)
  // Per-code walk record (gt_set_md_walk_record): on for the exact calls
  // that follow; walk_n = batch size of the last exact call, -1 when that
  // call was not recorded.
  bool walk_on = false;
  int32_t walk_n = -1;
  DevBuf walkbuf;
  // One-code call: its device state {minimum, waves done} ({64, 0} between
  // launches), the pinned host word the kernel writes (tag, result) to, and
  // the last tag.
  DevBuf sbuf;
  uint64_t* spin = nullptr;
  uint32_t* ctrl_host = nullptr;   // pinned: the control slot read back by check_ctrl
  uint32_t single_tag = 0;
  DevBuf masks, klo, khi, canon, stab;                        // canonical
  DevBuf children, bflag, bsum, bidx, bcnt, bmasks, brows, boffs;  // BFS level

  bool prof = false;
  std::map<std::string, ProfRec> prof_rec;
  // Stream of the last call that used the context's buffers, and an event
  // for handing them over to a call on another stream.
  hipStream_t last_s = nullptr;
  hipEvent_t handover = nullptr;
  BufState bstate;
  hipStream_t capture_stream = nullptr;

  std::vector<DevBuf*> all_bufs() {
    return {&mplanes, &params_buf, &ctrl, &idx, &offs, &tmax, &rows, &plan, &best, &mdcnt,
            &out, &zmax, &stopf, &sbuf, &masks, &klo, &khi, &canon, &stab,
            &children, &bflag, &bsum, &bidx, &bcnt, &bmasks, &brows, &boffs, &walkbuf};
  }

  int activate() {
    GT_HIP(hipSetDevice(device));
    return 0;
  }
  hipStream_t pick(void* s) const { return s ? reinterpret_cast<hipStream_t>(s) : stream; }

  // Calls on one context share its buffers (rows, plan, minima, canonical
  // scratch, ...), so a call on stream s must not start before the previous
  // call's work on another stream has finished.  A call on a caller's stream
  // records the handover event there before it returns (the caller may
  // destroy that stream afterwards); the context stream's work is recorded
  // when the next call moves to another stream.  The next call on another
  // stream waits on the event, and so does every call on a caller's stream
  // (a destroyed stream's handle may come back for a new one).  Nothing is
  // recorded while calls stay on the context stream, and nothing during
  // graph capture (capture runs on the context stream and is ordered when it
  // begins).
  int order_stream(hipStream_t s) {
    if (bstate.frozen) return 0;
    if (last_s && !(last_s == stream && s == stream)) {
      if (last_s == stream) GT_HIP(hipEventRecord(handover, stream));
      GT_HIP(hipStreamWaitEvent(s, handover, 0));
    }
    last_s = s;
    return 0;
  }
  int release_stream(hipStream_t s) {
    if (bstate.frozen || s == stream) return 0;
    GT_HIP(hipEventRecord(handover, s));
    return 0;
  }

  int upload_params(hipStream_t s) {
    if (!params_dirty) return 0;
    if (bstate.frozen) return fail("launch parameters changed during graph capture");
    GT_HIP(hipMemcpyAsync(params_buf.p, &params, sizeof(MdParams), hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
    params_dirty = false;
    return 0;
  }

  int prof_begin(const char* k, hipStream_t s, hipEvent_t* e0) {
    if (!prof || bstate.frozen) return 0;
    GT_HIP(hipEventCreate(e0));
    GT_HIP(hipEventRecord(*e0, s));
    (void)k;
    return 0;
  }
  int prof_end(const char* k, hipStream_t s, hipEvent_t e0) {
    if (!prof || bstate.frozen) return 0;
    hipEvent_t e1;
    GT_HIP(hipEventCreate(&e1));
    GT_HIP(hipEventRecord(e1, s));
    prof_rec[k].ev.emplace_back(e0, e1);
    return 0;
  }

  // The control slot is read back into pinned memory (a pageable
  // destination costs a staged copy, several times longer).
  int check_ctrl(hipStream_t s) {
    uint32_t* h = ctrl_host;
    GT_HIP(hipMemcpyAsync(h, ctrl_p, kCtrlSlotWords * 4, hipMemcpyDeviceToHost, s));
    GT_HIP(hipStreamSynchronize(s));
    if (h[2] != 0) {
      const uint32_t zero = 0;
      GT_HIP(hipMemcpyAsync(ctrl_p + 2, &zero, 4, hipMemcpyHostToDevice, s));
      GT_HIP(hipStreamSynchronize(s));
      if (h[2] & kErrBadIndex) return fail("input error: index outside 0..48 or repeated in a set");
      return fail("input error: set size outside the supported range, or too much work for one call");
    }
    return 0;
  }

  // ---- min distance on device pointers (stream-ordered, no host sync) ----
  // kmax: a bound on the sets' sizes known to the caller (larger sets are
  // input errors); the list search is enqueued only when kmax > fused_k.
  int md_enqueue(const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets, int32_t n,
                 const int32_t* d_tmax, int32_t tmax_s, bool thresholded,
                 int64_t sample_count, uint64_t seed, int32_t* d_out, hipStream_t s,
                 uint32_t set_period = 0, const uint32_t* d_gen = nullptr,
                 int kmax = kMaxKExact) {
    if (n <= 0) return 0;
    GT_OK(upload_params(s));
    if (sample_count > 0) {
      GT_OK(rows.ensure((size_t)std::max<int64_t>(n_idx, 1) * 3 * kRowWords * 4));
      GT_HIP(launch_md_rows(d_idx, n_idx, mplanes.as<uint32_t>(), rows.as<uint32_t>(),
                            ctrl_p, s));
      GT_OK(zmax.ensure((size_t)n * 4));
      GT_OK(stopf.ensure((size_t)n * 4));
      hipEvent_t e0 = nullptr;
      GT_OK(prof_begin("md_sampled", s, &e0));
      GT_HIP(launch_md_sampled(rows.as<uint32_t>(), d_offsets, n, d_tmax, tmax_s, sample_count,
                               seed, set_period, d_gen, zmax.as<int32_t>(), stopf.as<uint32_t>(),
                               d_out, ctrl_p, s));
      GT_OK(prof_end("md_sampled", s, e0));
      return 0;
    }
    GT_OK(rows.ensure((size_t)std::max<int64_t>(n_idx, 1) * kRowsPerIdx * kRowWords * 4));
    GT_OK(plan.ensure(kMdPlanAreaBytes + (size_t)n * kMdPlanBytesPerCode));
    GT_OK(best.ensure((size_t)n * 8));
    uint64_t* d_walked = nullptr;
    walk_n = -1;
    if (walk_on) {
      GT_OK(walkbuf.ensure((size_t)n * 8));
      d_walked = walkbuf.as<uint64_t>();
      walk_n = n;
    }
    return md_exact_launch(d_idx, n_idx, d_offsets, n, d_tmax, tmax_s, thresholded, d_out,
                           rows.as<uint32_t>(), plan.as<uint8_t>(), best.as<uint64_t>(),
                           mdcnt.as<uint32_t>(), d_walked, s, true, kmax);
  }

  // Exact or thresholded search with caller-given scratch: rows
  // (kRowsPerIdx * n_idx rows), the plan area (kMdPlanAreaBytes + 24 n bytes:
  // claim counters, MdPlan and the list counts, order, row offsets, aux
  // words, the three lists), the per-code state (n 64-bit words) and the
  // count of md_fused's list (one word, 0 between calls).  md_fused searches
  // the codes of k <= fused_k and lists the others; when kmax allows larger
  // codes the list search follows: md_prepl and md_main at the codes'
  // levels, then, for the codes the first pass could not certify, one level
  // lower (only when some k in (fused_k, kmax] has a middle level) and at
  // second_level(k).  Every word a kernel reads is written before by an
  // earlier kernel of the call, so no clearing is needed.
  int md_exact_launch(const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets, int32_t n,
                      const int32_t* d_tmax, int32_t tmax_s, bool thresholded, int32_t* d_out,
                      uint32_t* d_rows, uint8_t* d_plan_area, uint64_t* d_state,
                      uint32_t* d_big_count, uint64_t* d_walked, hipStream_t s, bool profile,
                      int kmax) {
    uint32_t* d_claims = reinterpret_cast<uint32_t*>(d_plan_area);
    MdPlan* d_plan = reinterpret_cast<MdPlan*>(d_plan_area + kMdClaimBytes);
    uint32_t* d_order = reinterpret_cast<uint32_t*>(d_plan_area + kMdPlanAreaBytes);
    uint32_t* d_roff = d_order + n;
    const MdAux ax{d_roff + n,
                   reinterpret_cast<uint32_t*>(d_plan_area + kMdFbCountOff),
                   d_roff + 2 * (size_t)n,
                   d_big_count,
                   d_roff + 3 * (size_t)n,
                   reinterpret_cast<uint32_t*>(d_plan_area + kMdMidCountOff),
                   d_roff + 4 * (size_t)n,
                   d_walked};
    const MdParams* dp = params_buf.as<MdParams>();
    const int grid = md_grid[thresholded ? 1 : 0];
    const bool timed = profile && prof && !bstate.frozen;
    hipEvent_t ev[8] = {};
    if (timed)
      for (hipEvent_t& e : ev) GT_HIP(hipEventCreate(&e));
    GT_HIP(launch_md_fused(d_idx, n_idx, d_offsets, n, mplanes.as<uint32_t>(), dp, d_tmax, tmax_s,
                           d_out, ax, d_order,
                           reinterpret_cast<uint32_t*>(d_plan_area + kMdFusedCountOff), fused_k,
                           kmax, ctrl_p, s, ev[0], ev[1]));
    if (timed) prof_rec["md_fused"].ev.emplace_back(ev[0], ev[1]);
    bool mid = false;
    for (int k = fused_k + 1; k <= std::min(kmax, kMaxKExact); ++k)
      mid = mid || mid_level(params.level, k) > 0;
    if (timed)   // events of the launches that do not follow
      for (int i = 2; i < 8; ++i)
        if (kmax <= fused_k || (!mid && (i == 4 || i == 5))) GT_HIP(hipEventDestroy(ev[i]));
    if (kmax <= fused_k) return 0;
    static const char* const kPassKernel[kPasses] = {"md_main", "md_main_mid", "md_main2"};
    for (int pass = kPassFirst; pass < kPasses; ++pass) {
      if (pass == kPassMid && !mid) continue;
      GT_HIP(launch_md_prepl(pass, d_idx, n_idx, mplanes.as<uint32_t>(), d_rows, d_offsets, n,
                             dp, d_plan, d_order, d_roff, d_claims, d_state, d_out, d_tmax,
                             tmax_s, ax, ctrl_p, s));
      hipEvent_t e0 = ev[2 + 2 * pass], e1 = ev[3 + 2 * pass];
      GT_HIP(launch_md_main(thresholded, grid, d_rows, d_plan, d_order, d_roff, d_claims, d_state,
                            d_tmax, tmax_s, d_out, dp, ax, pass, s, e0, e1));
      if (timed) prof_rec[kPassKernel[pass]].ev.emplace_back(e0, e1);
    }
    return 0;
  }
};

// Work of one call on stream s: ordered after the context's previous work
// (order_stream) and handed over when the call ends (release_stream).
struct StreamUse {
  gt_ctx* c;
  hipStream_t s;
  int rc;
  StreamUse(gt_ctx* ctx, hipStream_t st) : c(ctx), s(st) { rc = c->order_stream(s); }
  ~StreamUse() { (void)c->release_stream(s); }
};

// Scope of a host-pointer call: its kernels flag errors in the host slot.
struct HostCall {
  gt_ctx* c;
  explicit HostCall(gt_ctx* ctx) : c(ctx) { c->ctrl_p = c->ctrl.as<uint32_t>() + kCtrlHostSlot; }
  ~HostCall() { c->ctrl_p = c->ctrl.as<uint32_t>(); }
};

extern "C" {

const char* gt_last_error(void) { return g_err.c_str(); }

int gt_device_count(int32_t* n) {
  if (!n) return fail("null output pointer");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return 0;
}

int gt_open(int32_t device, gt_ctx** out) {
  if (!out) return fail("gt_open: null output pointer");
  *out = nullptr;
  int ndev = 0;
  GT_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail("gt_open: device %d not present (%d devices)", device, ndev);
  gt_ctx* c = new gt_ctx();
  c->device = device;
  auto bail = [&](int r) {
    if (c->ctrl_host) (void)hipHostFree(c->ctrl_host);
    delete c;
    return r;
  };
  if (c->activate()) return bail(-1);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bail(fail("hipGetDeviceProperties failed"));
  c->cu_count = prop.multiProcessorCount;
  c->clock_khz = prop.clockRate;
  c->lds_per_block = (int)prop.sharedMemPerBlock;
  c->name = prop.gcnArchName;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail("hipStreamCreate failed"));
  if (hipEventCreateWithFlags(&c->handover, hipEventDisableTiming) != hipSuccess)
    return bail(fail("hipEventCreate failed"));
  for (int k = 0; k <= kMaxKExact; ++k) c->params.level[k] = (uint8_t)default_level(k);
  for (int k = 0; k <= kMaxKExact; ++k) md_params_set(c->params, k, default_L(k), 2);
  c->md_grid[0] = c->cu_count * md_main_blocks_per_cu(false);
  c->md_grid[1] = c->cu_count * md_main_blocks_per_cu(true);
  if (scratch_address(device, &c->scratch_dev)) return bail(-1);
  uint32_t mp[kNpts * kRowWords];
  build_mplanes(mp);
  if (c->mplanes.ensure(sizeof(mp)) || c->params_buf.ensure(sizeof(MdParams)) ||
      c->ctrl.ensure(kCtrlWords * 4) || c->mdcnt.ensure(16))
    return bail(-1);
  if (hipHostMalloc((void**)&c->ctrl_host, kCtrlSlotWords * 4, hipHostMallocDefault) != hipSuccess)
    return bail(fail("gt_open: hipHostMalloc of the control readback failed"));
  uint32_t zeros[kCtrlWords] = {0};
  if (hipMemcpy(c->mplanes.p, mp, sizeof(mp), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->ctrl.p, zeros, sizeof(zeros), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->mdcnt.p, zeros, 16, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->params_buf.p, &c->params, sizeof(MdParams), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail("gt_open: upload of constant tables failed"));
  c->params_dirty = false;
  c->ctrl_p = c->ctrl.as<uint32_t>();
  for (DevBuf* b : c->all_bufs()) b->st = &c->bstate;
  *out = c;
  return 0;
}

int gt_close(gt_ctx* c) {
  if (!c) return 0;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->prof_rec)
      for (auto& pr : kv.second.ev) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (DevBuf* b : c->all_bufs()) b->release();
    if (c->spin) (void)hipHostFree(c->spin);
    if (c->ctrl_host) (void)hipHostFree(c->ctrl_host);
    if (c->handover) (void)hipEventDestroy(c->handover);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
  return 0;
}

int gt_device_info(gt_ctx* c, int32_t* cu_count, int32_t* clock_khz, int32_t* lds_per_block,
                   char* name, int32_t name_len) {
  if (!c) return fail("null context");
  if (cu_count) *cu_count = c->cu_count;
  if (clock_khz) *clock_khz = c->clock_khz;
  if (lds_per_block) *lds_per_block = c->lds_per_block;
  if (name && name_len > 0) {
    std::strncpy(name, c->name.c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  return 0;
}

int gt_set_md_params(gt_ctx* c, int32_t k, int32_t L, int32_t check_log2) {
  if (!c) return fail("null context");
  if (k < 0 || k > kMaxKExact) return fail("k=%d outside 0..%d", k, kMaxKExact);
  if (L < kU || L > 30) return fail("L=%d outside %d..30", L, kU);
  if (check_log2 < 0 || check_log2 > 16) return fail("check_log2=%d outside 0..16", check_log2);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->bstate.frozen) return fail("launch parameters changed during graph capture");
  md_params_set(c->params, k, L, check_log2);
  c->params_dirty = true;
  // Uploaded now (ordered on the context stream), so no later call has to.
  GT_OK(c->activate());
  return c->upload_params(c->stream);
}

int gt_get_md_params(gt_ctx* c, int32_t k, int32_t* L, int32_t* check_log2) {
  if (!c) return fail("null context");
  if (k < 0 || k > kMaxKExact) return fail("k=%d outside 0..%d", k, kMaxKExact);
  if (L) *L = c->params.L[k];
  if (check_log2) *check_log2 = c->params.check_log2[k];
  return 0;
}

int gt_set_md_level(gt_ctx* c, int32_t k, int32_t level) {
  if (!c) return fail("null context");
  if (k < 3 || k > kMaxKExact) return fail("k=%d outside 3..%d (k <= 2 walks at level 1)", k, kMaxKExact);
  if (level < 1 || level > std::min(k - 1, kMaxLevel))
    return fail("level=%d outside 1..%d for k=%d", level, std::min(k - 1, kMaxLevel), k);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->bstate.frozen) return fail("launch parameters changed during graph capture");
  MdParams p = c->params;
  md_level_set(p, k, level);
  for (int q = 3; q < kMaxKExact; ++q)
    if (p.level[q + 1] > p.level[q] + 1)
      return fail("level %d at k=%d would follow level %d at k=%d: the first-pass walk dimension "
                  "k - level must not decrease with k", (int)p.level[q + 1], q + 1,
                  (int)p.level[q], q);
  c->params = p;
  c->params_dirty = true;
  GT_OK(c->activate());
  return c->upload_params(c->stream);
}

int gt_md_level_info(int32_t m, int32_t s, uint32_t* rep_mask, int32_t* nsub, int32_t* cert) {
  if (m < 1 || m > kMaxLevel) return fail("level %d outside 1..%d", m, kMaxLevel);
  if (s < 0 || s >= level_nsub(m)) return fail("level %d has %d representatives", m, level_nsub(m));
  if (rep_mask) *rep_mask = level_rep(m, s);
  if (nsub) *nsub = level_nsub(m);
  if (cert) *cert = level_cert(m);
  return 0;
}

int gt_set_md_fused_max_k(gt_ctx* c, int32_t k) {
  if (!c) return fail("null context");
  if (k < 1 || k > kFusedMaxK) return fail("fused bound %d outside 1..%d", k, kFusedMaxK);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->bstate.frozen) return fail("launch parameters changed during graph capture");
  c->fused_k = k;
  return 0;
}

int gt_get_md_fused_max_k(gt_ctx* c, int32_t* k) {
  if (!c) return fail("null context");
  if (k) *k = c->fused_k;
  return 0;
}

int gt_set_md_walk_record(gt_ctx* c, int32_t on) {
  if (!c) return fail("null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->bstate.frozen) return fail("walk record switched during graph capture");
  c->walk_on = on != 0;
  return 0;
}

int gt_get_md_walk_record(gt_ctx* c, uint64_t* out, int32_t n) {
  if (!c) return fail("null context");
  if (n < 0 || (n > 0 && !out)) return fail("null out");
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->walk_n < 0) return fail("the last exact call on this context was not recorded");
  if (n != c->walk_n) return fail("the last recorded call had %d codes, not %d", c->walk_n, n);
  if (n == 0) return 0;
  GT_OK(c->activate());
  // The recorded call may have run on a caller's stream that no longer
  // exists: wait for the device, then copy (a diagnostic call, not a hot
  // path).
  GT_HIP(hipDeviceSynchronize());
  GT_HIP(hipMemcpy(out, c->walkbuf.p, (size_t)n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int gt_get_md_level(gt_ctx* c, int32_t k, int32_t* level) {
  if (!c) return fail("null context");
  if (k < 0 || k > kMaxKExact) return fail("k=%d outside 0..%d", k, kMaxKExact);
  if (level) *level = code_level(c->params.level, k, false);
  return 0;
}

int gt_md_batch(gt_ctx* c, const int32_t* idx, const int32_t* offsets, int32_t n,
                const int32_t* tmax_per_set, int32_t tmax, int64_t sample_count,
                uint64_t sample_seed, int32_t* out) {
  if (!c) return fail("null context");
  if (n < 0) return fail("negative batch size");
  if (n == 0) return 0;
  if (!offsets || !out) return fail("null offsets/out");
  const int64_t n_idx = offsets[n];
  if (n_idx > 0 && !idx) return fail("null idx");
  if (offsets[0] != 0) return fail("offsets[0] must be 0");
  bool thresholded = false;
  uint64_t chunks = 0;
  int kmax = 0;
  for (int32_t s = 0; s < n; ++s) {
    const int k = offsets[s + 1] - offsets[s];
    if (k < 0) return fail("offsets must be non-decreasing");
    kmax = std::max(kmax, k);
    if (sample_count > 0 ? k > kMaxK : k > kMaxKExact)
      return fail(sample_count > 0 ? "sampled evaluation supports k <= 49 (set %d has k=%d)"
                                   : "exact evaluation supports k <= 21 (set %d has k=%d)", s, k);
    const int t = tmax_per_set ? tmax_per_set[s] : tmax;
    if (t < 48) thresholded = true;
    if (sample_count <= 0 && k > 0) chunks += code_chunks(walk_dim(k), c->params.L[k]);
  }
  if (chunks >= (1ull << 31)) return fail("exact search too large for one call (%llu chunks)", (unsigned long long)chunks);
  for (int64_t i = 0; i < n_idx; ++i)
    if (idx[i] < 0 || idx[i] >= kNpts) return fail("index %d outside 0..48", idx[i]);
  std::lock_guard<std::mutex> lk(c->mu);
  HostCall hc(c);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  hipStream_t s = c->stream;
  GT_OK(c->idx.ensure((size_t)std::max<int64_t>(n_idx, 1) * 4));
  GT_OK(c->offs.ensure((size_t)(n + 1) * 4));
  GT_OK(c->out.ensure((size_t)n * 4));
  if (tmax_per_set) GT_OK(c->tmax.ensure((size_t)n * 4));
  if (n_idx > 0) GT_HIP(hipMemcpyAsync(c->idx.p, idx, (size_t)n_idx * 4, hipMemcpyHostToDevice, s));
  GT_HIP(hipMemcpyAsync(c->offs.p, offsets, (size_t)(n + 1) * 4, hipMemcpyHostToDevice, s));
  if (tmax_per_set)
    GT_HIP(hipMemcpyAsync(c->tmax.p, tmax_per_set, (size_t)n * 4, hipMemcpyHostToDevice, s));
  GT_OK(c->md_enqueue(c->idx.as<int32_t>(), n_idx, c->offs.as<int32_t>(), n,
                      tmax_per_set ? c->tmax.as<int32_t>() : nullptr, tmax, thresholded,
                      sample_count, sample_seed, c->out.as<int32_t>(), s, 0, nullptr, kmax));
  GT_HIP(hipMemcpyAsync(out, c->out.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  GT_OK(c->check_ctrl(s));
  return 0;
}

int gt_md_batch_dev2(gt_ctx* c, const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets,
                     int32_t n, const int32_t* d_tmax, int32_t tmax, int32_t thresholded,
                     int32_t kmax, int64_t sample_count, uint64_t sample_seed, int32_t* d_out,
                     void* stream) {
  if (!c) return fail("null context");
  if (n < 0 || n_idx < 0) return fail("negative sizes");
  if (n > 0 && (!d_offsets || !d_out)) return fail("null device offsets/out");
  if (n_idx > 0 && !d_idx) return fail("null device idx");
  if (kmax < 0 || kmax > kMaxKExact) return fail("kmax=%d outside 0..%d", kmax, kMaxKExact);
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  return c->md_enqueue(d_idx, n_idx, d_offsets, n, d_tmax, tmax, thresholded != 0, sample_count,
                       sample_seed, d_out, c->pick(stream), 0, nullptr,
                       sample_count > 0 ? kMaxKExact : kmax);
}

int gt_md_batch_dev(gt_ctx* c, const int32_t* d_idx, int64_t n_idx, const int32_t* d_offsets,
                    int32_t n, const int32_t* d_tmax, int32_t tmax, int32_t thresholded,
                    int64_t sample_count, uint64_t sample_seed, int32_t* d_out, void* stream) {
  return gt_md_batch_dev2(c, d_idx, n_idx, d_offsets, n, d_tmax, tmax, thresholded, kMaxKExact,
                          sample_count, sample_seed, d_out, stream);
}

int gt_md_batch_dev_noalloc(gt_ctx* c, const int32_t* d_idx, int64_t n_idx,
                            const int32_t* d_offsets, int32_t n, const int32_t* d_tmax,
                            int32_t tmax, int32_t thresholded, int32_t* d_out, void* stream) {
  if (!c) return fail("null context");
  if (n < 0 || n_idx < 0) return fail("negative sizes");
  if (n > kNaMaxCodes || n_idx > kNaMaxIdx)
    return fail("allocation-free call: %d codes / %lld indices exceed the maximum of %d codes / "
                "%lld indices", n, (long long)n_idx, kNaMaxCodes, (long long)kNaMaxIdx);
  if (n == 0) return 0;
  if (!d_offsets || !d_out) return fail("null device offsets/out");
  if (n_idx > 0 && !d_idx) return fail("null device idx");
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->params_dirty) return fail("launch parameters not uploaded");
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  uint8_t* base = c->scratch_dev;
  c->walk_n = -1;   // not recorded (the record's buffer is allocated)
  return c->md_exact_launch(d_idx, n_idx, d_offsets, n, d_tmax, tmax, thresholded != 0, d_out,
                            reinterpret_cast<uint32_t*>(base + kNaRowsOff), base + kNaPlanOff,
                            reinterpret_cast<uint64_t*>(base + kNaBestOff),
                            reinterpret_cast<uint32_t*>(base + kNaCountOff), nullptr,
                            c->pick(stream), false, kMaxKExact);
}

int gt_md_batch_dev_noalloc_limits(int32_t* max_codes, int64_t* max_idx) {
  if (max_codes) *max_codes = kNaMaxCodes;
  if (max_idx) *max_idx = kNaMaxIdx;
  return 0;
}

constexpr int kSingleBatchK = 12;   // one-code calls of at least this k use the batch path

int gt_md_single(gt_ctx* c, const int32_t* idx, int32_t k, int32_t tmax, int32_t* out) {
  if (!c) return fail("null context");
  if (!out) return fail("null out");
  if (k <= 0) { *out = 0; return 0; }
  if (k > kMaxKExact) return fail("exact evaluation supports k <= 21 (k=%d)", k);
  for (int i = 0; i < k; ++i)
    if (idx[i] < 0 || idx[i] >= kNpts) return fail("index %d outside 0..48", idx[i]);
  const int thr = 48 - tmax;
  const int w = walk_dim(k);
  if (w == 0) {   // k = 1: a monomial has weight 49
    *out = (49 <= thr) ? tmax + 1 : 0;
    return 0;
  }
  // From k = kSingleBatchK on, the level-1 walk of one launch (8^(k-1) / 7
  // codewords: 0.33 ms at k = 12, 21 ms at k = 14) costs more than the batch
  // path's passes over anchored subcodes (level 4, then 3, then 2), launches
  // included: the code goes through gt_md_batch as a batch of one.
  if (k >= kSingleBatchK) {
    const int32_t offsets[2] = {0, k};
    return gt_md_batch(c, idx, offsets, 1, nullptr, tmax, 0, 0, out);
  }
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  hipStream_t s = c->stream;
  // Host-built bit rows of the subcode vanishing at (1,1) (rows 3i + b =
  // x^b * (g_i + g_{k-1}), i < k - 1), passed in the kernel arguments.
  uint32_t mp[kNpts * kRowWords];
  build_mplanes(mp);
  SingleRows rows{};
  for (int i = 0; i < w; ++i) {
    uint32_t v[6];
    for (int q = 0; q < 6; ++q) v[q] = mp[idx[i] * kRowWords + q] ^ mp[idx[k - 1] * kRowWords + q];
    for (int b = 0; b < 3; ++b) {
      for (int q = 0; q < 6; ++q) rows.r[(3 * i + b) * kRowWords + q] = v[q];
      mul_x(v);
    }
  }
  if (!c->spin) {
    GT_HIP(hipHostMalloc((void**)&c->spin, 256, hipHostMallocCoherent | hipHostMallocMapped));
    __atomic_store_n(c->spin, 0ull, __ATOMIC_RELEASE);   // no tag yet (tags start at 1)
    GT_OK(c->sbuf.ensure(8));
    const uint32_t init[2] = {64u, 0u};
    GT_HIP(hipMemcpyAsync(c->sbuf.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    GT_HIP(hipStreamSynchronize(s));
  }
  // One-code launches want many waves: about 2^11 chunks for the largest piece.
  const int L = std::max<int>(kU, std::min<int>(c->params.L[k], 3 * (w - 1) - kLaneBits - 11));
  hipEvent_t e0 = nullptr;
  GT_OK(c->prof_begin("md_single", s, &e0));
  const uint32_t tag = ++c->single_tag;
  GT_HIP(launch_md_single(tmax < 48, rows, w, L, c->params.check_log2[k], thr, tmax, tag,
                          c->sbuf.as<uint32_t>(), c->spin, s));
  GT_OK(c->prof_end("md_single", s, e0));
  // A short launch is awaited by polling the pinned word (a stream
  // synchronisation costs more than most one-code launches run); after
  // 200 us the call blocks on the stream instead.
  uint64_t word = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    word = __atomic_load_n(c->spin, __ATOMIC_ACQUIRE);
    if ((uint32_t)(word >> 32) == tag) break;
    if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) {
      GT_HIP(hipStreamSynchronize(s));
      word = __atomic_load_n(c->spin, __ATOMIC_ACQUIRE);
      if ((uint32_t)(word >> 32) != tag) return fail("one-code launch left no result");
      break;
    }
  }
  *out = (int32_t)(uint32_t)word;
  return 0;
}

// Canonical forms of n device masks in input order (the kernel validates
// them: a bad mask is flagged and its outputs are unspecified).
static int canon_enqueue(gt_ctx* c, const uint64_t* d_masks, int32_t n, uint64_t* d_lo,
                         uint64_t* d_hi, uint64_t* d_canon, int32_t* d_stab, hipStream_t s) {
  if (n <= 0) return 0;
  hipEvent_t e0 = nullptr;
  GT_OK(c->prof_begin("canon", s, &e0));
  GT_HIP(launch_canon(d_masks, n, d_lo, d_hi, d_canon, d_stab, c->ctrl_p, s));
  GT_OK(c->prof_end("canon", s, e0));
  return 0;
}

static int canon_host(gt_ctx* c, const uint64_t* masks, const int32_t* sets, int32_t n, int32_t k,
                      uint64_t* key_lo, uint64_t* key_hi, uint64_t* canon_mask, int32_t* stab) {
  std::lock_guard<std::mutex> lk(c->mu);
  HostCall hc(c);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  hipStream_t s = c->stream;
  GT_OK(c->masks.ensure((size_t)n * 8));
  GT_OK(c->klo.ensure((size_t)n * 8));
  GT_OK(c->khi.ensure((size_t)n * 8));
  GT_OK(c->canon.ensure((size_t)n * 8));
  GT_OK(c->stab.ensure((size_t)n * 4));
  if (sets) {
    GT_OK(c->idx.ensure((size_t)n * k * 4));
    GT_HIP(hipMemcpyAsync(c->idx.p, sets, (size_t)n * k * 4, hipMemcpyHostToDevice, s));
    GT_HIP(launch_sets_to_masks(c->idx.as<int32_t>(), n, k, c->masks.as<uint64_t>(),
                                c->ctrl_p, s));
  } else {
    GT_HIP(hipMemcpyAsync(c->masks.p, masks, (size_t)n * 8, hipMemcpyHostToDevice, s));
  }
  GT_OK(canon_enqueue(c, c->masks.as<uint64_t>(), n, c->klo.as<uint64_t>(), c->khi.as<uint64_t>(),
                      c->canon.as<uint64_t>(), c->stab.as<int32_t>(), s));
  if (key_lo) GT_HIP(hipMemcpyAsync(key_lo, c->klo.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (key_hi) GT_HIP(hipMemcpyAsync(key_hi, c->khi.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (canon_mask) GT_HIP(hipMemcpyAsync(canon_mask, c->canon.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (stab) GT_HIP(hipMemcpyAsync(stab, c->stab.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  GT_OK(c->check_ctrl(s));
  return 0;
}

int gt_canon_batch(gt_ctx* c, const int32_t* sets, int32_t n, int32_t k, uint64_t* key_lo,
                   uint64_t* key_hi, uint64_t* canon_mask, int32_t* stab) {
  if (!c) return fail("null context");
  if (k < 1 || k > kMaxKCanon) return fail("canonical form supports 1 <= k <= 16 (k=%d)", k);
  if (n < 0) return fail("negative batch size");
  if (n == 0) return 0;
  if (!sets) return fail("null sets");
  return canon_host(c, nullptr, sets, n, k, key_lo, key_hi, canon_mask, stab);
}

int gt_canon_masks(gt_ctx* c, const uint64_t* masks, int32_t n, uint64_t* key_lo, uint64_t* key_hi,
                   uint64_t* canon_mask, int32_t* stab) {
  if (!c) return fail("null context");
  if (n < 0) return fail("negative batch size");
  if (n == 0) return 0;
  if (!masks) return fail("null masks");
  return canon_host(c, masks, nullptr, n, 0, key_lo, key_hi, canon_mask, stab);
}

int gt_canon_masks_dev(gt_ctx* c, const uint64_t* d_masks, int32_t n, uint64_t* d_key_lo,
                       uint64_t* d_key_hi, uint64_t* d_canon_mask, int32_t* d_stab, void* stream) {
  if (!c) return fail("null context");
  if (n < 0) return fail("negative batch size");
  if (n > 0 && !d_masks) return fail("null device masks");
  if (!d_key_lo || !d_key_hi || !d_canon_mask || !d_stab) return fail("null device output");
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  return canon_enqueue(c, d_masks, n, d_key_lo, d_key_hi, d_canon_mask, d_stab, c->pick(stream));
}

struct gt_graph {
  gt_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t gen = 0;
};

int gt_capture_begin(gt_ctx* c) {
  if (!c) return fail("null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->bstate.frozen) return fail("capture already in progress");
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  GT_OK(c->upload_params(c->stream));
  GT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->bstate.frozen = true;
  return 0;
}

int gt_capture_end(gt_ctx* c, gt_graph** out) {
  if (!c || !out) return fail("null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bstate.frozen) return fail("no capture in progress");
  c->bstate.frozen = false;
  *out = nullptr;
  hipGraph_t g = nullptr;
  GT_HIP(hipStreamEndCapture(c->stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail("hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  gt_graph* gr = new gt_graph();
  gr->ctx = c;
  gr->graph = g;
  gr->exec = ex;
  gr->gen = c->bstate.gen;
  *out = gr;
  return 0;
}

int gt_graph_launch(gt_graph* g, void* stream) {
  if (!g || !g->ctx) return fail("null graph");
  gt_ctx* c = g->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (g->gen != c->bstate.gen)
    return fail("context buffers were reallocated after this graph was captured; recapture");
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  GT_HIP(hipGraphLaunch(g->exec, c->pick(stream)));
  return 0;
}

int gt_graph_destroy(gt_graph* g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return 0;
}

// ------------------------------------------------------- hash set / BFS --

struct gt_hashset {
  gt_ctx* ctx = nullptr;
  uint64_t cap = 0;
  uint64_t* table = nullptr;
  uint64_t* tags = nullptr;
  uint32_t* count = nullptr;
  uint64_t call = 0;
};

int gt_hashset_create(gt_ctx* c, int32_t log2_cap, gt_hashset** out) {
  if (!c || !out) return fail("null argument");
  if (log2_cap < 4 || log2_cap > 30) return fail("log2 capacity must be in 4..30 (got %d)", log2_cap);
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  gt_hashset* h = new gt_hashset();
  h->ctx = c;
  h->cap = 1ull << log2_cap;
  if (hipMalloc(&h->table, h->cap * 8) != hipSuccess || hipMalloc(&h->tags, h->cap * 8) != hipSuccess ||
      hipMalloc(&h->count, 256) != hipSuccess) {
    if (h->table) (void)hipFree(h->table);
    if (h->tags) (void)hipFree(h->tags);
    delete h;
    return fail("hipMalloc of a hash set with 2^%d slots failed", log2_cap);
  }
  if (launch_hs_clear(h->table, h->tags, h->cap, h->count, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    (void)hipFree(h->table);
    (void)hipFree(h->tags);
    (void)hipFree(h->count);
    delete h;
    return fail("hash set initialisation failed");
  }
  *out = h;
  return 0;
}

int gt_hashset_destroy(gt_hashset* h) {
  if (!h) return 0;
  {
    std::lock_guard<std::mutex> lk(h->ctx->mu);
    (void)h->ctx->activate();
    (void)hipStreamSynchronize(h->ctx->stream);
    (void)hipFree(h->table);
    (void)hipFree(h->tags);
    (void)hipFree(h->count);
  }
  delete h;
  return 0;
}

int gt_hashset_clear(gt_hashset* h, void* stream) {
  if (!h) return fail("null hash set");
  gt_ctx* c = h->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_HIP(launch_hs_clear(h->table, h->tags, h->cap, h->count, c->pick(stream)));
  h->call = 0;
  return 0;
}

int gt_hashset_count(gt_hashset* h, int64_t* n) {
  if (!h || !n) return fail("null argument");
  gt_ctx* c = h->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  uint32_t v = 0;
  GT_HIP(hipMemcpyAsync(&v, h->count, 4, hipMemcpyDeviceToHost, c->stream));
  GT_HIP(hipStreamSynchronize(c->stream));
  *n = v;
  return 0;
}

static int dedup_enqueue(gt_hashset* h, const uint64_t* d_keys, int32_t n, uint32_t* d_flag,
                         hipStream_t s) {
  if ((uint64_t)n >= (1ull << 32)) return fail("batch too large");
  ++h->call;
  GT_HIP(launch_hs_dedup(d_keys, n, h->table, h->tags, h->cap, h->call, h->count, d_flag,
                         h->ctx->ctrl_p, s));
  return 0;
}

int gt_dedup_dev(gt_hashset* h, const uint64_t* d_keys, int32_t n, uint32_t* d_flag, void* stream) {
  if (!h || (!d_keys && n > 0) || (!d_flag && n > 0) || n < 0) return fail("bad argument");
  if (n == 0) return 0;
  gt_ctx* c = h->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  return dedup_enqueue(h, d_keys, n, d_flag, c->pick(stream));
}

int gt_dedup(gt_hashset* h, const uint64_t* keys, int32_t n, uint32_t* flag) {
  if (!h || !keys || !flag || n < 0) return fail("bad argument");
  if (n == 0) return 0;
  gt_ctx* c = h->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  HostCall hc(c);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  hipStream_t s = c->stream;
  GT_OK(c->bmasks.ensure((size_t)n * 8));
  GT_OK(c->bflag.ensure((size_t)n * 4));
  GT_HIP(hipMemcpyAsync(c->bmasks.p, keys, (size_t)n * 8, hipMemcpyHostToDevice, s));
  GT_OK(dedup_enqueue(h, c->bmasks.as<uint64_t>(), n, c->bflag.as<uint32_t>(), s));
  GT_HIP(hipMemcpyAsync(flag, c->bflag.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  GT_OK(c->check_ctrl(s));
  return 0;
}

int gt_compact_dev(gt_ctx* c, const uint32_t* d_flag, int32_t n, uint32_t* d_out_idx,
                   uint32_t* d_count, void* stream) {
  if (!c || n < 0 || !d_count || (n > 0 && (!d_flag || !d_out_idx))) return fail("bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  StreamUse su(c, c->pick(stream));
  GT_OK(su.rc);
  GT_OK(c->bsum.ensure((size_t)((n + 1023) / 1024 + 1) * 4));
  GT_HIP(launch_compact(d_flag, n, c->bsum.as<uint32_t>(), d_out_idx, d_count, c->pick(stream)));
  return 0;
}

// One BFS level: children of the n_par parents (masks of size kp), their
// canonical masks (canonical != 0) or the children themselves (raw BFS), the
// classes not yet in `seen` (first occurrence in child order), and their
// max_zeros at threshold tmax, exact or, with sample_count > 0, from the
// sampled stream (seed sample_seed, set ids restarting every set_period
// classes when it is nonzero).  Writes at most cap classes; *n_new is the
// full count (the call fails if it exceeds cap).
int gt_bfs_level_ex(gt_ctx* c, gt_hashset* seen, const uint64_t* d_par, int32_t n_par, int32_t kp,
                    int32_t tmax, int32_t canonical, int64_t sample_count, uint64_t sample_seed,
                    int32_t set_period, uint64_t* d_out_masks, int32_t* d_out_mz, int32_t cap,
                    int32_t* n_new) {
  const int kmax = canonical ? kMaxKCanon : (sample_count > 0 ? kMaxK : kMaxKExact);
  if (!c || !seen || !n_new || n_par < 0 || kp < 0 || kp + 1 > kmax || set_period < 0)
    return fail("bad argument (parent size must be 0..%d for this mode)", kmax - 1);
  *n_new = 0;
  if (n_par == 0) return 0;
  if (!d_par || cap < 0 || (cap > 0 && (!d_out_masks || !d_out_mz))) return fail("bad argument");
  if (seen->ctx != c) return fail("hash set belongs to another context");
  const int64_t nch = (int64_t)n_par * (kNpts - kp);
  if (nch >= (1ll << 31)) return fail("level too large (%lld children)", (long long)nch);
  // One lock for the whole level: the expand/canonical/dedup scratch read by
  // the second half must not be reused by another call on this context.
  uint32_t nu = 0;
  std::lock_guard<std::mutex> lk(c->mu);
  HostCall hc(c);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  hipStream_t s = c->stream;
  const uint64_t* keys = nullptr;
  {
    const size_t n = (size_t)nch;
    GT_OK(c->children.ensure(n * 8));
    GT_OK(c->bflag.ensure(n * 4));
    GT_OK(c->bidx.ensure(n * 4));
    GT_OK(c->bcnt.ensure(64));
    GT_OK(c->bsum.ensure((n + 1023) / 1024 * 4 + 4));
    GT_HIP(launch_bfs_expand(d_par, n_par, kp, c->children.as<uint64_t>(), c->ctrl_p, s));
    keys = c->children.as<uint64_t>();
    if (canonical) {
      GT_OK(c->klo.ensure(n * 8));
      GT_OK(c->khi.ensure(n * 8));
      GT_OK(c->canon.ensure(n * 8));
      GT_OK(c->stab.ensure(n * 4));
      GT_OK(canon_enqueue(c, c->children.as<uint64_t>(), (int32_t)n, c->klo.as<uint64_t>(),
                          c->khi.as<uint64_t>(), c->canon.as<uint64_t>(), c->stab.as<int32_t>(), s));
      keys = c->canon.as<uint64_t>();
    }
    GT_OK(dedup_enqueue(seen, keys, (int32_t)n, c->bflag.as<uint32_t>(), s));
    GT_HIP(launch_compact(c->bflag.as<uint32_t>(), (int32_t)n, c->bsum.as<uint32_t>(),
                          c->bidx.as<uint32_t>(), c->bcnt.as<uint32_t>(), s));
    GT_HIP(hipMemcpyAsync(&nu, c->bcnt.p, 4, hipMemcpyDeviceToHost, s));
    GT_OK(c->check_ctrl(s));
  }
  *n_new = (int32_t)nu;
  if ((int64_t)nu > cap) return fail("%u new classes exceed the output capacity %d", nu, cap);
  if (nu == 0) return 0;
  const int k = kp + 1;
  GT_OK(c->bmasks.ensure((size_t)nu * 8));
  GT_OK(c->brows.ensure((size_t)nu * k * 4));
  GT_OK(c->boffs.ensure((size_t)(nu + 1) * 4));
  GT_HIP(launch_gather_u64(keys, c->bidx.as<uint32_t>(), c->bcnt.as<uint32_t>(), d_out_masks,
                           (int32_t)nu, s));
  GT_HIP(launch_masks_to_rows(d_out_masks, (int32_t)nu, k, c->brows.as<int32_t>(),
                              c->boffs.as<int32_t>(), c->ctrl_p, s));
  GT_OK(c->md_enqueue(c->brows.as<int32_t>(), (int64_t)nu * k, c->boffs.as<int32_t>(), (int32_t)nu,
                      nullptr, tmax, tmax < 48, sample_count > 0 ? sample_count : 0, sample_seed,
                      d_out_mz, s, (uint32_t)set_period, nullptr, k));
  GT_OK(c->check_ctrl(s));
  return 0;
}

int gt_bfs_level(gt_ctx* c, gt_hashset* seen, const uint64_t* d_par, int32_t n_par, int32_t kp,
                 int32_t tmax, uint64_t* d_out_masks, int32_t* d_out_mz, int32_t cap,
                 int32_t* n_new) {
  return gt_bfs_level_ex(c, seen, d_par, n_par, kp, tmax, 1, 0, 0, 0, d_out_masks, d_out_mz, cap,
                         n_new);
}

// --------------------------------------------------- evolutionary search --
// Every generation runs on the device: one captured graph per search,
// replayed once per generation (the population, the set of evaluated classes
// and the log stay in HBM).
}  // extern "C"

struct gt_evolve {
  gt_ctx* ctx = nullptr;
  gt_mlp* ranker = nullptr;
  MlpBufs mb{};
  hipStream_t rstream = nullptr;
  int k = 0, pop = 0, n_prop = 0, max_gen = 0;
  uint64_t seed = 0, scap = 0;
  int64_t sample_count = 0;   // > 0: sampled scores (seed + (g + 1) * 1,000,003 per generation)
  bool canonical = true;      // false: sets deduplicated as they are (no canonical forms)
  int64_t launched = 0;
  std::vector<void*> allocs;
  uint64_t *pop_m = nullptr, *props = nullptr, *plo = nullptr, *phi = nullptr, *pcm = nullptr;
  uint64_t *chosen = nullptr, *clo = nullptr, *chi = nullptr, *ccm = nullptr;
  uint64_t *skeys = nullptr, *stags = nullptr;
  uint64_t* log_cls = nullptr;
  int32_t *pop_d = nullptr, *pop_age = nullptr, *pstab = nullptr, *cstab = nullptr;
  int32_t *idx = nullptr, *offs = nullptr, *mz = nullptr;
  int32_t *log_d = nullptr, *log_best = nullptr, *log_new = nullptr;
  uint32_t *keep = nullptr, *count = nullptr, *scount = nullptr, *state = nullptr;
  float* score = nullptr;
  uint16_t* feat = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t graph_gen = 0;

  template <class T> int alloc(T** p, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n * sizeof(T), 256)) != hipSuccess)
      return fail("hipMalloc of %zu evolve bytes failed", n * sizeof(T));
    allocs.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return 0;
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  void release() {
    drop_graph();
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
  }

  // One generation, stream-ordered.
  int enqueue(hipStream_t s) {
    gt_ctx* c = ctx;
    uint32_t* ctrl = c->ctrl.as<uint32_t>();
    GT_HIP(launch_evo_propose(pop_m, pop_d, pop_age, pop, n_prop, seed, state, props, s));
    const uint64_t* pkey = canonical ? pcm : props;     // class of each proposal
    if (canonical) GT_OK(canon_enqueue(c, props, n_prop, plo, phi, pcm, pstab, s));
    GT_HIP(launch_evo_dedup(pkey, n_prop, skeys, scap, keep, count, s));
    if (ranker) {   // scores are read only when more than pop new proposals remain
      GT_HIP(launch_mlp_features(pkey, n_prop, feat, s, count, (uint32_t)pop));
      GT_HIP(launch_mlp_predict(mb, feat, n_prop, score, s, count, (uint32_t)pop));
    }
    GT_HIP(launch_evo_select(props, keep, count, ranker ? score : nullptr, pop, n_prop, k, seed,
                             state, chosen, s));
    const uint64_t* ckey = canonical ? ccm : chosen;
    if (canonical) GT_OK(canon_enqueue(c, chosen, pop, clo, chi, ccm, cstab, s));
    GT_HIP(launch_hs_insert(ckey, pop, skeys, stags, scap, 0, scount, ctrl, s));
    GT_HIP(launch_masks_to_rows(chosen, pop, k, idx, offs, ctrl, s));
    GT_OK(c->md_enqueue(idx, (int64_t)pop * k, offs, pop, nullptr, 48, false, sample_count, seed,
                        mz, s, 0, sample_count > 0 ? state : nullptr, k));
    GT_HIP(launch_evo_survive(pop_m, pop_d, pop_age, chosen, ckey, mz, count, pop, false, max_gen,
                              state, log_cls, log_d, log_best, log_new, s));
    return 0;
  }

  // (Re)capture the generation graph on the context stream; needed again
  // after a context buffer it reads was reallocated.
  int capture() {
    gt_ctx* c = ctx;
    drop_graph();
    GT_OK(c->upload_params(c->stream));
    GT_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->bstate.frozen = true;
    const int r = enqueue(c->stream);
    c->bstate.frozen = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != 0) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    GT_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      (void)hipGraphDestroy(g);
      exec = nullptr;
      return fail("hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    graph = g;
    graph_gen = c->bstate.gen;
    return 0;
  }
};

extern "C" {

int gt_evolve_create2(gt_ctx* c, int32_t k, int32_t pop, int32_t n_prop, uint64_t seed,
                      int32_t max_generations, int32_t log2_seen, gt_mlp* ranker,
                      const uint64_t* init_masks, int64_t sample_count, int32_t canonical,
                      gt_evolve** out) {
  if (!c || !out || !init_masks) return fail("null argument");
  *out = nullptr;
  if (sample_count < 0) return fail("negative sample count");
  // Canonical forms reach k = 16, the exact kernel k = 21; a 2-point swap
  // needs two points inside and two outside (k <= 47).
  const int kmax = canonical ? kMaxKCanon : (sample_count > 0 ? kNpts - 2 : kMaxKExact);
  if (k < 2 || k > kmax)
    return fail("set size must be in 2..%d for this mode (got %d)", kmax, k);
  if (pop < 1 || pop > kEvoMaxPop) return fail("population must be in 1..%d", kEvoMaxPop);
  if (n_prop < pop || n_prop > kEvoMaxProp)
    return fail("proposal count must be in pop..%d", kEvoMaxProp);
  if (max_generations < 0 || max_generations >= (1 << 24))
    return fail("max_generations must be in 0..2^24-1");
  const int64_t entries = (int64_t)pop * (max_generations + 1);
  if (log2_seen == 0) {
    log2_seen = 12;
    while ((1ll << log2_seen) < 2 * entries && log2_seen < 30) ++log2_seen;
  }
  if (log2_seen < 4 || log2_seen > 30 || (1ll << log2_seen) <= entries)
    return fail("seen-set capacity 2^%d cannot hold %lld classes", log2_seen, (long long)entries);
  for (int i = 0; i < pop; ++i)
    if (__builtin_popcountll(init_masks[i]) != k || (init_masks[i] >> kNpts) != 0)
      return fail("initial set %d is not a %d-subset of the 49 points", i, k);
  if (ranker && mlp_device(ranker) != c->device)
    return fail("the ranker lives on another device");
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  gt_evolve* e = new gt_evolve();
  e->ctx = c;
  e->ranker = ranker;
  if (ranker) {
    e->mb = *mlp_buffers(ranker);
    e->rstream = mlp_stream(ranker);
  }
  e->k = k;
  e->pop = pop;
  e->n_prop = n_prop;
  e->max_gen = max_generations;
  e->seed = seed;
  e->sample_count = sample_count;
  e->canonical = canonical != 0;
  e->scap = 1ull << log2_seen;
  const size_t P = (size_t)pop, N = (size_t)n_prop, slots = (size_t)max_generations + 1;
  auto bail = [&](int r) {
    e->release();
    delete e;
    return r;
  };
  int r = 0;
  if ((r = e->alloc(&e->pop_m, P)) || (r = e->alloc(&e->pop_d, P)) || (r = e->alloc(&e->pop_age, P)) ||
      (r = e->alloc(&e->props, N)) || (r = e->alloc(&e->plo, N)) || (r = e->alloc(&e->phi, N)) ||
      (r = e->alloc(&e->pcm, N)) || (r = e->alloc(&e->pstab, N)) ||
      (r = e->alloc(&e->keep, N)) || (r = e->alloc(&e->count, 64)) || (r = e->alloc(&e->chosen, P)) ||
      (r = e->alloc(&e->clo, P)) || (r = e->alloc(&e->chi, P)) || (r = e->alloc(&e->ccm, P)) ||
      (r = e->alloc(&e->cstab, P)) || (r = e->alloc(&e->idx, P * k)) ||
      (r = e->alloc(&e->offs, P + 1)) || (r = e->alloc(&e->mz, P)) ||
      (r = e->alloc(&e->skeys, e->scap)) ||
      (r = e->alloc(&e->stags, e->scap)) || (r = e->alloc(&e->scount, 64)) ||
      (r = e->alloc(&e->state, 64)) || (r = e->alloc(&e->log_cls, slots * P)) ||
      (r = e->alloc(&e->log_d, slots * P)) || (r = e->alloc(&e->log_best, slots)) ||
      (r = e->alloc(&e->log_new, slots)))
    return bail(r);
  if (ranker && ((r = e->alloc(&e->score, N)) || (r = e->alloc(&e->feat, N * kMlpIn))))
    return bail(r);
  // Initial population: canonical forms, seen set, exact distances, log slot
  // 0.  The same calls size the context buffers the captured graph uses.
  hipStream_t s = c->stream;
  uint32_t* ctrl = c->ctrl.as<uint32_t>();
  auto init = [&]() -> int {
    GT_HIP(hipMemcpyAsync(e->pop_m, init_masks, P * 8, hipMemcpyHostToDevice, s));
    GT_HIP(hipMemcpyAsync(e->chosen, init_masks, P * 8, hipMemcpyHostToDevice, s));
    for (size_t o = 0; o < N; o += P)
      GT_HIP(hipMemcpyAsync(e->props + o, init_masks, std::min(P, N - o) * 8,
                            hipMemcpyHostToDevice, s));
    GT_HIP(launch_hs_clear(e->skeys, e->stags, e->scap, e->scount, s));
    const uint64_t* ckey = e->canonical ? e->ccm : e->chosen;
    if (e->canonical) {
      GT_OK(canon_enqueue(c, e->props, n_prop, e->plo, e->phi, e->pcm, e->pstab, s));
      GT_OK(canon_enqueue(c, e->chosen, pop, e->clo, e->chi, e->ccm, e->cstab, s));
    }
    GT_HIP(launch_hs_insert(ckey, pop, e->skeys, e->stags, e->scap, 0, e->scount, ctrl, s));
    GT_HIP(launch_masks_to_rows(e->chosen, pop, k, e->idx, e->offs, ctrl, s));
    GT_OK(c->md_enqueue(e->idx, (int64_t)pop * k, e->offs, pop, nullptr, 48, false, sample_count,
                        seed, e->mz, s, 0, nullptr, k));
    GT_HIP(launch_evo_survive(e->pop_m, e->pop_d, e->pop_age, e->chosen, ckey, e->mz, e->count,
                              pop, true, max_generations, e->state, e->log_cls, e->log_d,
                              e->log_best, e->log_new, s));
    GT_OK(c->check_ctrl(s));
    return e->capture();
  };
  if ((r = init()) != 0) return bail(r);
  *out = e;
  return 0;
}

int gt_evolve_create(gt_ctx* c, int32_t k, int32_t pop, int32_t n_prop, uint64_t seed,
                     int32_t max_generations, int32_t log2_seen, gt_mlp* ranker,
                     const uint64_t* init_masks, gt_evolve** out) {
  return gt_evolve_create2(c, k, pop, n_prop, seed, max_generations, log2_seen, ranker, init_masks,
                           0, 1, out);
}

int gt_evolve_run(gt_evolve* e, int32_t generations) {
  if (!e) return fail("null search");
  if (generations < 0) return fail("negative generation count");
  gt_ctx* c = e->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (e->launched + generations > e->max_gen)
    return fail("%lld generations would exceed max_generations = %d",
                (long long)(e->launched + generations), e->max_gen);
  GT_OK(c->activate());
  if (generations == 0) return 0;
  GT_OK(c->order_stream(c->stream));
  if (!e->exec || e->graph_gen != c->bstate.gen) GT_OK(e->capture());
  GT_OK(c->upload_params(c->stream));
  if (e->rstream) {
    // Order the generations after ranker work already queued (training).
    hipEvent_t ev;
    GT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    GT_HIP(hipEventRecord(ev, e->rstream));
    GT_HIP(hipStreamWaitEvent(c->stream, ev, 0));
    GT_HIP(hipEventDestroy(ev));
  }
  for (int32_t i = 0; i < generations; ++i) GT_HIP(hipGraphLaunch(e->exec, c->stream));
  e->launched += generations;
  return 0;
}

int gt_evolve_train_ranker(gt_evolve* e, int32_t slot0, int32_t slot1, int32_t batch, float lr,
                           int32_t epochs) {
  if (!e) return fail("null search");
  if (!e->ranker) return fail("the search has no ranker");
  gt_ctx* c = e->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (slot0 < 0 || slot1 <= slot0 || slot1 > e->launched + 1)
    return fail("log slots [%d, %d) outside [0, %lld]", slot0, slot1, (long long)e->launched + 1);
  if (batch < 1 || epochs < 1) return fail("batch and epochs must be positive");
  GT_OK(c->activate());
  GT_OK(c->order_stream(c->stream));
  const size_t P = (size_t)e->pop;
  return mlp_train_device(e->ranker, e->log_cls + slot0 * P, e->log_d + slot0 * P,
                          (int32_t)((slot1 - slot0) * P), batch, lr, epochs, c->stream);
}

int gt_evolve_sync(gt_evolve* e, int32_t* generations_done) {
  if (!e) return fail("null search");
  gt_ctx* c = e->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_OK(c->check_ctrl(c->stream));
  if (generations_done) *generations_done = (int32_t)e->launched;
  return 0;
}

int gt_evolve_log(gt_evolve* e, int32_t slot0, int32_t slot1, uint64_t* classes, int32_t* dists,
                  int32_t* best, int32_t* n_new) {
  if (!e) return fail("null search");
  gt_ctx* c = e->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  if (slot0 < 0 || slot1 < slot0 || slot1 > e->launched + 1)
    return fail("log slots [%d, %d) outside [0, %lld]", slot0, slot1, (long long)e->launched + 1);
  GT_OK(c->activate());
  hipStream_t s = c->stream;
  const size_t P = (size_t)e->pop, ns = (size_t)(slot1 - slot0);
  if (ns == 0) return 0;
  if (classes)
    GT_HIP(hipMemcpyAsync(classes, e->log_cls + slot0 * P, ns * P * 8, hipMemcpyDeviceToHost, s));
  if (dists)
    GT_HIP(hipMemcpyAsync(dists, e->log_d + slot0 * P, ns * P * 4, hipMemcpyDeviceToHost, s));
  if (best) GT_HIP(hipMemcpyAsync(best, e->log_best + slot0, ns * 4, hipMemcpyDeviceToHost, s));
  if (n_new) GT_HIP(hipMemcpyAsync(n_new, e->log_new + slot0, ns * 4, hipMemcpyDeviceToHost, s));
  GT_HIP(hipStreamSynchronize(s));
  return 0;
}

int gt_evolve_population(gt_evolve* e, uint64_t* masks, int32_t* dists, int32_t* age,
                         int64_t* seen_classes) {
  if (!e) return fail("null search");
  gt_ctx* c = e->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  hipStream_t s = c->stream;
  const size_t P = (size_t)e->pop;
  uint32_t seen = 0;
  if (masks) GT_HIP(hipMemcpyAsync(masks, e->pop_m, P * 8, hipMemcpyDeviceToHost, s));
  if (dists) GT_HIP(hipMemcpyAsync(dists, e->pop_d, P * 4, hipMemcpyDeviceToHost, s));
  if (age) GT_HIP(hipMemcpyAsync(age, e->pop_age, P * 4, hipMemcpyDeviceToHost, s));
  GT_HIP(hipMemcpyAsync(&seen, e->scount, 4, hipMemcpyDeviceToHost, s));
  GT_HIP(hipStreamSynchronize(s));
  if (seen_classes) *seen_classes = seen;
  return 0;
}

int gt_evolve_destroy(gt_evolve* e) {
  if (!e) return 0;
  {
    std::lock_guard<std::mutex> lk(e->ctx->mu);
    (void)e->ctx->activate();
    (void)hipStreamSynchronize(e->ctx->stream);
    e->release();
  }
  delete e;
  return 0;
}

// ------------------------------------------------------ device memory --

int gt_dev_alloc(gt_ctx* c, int64_t bytes, void** out) {
  if (!c || !out || bytes < 0) return fail("bad argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_HIP(hipMalloc(out, (size_t)std::max<int64_t>(bytes, 256)));
  return 0;
}

int gt_dev_free(gt_ctx* c, void* p) {
  if (!c) return fail("null context");
  if (!p) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_HIP(hipStreamSynchronize(c->stream));
  GT_HIP(hipFree(p));
  return 0;
}

int gt_memcpy_h2d(gt_ctx* c, void* d, const void* h, int64_t bytes, void* stream) {
  if (!c || (bytes > 0 && (!d || !h)) || bytes < 0) return fail("bad argument");
  if (bytes == 0) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  hipStream_t s = c->pick(stream);
  GT_HIP(hipMemcpyAsync(d, h, (size_t)bytes, hipMemcpyHostToDevice, s));
  GT_HIP(hipStreamSynchronize(s));   // the host buffer may be reused on return
  return 0;
}

int gt_memcpy_d2h(gt_ctx* c, void* h, const void* d, int64_t bytes, void* stream) {
  if (!c || (bytes > 0 && (!d || !h)) || bytes < 0) return fail("bad argument");
  if (bytes == 0) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  hipStream_t s = c->pick(stream);
  GT_HIP(hipMemcpyAsync(h, d, (size_t)bytes, hipMemcpyDeviceToHost, s));
  GT_HIP(hipStreamSynchronize(s));
  return 0;
}

int gt_fill_u32_dev(gt_ctx* c, uint32_t* d, int64_t n, uint32_t value, void* stream) {
  if (!c || (n > 0 && !d) || n < 0) return fail("bad argument");
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  GT_HIP(launch_fill_u32(d, n, value, c->pick(stream)));
  return 0;
}

int gt_copy_dev(gt_ctx* c, void* dst, const void* src, int64_t bytes, void* stream) {
  if (!c || bytes < 0 || (bytes % 16) != 0) return fail("copy size must be a multiple of 16");
  if (bytes == 0) return 0;
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  hipEvent_t e0 = nullptr;
  GT_OK(c->prof_begin("copy", c->pick(stream), &e0));
  GT_HIP(launch_copy16(dst, src, bytes, c->pick(stream)));
  GT_OK(c->prof_end("copy", c->pick(stream), e0));
  return 0;
}

int gt_sync(gt_ctx* c, void* stream) {
  if (!c) return fail("null context");
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  hipStream_t s = c->pick(stream);
  GT_HIP(hipStreamSynchronize(s));
  return c->check_ctrl(s);
}

int gt_prof_enable(gt_ctx* c, int32_t on) {
  if (!c) return fail("null context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->prof = on != 0;
  return 0;
}

int gt_prof_read(gt_ctx* c, const char* kernel, int32_t* count, double* total_ms) {
  if (!c || !kernel) return fail("null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  GT_OK(c->activate());
  auto it = c->prof_rec.find(kernel);
  int32_t n = 0;
  double tot = 0.0;
  if (it != c->prof_rec.end()) {
    for (auto& pr : it->second.ev) {
      GT_HIP(hipEventSynchronize(pr.second));
      float ms = 0.f;
      GT_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      tot += ms;
      ++n;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    it->second.ev.clear();
  }
  if (count) *count = n;
  if (total_ms) *total_ms = tot;
  return 0;
}

}  // extern "C"
