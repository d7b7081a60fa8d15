// Host side of the persistent serving kernel (C ABI in include/gtoric.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gtoric.h"
#include "gt_serve.h"

using namespace gt;

namespace {

int sfail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return set_error(buf);
}

#define SV_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return sfail("%s: %s", #call, hipGetErrorString(e_));    \
  } while (0)

#define SV_OK(call)         \
  do {                      \
    int r_ = (call);        \
    if (r_ != 0) return r_; \
  } while (0)

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Projective codewords the search of a k-point request walks: its first
// level's subcodes (k <= 8, searched whole by one wave; more only when a
// lower level follows), or, for larger k, its first phase and its middle
// phase (serve_phases; the level-2 phase is walked only when neither
// certifies the request).
uint64_t walk_cw(int k) {
  if (k <= 1) return 0;
  const int m = (k <= 2) ? 1 : default_level(k);
  uint64_t w = (uint64_t)level_nsub(m) * proj_codewords(k - m);
  if (k > kFusedMaxK && m - 1 > second_level(k))
    w += (uint64_t)level_nsub(m - 1) * proj_codewords(k - m + 1);
  return w;
}

}  // namespace

struct gt_server {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  uint32_t slots = 0;
  uint64_t budget = 0;
  int grid = 0;
  // fine-grained host memory
  ServeHost* h = nullptr;
  ServeReq* ring = nullptr;
  ServeRes* res = nullptr;
  ServeHost* h_dev = nullptr;
  // device memory
  ServeDev dh{};            // host copy of the device control block
  ServeDev* d = nullptr;
  std::vector<void*> dev_allocs;
  uint32_t* mplanes = nullptr;
  // host bookkeeping
  uint64_t next = 0;        // next request index
  uint64_t oldest = 0;      // oldest request whose result is not consumed
  uint64_t outstanding = 0; // codewords of published, unconsumed requests
  // Load-dependent chunk length: a request admitted while fewer than
  // short_below codewords are outstanding walks chunks of short_L[k] steps.
  bool short_on = false;
  uint64_t short_below = 0;
  uint8_t short_L[kServeMaxK + 1] = {};
  std::vector<uint64_t> tag, work;
  std::vector<uint8_t> pending;
  bool running = false;
  uint64_t idle_ticks = 0, life_ticks = 0;
  double tick_s = 1e-8;                     // device wall-clock period
  uint64_t last_stats[kServeStats] = {};    // counters of the last kernel that stopped

  template <class T>
  int dalloc(T** p, size_t n) {
    void* q = nullptr;
    SV_HIP(hipMalloc(&q, std::max<size_t>(n * sizeof(T), 64)));
    dev_allocs.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return 0;
  }

  int launch() {
    // Fresh numbering for a fresh kernel (only with nothing in flight).
    __atomic_store_n(&h->published, 0ull, __ATOMIC_RELEASE);
    __atomic_store_n(&h->consumed, 0ull, __ATOMIC_RELEASE);
    for (uint32_t s = 0; s < slots; ++s) __atomic_store_n(&res[s].word, 0ull, __ATOMIC_RELEASE);
    next = oldest = outstanding = 0;
    std::fill(pending.begin(), pending.end(), 0);
    return launch_at(0);
  }

  // A kernel that starts dispatching at request `first` (the numbering, the
  // ring and `published` are kept).
  int launch_at(uint64_t first) {
    __atomic_store_n(&h->shutdown, 0u, __ATOMIC_RELEASE);
    __atomic_store_n(&h->exited, 0u, __ATOMIC_RELEASE);
    SV_HIP(hipSetDevice(device));
    SV_HIP(launch_serve(grid, d, h_dev, mplanes, first, stream));
    running = true;
    return 0;
  }

  // Make sure a kernel is running.  A kernel that stopped by itself (idle or
  // lifetime limit) still finishes every request it dispatched; requests
  // published after its last dispatcher pass ([consumed, next)) are handed
  // to a new kernel once the dispatched ones are done, so none is lost.
  // *live = false: the old kernel's dispatched requests are still finishing.
  int ensure_live(bool* live) {
    *live = false;
    if (running && !__atomic_load_n(&h->exited, __ATOMIC_ACQUIRE)) {
      *live = true;
      return 0;
    }
    if (running) {
      const uint64_t consumed = __atomic_load_n(&h->consumed, __ATOMIC_ACQUIRE);
      if (oldest < consumed) return 0;
      SV_OK(stop());
      if (next != oldest) {
        SV_OK(launch_at(consumed));
        *live = true;
        return 0;
      }
    }
    SV_OK(launch());
    *live = true;
    return 0;
  }

  int stop() {
    if (!running) return 0;
    __atomic_store_n(&h->shutdown, 1u, __ATOMIC_RELEASE);
    SV_HIP(hipSetDevice(device));
    SV_HIP(hipStreamSynchronize(stream));
    running = false;
    SV_HIP(hipMemcpyAsync(last_stats, d->stats, sizeof(last_stats), hipMemcpyDeviceToHost, stream));
    SV_HIP(hipStreamSynchronize(stream));
    return 0;
  }

  // Collect finished requests in [oldest, next) into *n_done entries; hands
  // undispatched requests of a kernel that stopped by itself to a new one.
  int poll(uint64_t* tags, int32_t* results, int32_t max_n, double* t_done, int32_t* n_done) {
    int n = 0;
    const double t = t_done ? now_s() : 0.0;
    for (uint64_t r = oldest; r < next && n < max_n; ++r) {
      const uint32_t s = (uint32_t)(r & (slots - 1));
      if (!pending[s]) continue;
      const uint64_t w = __atomic_load_n(&res[s].word, __ATOMIC_ACQUIRE);
      if ((w >> 32) != ((r + 1) & 0xFFFFFFFFull)) continue;
      if (tags) tags[n] = tag[s];
      if (results) results[n] = (int32_t)(uint32_t)w;
      if (t_done) t_done[n] = t;
      pending[s] = 0;
      outstanding -= work[s];
      ++n;
    }
    while (oldest < next && !pending[oldest & (slots - 1)]) ++oldest;
    *n_done = n;
    if (running && oldest != next && __atomic_load_n(&h->exited, __ATOMIC_ACQUIRE)) {
      bool live = false;
      SV_OK(ensure_live(&live));
    }
    return 0;
  }

  // Admission: a slot inside the window and room in the work budget.
  bool can_admit(uint64_t w) const {
    if (next - oldest >= slots) return false;
    return outstanding == 0 || outstanding + w <= budget;
  }

  int submit(const int32_t* idx, int32_t k, int32_t tmax, uint64_t t, int32_t* accepted) {
    *accepted = 0;
    if (k < 0 || k > kServeMaxK) return sfail("served requests need 0 <= k <= %d (k=%d)", kServeMaxK, k);
    for (int i = 0; i < k; ++i)
      if (idx[i] < 0 || idx[i] >= kNpts) return sfail("index %d outside 0..48", idx[i]);
    bool live = false;
    SV_OK(ensure_live(&live));
    if (!live) return 0;   // the stopped kernel's requests are still finishing
    const uint64_t w = walk_cw(k);
    if (!can_admit(w)) return 0;
    const uint32_t s = (uint32_t)(next & (slots - 1));
    ServeReq& q = ring[s];
    int32_t Lq = 0;
    if (short_on && k >= 2 && outstanding < short_below && short_L[k] < dh.L[k]) Lq = short_L[k];
    q.k = k | (Lq << 8);
    q.tmax = tmax;
    q.tag = t;
    for (int i = 0; i < (int)sizeof(q.idx); ++i) q.idx[i] = (uint8_t)((i < k) ? idx[i] : 0);
    tag[s] = t;
    work[s] = w;
    pending[s] = 1;
    outstanding += w;
    ++next;
    __atomic_store_n(&h->published, next, __ATOMIC_RELEASE);
    *accepted = 1;
    return 0;
  }
};

extern "C" {

int gt_serve_start(int32_t device, int32_t slots, int64_t work_budget, const int32_t* L_by_k,
                   const int32_t* check_by_k, double idle_s, double life_s, gt_server** out) {
  if (!out) return sfail("null output");
  *out = nullptr;
  if (slots < 1 || slots > (1 << 20) || (slots & (slots - 1)))
    return sfail("slots must be a power of two in 1..2^20 (got %d)", slots);
  gt_server* sv = new gt_server();
  auto bail = [&](int r) {
    for (void* p : sv->dev_allocs) (void)hipFree(p);
    if (sv->h) (void)hipHostFree(sv->h);
    if (sv->ring) (void)hipHostFree(sv->ring);
    if (sv->res) (void)hipHostFree(sv->res);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
    return r;
  };
  sv->device = device;
  sv->slots = (uint32_t)slots;
  sv->budget = (uint64_t)std::max<int64_t>(work_budget, 1);
  if (hipSetDevice(device) != hipSuccess) return bail(sfail("hipSetDevice(%d) failed", device));
  if (hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(sfail("hipStreamCreate failed"));
  const unsigned fl = hipHostMallocCoherent | hipHostMallocMapped;
  if (hipHostMalloc((void**)&sv->h, sizeof(ServeHost), fl) != hipSuccess ||
      hipHostMalloc((void**)&sv->ring, sizeof(ServeReq) * slots, fl) != hipSuccess ||
      hipHostMalloc((void**)&sv->res, sizeof(ServeRes) * slots, fl) != hipSuccess)
    return bail(sfail("hipHostMalloc of the request ring failed"));
  std::memset(sv->h, 0, sizeof(ServeHost));
  std::memset(sv->ring, 0, sizeof(ServeReq) * slots);
  std::memset(sv->res, 0, sizeof(ServeRes) * slots);
  ServeReq* ring_dev = nullptr;
  ServeRes* res_dev = nullptr;
  if (hipHostGetDevicePointer((void**)&sv->h_dev, sv->h, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&ring_dev, sv->ring, 0) != hipSuccess ||
      hipHostGetDevicePointer((void**)&res_dev, sv->res, 0) != hipSuccess)
    return bail(sfail("hipHostGetDevicePointer failed"));
  sv->h->ring = ring_dev;
  sv->h->results = res_dev;
  // Device control block and slot arrays.
  ServeDev& dh = sv->dh;
  dh.slots = (uint32_t)slots;
  int clk_khz = 0;
  if (hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
      clk_khz <= 0)
    clk_khz = 100000;
  sv->tick_s = 1.0 / (clk_khz * 1e3);
  dh.idle_ticks = (uint64_t)(std::max(idle_s, 0.01) * clk_khz * 1e3);
  dh.life_ticks = (uint64_t)(std::max(life_s, 0.01) * clk_khz * 1e3);
  for (int k = 0; k <= kServeMaxK; ++k) {
    const int L = L_by_k ? L_by_k[k] : 10;
    const int c = check_by_k ? check_by_k[k] : 2;
    if (L < kU || L > 30 || c < 0 || c > 16) return bail(sfail("bad serving parameters for k=%d", k));
    dh.L[k] = (uint8_t)L;
    dh.check_log2[k] = (uint8_t)c;
    dh.level[k] = (uint8_t)default_level(k);
  }
  if (sv->dalloc(&dh.s_k, slots) || sv->dalloc(&dh.s_tmax, slots) ||
      sv->dalloc(&dh.s_state, slots) ||
      sv->dalloc(&dh.s_nch, slots) || sv->dalloc(&dh.s_seq, slots) ||
      sv->dalloc(&dh.ord_start, (size_t)4 * slots) ||
      sv->dalloc(&dh.s_idx, (size_t)slots) || sv->dalloc(&dh.s_ph, (size_t)slots) ||
      sv->dalloc(&dh.s_flags, (size_t)slots) || sv->dalloc(&sv->d, 1) ||
      sv->dalloc(&sv->mplanes, (size_t)kNpts * kRowWords))
    return bail(-1);
  // Evaluation-matrix bit planes (same construction as the batch path).
  std::vector<uint32_t> mp((size_t)kNpts * kRowWords, 0u);
  auto mul = [](int a, int b) {
    int r = 0;
    for (int i = 0; i < 3; ++i) { if (b & 1) r ^= a; a <<= 1; if (a & 8) a ^= 0xB; b >>= 1; }
    return r;
  };
  auto pw = [&](int a, int e) { int r = 1; for (int i = 0; i < e; ++i) r = mul(r, a); return r; };
  for (int i = 0; i < kNpts; ++i) {
    uint64_t pl[3] = {0, 0, 0};
    for (int j = 0; j < kNpts; ++j) {
      const int v = mul(pw(j / 7 + 1, i / 7), pw(j % 7 + 1, i % 7));
      for (int p = 0; p < 3; ++p) if ((v >> p) & 1) pl[p] |= 1ull << j;
    }
    for (int p = 0; p < 3; ++p) {
      mp[i * kRowWords + 2 * p] = (uint32_t)pl[p];
      mp[i * kRowWords + 2 * p + 1] = (uint32_t)(pl[p] >> 32);
    }
  }
  if (hipMemcpy(sv->mplanes, mp.data(), mp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(sv->d, &dh, sizeof(ServeDev), hipMemcpyHostToDevice) != hipSuccess)
    return bail(sfail("upload of serving state failed"));
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cu <= 0)
    return bail(sfail("CU count query failed"));
  sv->grid = cu * serve_blocks_per_cu();
  sv->tag.assign(slots, 0);
  sv->work.assign(slots, 0);
  sv->pending.assign(slots, 0);
  if (sv->launch() != 0) return bail(-1);
  *out = sv;
  return 0;
}

int gt_serve_submit(gt_server* sv, const int32_t* idx, int32_t k, int32_t tmax, uint64_t tag,
                    int32_t* accepted) {
  if (!sv || !accepted || (k > 0 && !idx)) return sfail("null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  return sv->submit(idx, k, tmax, tag, accepted);
}

int gt_serve_set_short_chunks(gt_server* sv, const int32_t* L_short_by_k, int64_t below_work) {
  if (!sv) return sfail("null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!L_short_by_k) {
    sv->short_on = false;
    return 0;
  }
  if (below_work < 0) return sfail("negative work threshold");
  for (int k = 0; k <= kServeMaxK; ++k)
    if (L_short_by_k[k] < kU || L_short_by_k[k] > 30)
      return sfail("bad short chunk length for k=%d", k);
  for (int k = 0; k <= kServeMaxK; ++k) sv->short_L[k] = (uint8_t)L_short_by_k[k];
  sv->short_below = (uint64_t)below_work;
  sv->short_on = true;
  return 0;
}

int gt_serve_poll(gt_server* sv, uint64_t* tags, int32_t* results, int32_t max_n, int32_t* n_done) {
  if (!sv || !n_done || max_n < 0) return sfail("bad argument");   // null tags/results: discarded
  std::lock_guard<std::mutex> lk(sv->mu);
  return sv->poll(tags, results, max_n, nullptr, n_done);
}

int gt_serve_in_flight(gt_server* sv, int64_t* n) {
  if (!sv || !n) return sfail("null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *n = (int64_t)(sv->next - sv->oldest);
  return 0;
}

// Drive a whole request stream from this thread: request i is submitted no
// earlier than t0 + arrival_s[i] (in order, as admission allows) and its
// latency is (completion seen by the host) - (t0 + arrival_s[i]).
int gt_serve_run_stream(gt_server* sv, const int32_t* idx, const int32_t* offsets, int32_t n,
                        const int32_t* tmax, const double* arrival_s, int32_t* results,
                        double* latency_s, double* wall_s, double timeout_s) {
  if (!sv || n < 0 || !offsets || !tmax || !arrival_s || !results || !latency_s)
    return sfail("null argument");
  if (n > 0 && offsets[n] > offsets[0] && !idx) return sfail("null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  for (int32_t i = 0; i < n; ++i)
    if (offsets[i + 1] - offsets[i] > kServeMaxK || offsets[i + 1] < offsets[i])
      return sfail("request %d has an unsupported size", i);
  std::vector<uint64_t> tags(1024);
  std::vector<int32_t> rs(1024);
  std::vector<double> td(1024);
  const double t0 = now_s();
  int32_t sent = 0, done = 0;
  while (done < n) {
    const double t = now_s();
    if (t - t0 > timeout_s) return sfail("serving stream timed out after %.1f s (%d/%d done)", timeout_s, done, n);
    while (sent < n && arrival_s[sent] <= t - t0) {
      int32_t acc = 0;
      SV_OK(sv->submit(idx + offsets[sent], offsets[sent + 1] - offsets[sent], tmax[sent],
                       (uint64_t)sent, &acc));
      if (!acc) break;
      ++sent;
    }
    int32_t got = 0;
    SV_OK(sv->poll(tags.data(), rs.data(), 1024, td.data(), &got));
    for (int i = 0; i < got; ++i) {
      const uint64_t r = tags[i];
      results[r] = rs[i];
      latency_s[r] = td[i] - (t0 + arrival_s[r]);
      ++done;
    }
    if (got == 0) std::this_thread::yield();
  }
  if (wall_s) *wall_s = now_s() - t0;
  return 0;
}

int gt_serve_stats(gt_server* sv, double* out, int32_t n) {
  if (!sv || !out) return sfail("null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->next != sv->oldest) return sfail("serve stats: %llu requests still in flight",
                                           (unsigned long long)(sv->next - sv->oldest));
  SV_OK(sv->stop());   // the next submit relaunches the kernel
  for (int i = 0; i < n && i < kServeStats; ++i) {
    const bool ticks = i == kStDispTicks || i == kStWaitTicks || i == kStLookupTicks ||
                       i == kStWalkTicks || i == kStFoldTicks || i == kStLifeTicks;
    out[i] = ticks ? sv->last_stats[i] * sv->tick_s : (double)sv->last_stats[i];
  }
  return 0;
}

int gt_serve_stop(gt_server* sv) {
  if (!sv) return 0;
  int rc;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    rc = sv->stop();
  }
  for (void* p : sv->dev_allocs) (void)hipFree(p);
  (void)hipHostFree(sv->h);
  (void)hipHostFree(sv->ring);
  (void)hipHostFree(sv->res);
  (void)hipStreamDestroy(sv->stream);
  delete sv;
  return rc;
}

}  // extern "C"
