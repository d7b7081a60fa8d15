// AGL2(F7) canonical form of exponent sets S in (Z/7)^2 on gfx950.
//
// Replaces canonical_forms (kernel_canon.py:44-129) and the CPU path
// canonical_forms_batch (champion_search_canonical.py:115-140).
//
// Definition kept from the reference: the key is the minimum, over the 2016
// matrices A of GL2(F7) and the k anchors t of S, of the sorted tuple
// rel(A s_t, A s_i) = A (s_i - s_t) packed 6 bits per position.  For sets of
// one size the packed order equals the order of the 49-bit image masks
// (both compare the largest element first), so the kernel minimises masks:
//   image(A, t) = translate(mask(A S), -(A s_t)),
// held in a byte-row layout (below) that makes the translation cheap.
// Only images that contain point 0 are candidates (the anchored images); the
// minimum over all 98,784 affine images is a different function.  The number
// of (A, t) pairs attaining the minimum is |Stab_AGL2(S)|.
//
// Search (one wavefront per set).  Write A = (r1; r2): r1 = (a, b) gives the
// row u = r1.p of an image point, r2 = (c, d) its column v.  The highest bit of
// an image is in its top occupied row, and after translating by anchor t the
// rows are u_i - u_t (mod 7), so the top row is 6 - (the run of empty rows
// just below row u_t), which depends on r1 and t only.  Phase 1: lane r1
// (48 nonzero first rows) reads its row values, their row set R, the longest
// cyclic run of empty rows below an occupied row of R, and the rows just
// below such a run (the rows that become the top row); the first rows with
// the wave's longest run give the top row of the canonical image.  Next in
// the order is the content of that top row.  For fixed (r1, t) the 42 second
// rows r2 = alpha r1 + beta r1' (beta != 0, r1' = (-b, a)) place the top-row
// points q at columns alpha (v - u_t) - beta s_t + beta s_q, s = r1'.p: every
// affine map x -> a + b x (b != 0) of the points' positions s_q, and nothing
// else, so the smallest top-row content a pair can reach is a table value of
// the 7-bit position pattern, the same for every anchor in the row.  Only the
// first rows and anchor rows reaching the smallest content go on.  Phase 2:
// for every such r1, lane r2 builds the image for A = (r1; r2) (42 valid
// second rows) from its own phase-1 values (the columns) and r1's row shifts
// (read from lane r1 in packed form), and translates it by every remaining
// anchor; those are the only (A, t) pairs whose image can be minimal, so both
// the minimum and the count of pairs that attain it (|Stab|) are exact.
// Ballots combine the lanes and the key is packed in scalar registers.  The
// waves take the sets in input order, dealt so that each XCD works on
// contiguous runs (canon_kernel).
#include <algorithm>

#include "gt_launch.h"

namespace gt {

constexpr int kCanonBlock = 256;
constexpr int kCanonMaxBlocks = 2048;  // persistent grid: 8 waves per SIMD on 256 CUs
constexpr int kRows = 48;              // nonzero vectors (a, b) of F7^2, index 7a + b - 1
constexpr int kRowStride = 52;

// Byte-row layout.  Inside the kernel a subset of the 7x7 torus is a 56-bit
// value with row u in byte u and column v in bit v of that byte (position
// 8u + v; bit 7 of every byte and byte 7 stay 0).  Integer order on this
// layout equals integer order on the 49-bit mask (position 7u + v), because
// both orders compare row 6 first and, inside a row, column 6 first.
//
// Translating by (-u, -v) is then a byte rotation of the 7 rows (two
// v_perm_b32, whose selectors are the row numbers (r + u) mod 7 and 12 for a
// zero byte) and a rotation of every 7-bit row by v: v_bfi(lo, x >> v,
// x << (7 - v)) with lo = bits 0..6-v of every row, then an AND that clears
// bit 7 of every row and byte 7, where the left shift leaves stray bits.
// One 16-byte entry per anchor point (one ds_read_b128 per image).
struct Anchor {
  uint32_t sel_lo, sel_hi;   // v_perm selectors for rows 0..3 and 4..6
  uint32_t vv;               // v | (7 - v) << 8: the 64-bit shifts read bits 0..5
  uint32_t lo;               // bits 0..6-v of every byte
};

constexpr uint32_t kRowBitsLo = 0x7F7F7F7Fu;   // rows 0..3
constexpr uint32_t kRowBitsHi = 0x007F7F7Fu;   // rows 4..6

__device__ __forceinline__ uint64_t translate(uint64_t M, const Anchor& e) {
  const uint32_t m0 = (uint32_t)M, m1 = (uint32_t)(M >> 32);
  const uint64_t X = ((uint64_t)__builtin_amdgcn_perm(m1, m0, e.sel_hi) << 32) |
                     __builtin_amdgcn_perm(m1, m0, e.sel_lo);
  const uint64_t a = X >> (e.vv & 63u), b = X << ((e.vv >> 8) & 63u);
  // v_bitop3 truth table 0xE4: (a & lo) | (b & ~lo).
  const uint32_t r0 = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, e.lo, 0xE4) &
                      kRowBitsLo;
  const uint32_t r1 = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32), e.lo,
                                                  0xE4) & kRowBitsHi;
  return ((uint64_t)r1 << 32) | r0;
}

__host__ __device__ __forceinline__ Anchor make_anchor(uint32_t u, uint32_t v) {
  Anchor e{};
  uint32_t sel[8];
  for (uint32_t r = 0; r < 7; ++r) sel[r] = (r + u) % 7u;
  sel[7] = 12u;
  e.sel_lo = sel[0] | (sel[1] << 8) | (sel[2] << 16) | (sel[3] << 24);
  e.sel_hi = sel[4] | (sel[5] << 8) | (sel[6] << 16) | (sel[7] << 24);
  e.vv = v | ((7u - v) << 8);
  e.lo = (0x7Fu >> v) * 0x01010101u;
  return e;
}

// Tables of one workgroup (4.1 KB of LDS).  Every table a lane indexes with
// its own value is read conflict-free: rowpos rows are 13 dwords apart (odd),
// the 128-byte tables are 32 dwords each, and the 7 anchors one image can
// use are 7 consecutive 16-byte entries.
struct alignas(16) CanonTables {
  Anchor anc[56];                      // translation by byte-row position 8u + v
  uint8_t rowpos[kRows * kRowStride];  // 8 w + s for r = 7a + b - 1 and p = 7x + y:
                                       // w = (a x + b y) mod 7, s = (-b x + a y) mod 7
  uint64_t par[kRows];                 // bit r2 set when vector r2 is parallel to vector r
  uint8_t longest[128];                // longest cyclic run of empty rows just below an occupied
                                       // row of the row set R
  uint8_t toprows[128];                // occupied rows of R just below such a run
  uint8_t amin[128];                   // smallest image of a 7-bit position set under x -> a + b x
};
static_assert(sizeof(Anchor) == 16 && offsetof(CanonTables, rowpos) % 4 == 0 &&
                  (kRows * kRowStride) % 4 == 0 && offsetof(CanonTables, longest) % 4 == 0 &&
                  offsetof(CanonTables, toprows) % 4 == 0 && offsetof(CanonTables, amin) % 4 == 0,
              "dword-written tables must be dword aligned");

__device__ __forceinline__ void build_canon_tables(CanonTables& T) {
  // Whole dwords: byte stores from neighbouring lanes into one dword conflict.
  for (int q = threadIdx.x; q < kRows * kRowStride / 4; q += blockDim.x) {
    uint32_t wd = 0;
    for (int c = 0; c < 4; ++c) {
      const int e = 4 * q + c, r = e / kRowStride, p = e % kRowStride;
      const int a = (r + 1) / 7, b = (r + 1) % 7, x = p / 7, y = p % 7;
      const int w = (a * x + b * y) % 7, sv = ((7 - b) * x + a * y) % 7;
      wd |= (uint32_t)((p < kNpts) ? 8 * w + sv : 0) << (8 * c);
    }
    reinterpret_cast<uint32_t*>(T.rowpos)[q] = wd;
  }
  for (int r = threadIdx.x; r < kRows; r += blockDim.x) {
    const int a = (r + 1) / 7, b = (r + 1) % 7;
    uint64_t m = 0;
    for (int r2 = 0; r2 < kRows; ++r2) {
      const int c = (r2 + 1) / 7, d = (r2 + 1) % 7;
      if ((a * d - b * c + 49) % 7 == 0) m |= 1ull << r2;
    }
    T.par[r] = m;
  }
  for (int q = threadIdx.x; q < 32; q += blockDim.x) {
    uint32_t lw = 0, tw = 0, aw = 0;
    for (int c = 0; c < 4; ++c) {
      const int R = 4 * q + c;
      int gap[7], longest = 0;
      for (int u = 0; u < 7; ++u) {
        int g = 0;
        while (g < 6 && !((R >> ((u + 6 - g) % 7)) & 1)) ++g;
        gap[u] = g;
        if ((R >> u) & 1) longest = max(longest, g);
      }
      uint32_t top = 0;
      for (int u = 0; u < 7; ++u)
        if (((R >> u) & 1) && gap[u] == longest) top |= 1u << ((u + 6 - longest) % 7);
      // Smallest image of the position set R under the 42 maps x -> a + b x.
      uint32_t best = (R == 0) ? 0u : 0x7Fu;
      for (int bm = 1; bm < 7 && R != 0; ++bm) {
        uint32_t img = 0;
        for (int x = 0; x < 7; ++x)
          if ((R >> x) & 1) img |= 1u << ((bm * x) % 7);
        for (int am = 0; am < 7; ++am) {
          const uint32_t rot = ((img << am) | (img >> (7 - am))) & 0x7Fu;
          best = min(best, rot);
        }
      }
      lw |= (uint32_t)longest << (8 * c);
      tw |= top << (8 * c);
      aw |= best << (8 * c);
    }
    reinterpret_cast<uint32_t*>(T.longest)[q] = lw;
    reinterpret_cast<uint32_t*>(T.toprows)[q] = tw;
    reinterpret_cast<uint32_t*>(T.amin)[q] = aw;
  }
  if (threadIdx.x < 56) {
    const uint32_t u = threadIdx.x >> 3, v = threadIdx.x & 7u;
    T.anc[threadIdx.x] = (u < 7 && v < 7) ? make_anchor(u, v) : Anchor{};
  }
}

// Byte-row layout -> 49-bit mask.
__device__ __forceinline__ uint64_t rows_to_mask(uint64_t x) {
  uint64_t m = 0;
#pragma unroll
  for (int u = 0; u < 7; ++u) m |= ((x >> (8 * u)) & 0x7Full) << (7 * u);
  return m;
}

// v_lshl_or_b32 with the shifted operand in a scalar register: (s << v) | acc.
// Written out because the compiler otherwise pairs the shifts into v_or3_b32,
// one more instruction per image point.
__device__ __forceinline__ uint32_t lshl_or_s(uint32_t s, uint32_t v, uint32_t acc) {
  uint32_t d;
  asm("v_lshl_or_b32 %0, %1, %2, %3" : "=v"(d) : "s"(s), "v"(v), "v"(acc));
  return d;
}

struct CanonOut {
  uint64_t key_lo, key_hi, canon;
  uint32_t stab;
};

// Canonical key, 49-bit canonical mask and |Stab| of the set `mask` (|S| = K,
// wave-uniform); every lane returns the same values.  Without kPack only
// |Stab| and the canonical image in byte-row layout (in `canon`) are
// returned: the conversion and the key packing are scalar work that
// canon_keys_kernel then does per lane.  par_mine = T.par[lane]
// (the relation is symmetric: bit r1 says whether this lane's vector is
// parallel to r1).
template <int K, bool kPack>
__device__ __forceinline__ CanonOut canon_set(const CanonTables& T, uint64_t mask, uint32_t lane,
                                              uint64_t par_mine) {
  constexpr int NP = (K + 4) / 5;   // 6-bit fields, five per dword
  uint32_t p[K];
  {
    uint64_t m = mask;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      p[i] = (uint32_t)__builtin_ctzll(m);
      m &= ~(1ull << p[i]);   // one scalar bit clear; m &= m - 1 takes four
    }
  }
  // Phase 1, lane r1 < 48: values z_i = r1.p_i (the row values when the lane
  // stands for r1, the column values when it stands for r2), their row set R,
  // the shifts 8 z_i packed for the readlanes of phase 2, and the lane's image
  // under (r1; r1') in byte-row layout (row z_i, column s_i = r1'.p_i).
  // Lanes 48..63 stand for no vector; they read rows 32..47, the addresses
  // lanes 32..47 read in the same half-wave (a broadcast; row 0 would share
  // a bank with row 32 at another address).  Their values are never used.
  const bool row_lane = lane < (uint32_t)kRows;
  const uint32_t base = (row_lane ? lane : lane - 16u) * kRowStride;
  uint32_t z[K];
  uint32_t P[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) P[j] = 0;
  uint32_t R = 0;
  uint64_t ML = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const uint32_t y = T.rowpos[base + p[i]];
    z[i] = y >> 3;
    R |= 1u << z[i];
    P[i / 5] |= z[i] << (6 * (i % 5) + 3);
    ML |= 1ull << y;
  }
  const uint32_t lg = T.longest[R];
  // First rows reaching the wave's longest run.  Every row lane qualifies at
  // L = 0, so the loop ends.
  uint64_t cand;
  uint32_t L = 6;
  while ((cand = __ballot(row_lane && lg >= L)) == 0) --L;
  const bool is_cand = (cand >> lane) & 1ull;
  // Smallest top-row content each of them allows, over its top rows (usually
  // one), and the top rows that reach it.
  uint32_t cmin = 0xFFu, vq = 0;
  uint32_t rem = is_cand ? (uint32_t)T.toprows[R] : 0u;
  while (__ballot(rem != 0) != 0) {
    if (rem != 0) {
      const uint32_t v = __builtin_ctz(rem);
      rem &= rem - 1;
      const uint32_t c = T.amin[(uint32_t)(ML >> (8 * v)) & 0x7Fu];
      vq = (c < cmin) ? (1u << v) : ((c == cmin) ? (vq | (1u << v)) : vq);
      cmin = min(cmin, c);
    }
  }
  // The wave's smallest content (follow lanes holding a smaller one), then
  // the first rows reaching it and, per lane, the anchor rows: a top row v
  // comes from anchors in row v + lg + 1.
  uint32_t cs = (uint32_t)__builtin_amdgcn_readlane(cmin, __builtin_ctzll(cand));
  for (;;) {
    const uint64_t lower = __ballot(is_cand && cmin < cs);
    if (lower == 0) break;
    cs = (uint32_t)__builtin_amdgcn_readlane(cmin, __builtin_ctzll(lower));
  }
  cand = __ballot(is_cand && cmin == cs);
  const uint32_t n_rot = (lg + 1u) % 7u;
  const uint32_t uq = ((vq << n_rot) | (vq >> (7u - n_rot))) & 0x7Fu;
  uint32_t am = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) am |= ((uq >> z[t]) & 1u) << t;
  // Phase 2, lane r2 < 48 (not parallel to r1), for every such r1: image of
  // A = (r1; r2), translated by each candidate anchor.  Row i of the image
  // enters as (1 << 8 w_i) << z_i, split in 32-bit halves: the shifted value
  // is the uniform one, so the half that does not hold row w_i receives 0.
  uint64_t best = ~0ull;
  uint32_t cnt = 0;
  while (cand != 0) {
    const int r1 = __builtin_ctzll(cand);
    cand &= ~(1ull << r1);
    const uint32_t amr = __builtin_amdgcn_readlane(am, r1);
    uint32_t Pr[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) Pr[j] = __builtin_amdgcn_readlane(P[j], r1);
    if (row_lane & !((par_mine >> r1) & 1ull)) {   // one exec mask, not two nested
      uint32_t lo = 0, hi = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const uint64_t pw = 1ull << ((Pr[i / 5] >> (6 * (i % 5))) & 63u);
        lo = lshl_or_s((uint32_t)pw, z[i], lo);
        hi = lshl_or_s((uint32_t)(pw >> 32), z[i], hi);
      }
      const uint64_t MA = ((uint64_t)hi << 32) | lo;
#pragma unroll
      for (int t = 0; t < K; ++t) {
        if ((amr >> t) & 1u) {
          const uint32_t sh = (Pr[t / 5] >> (6 * (t % 5))) & 63u;
          const uint64_t img = translate(MA, T.anc[sh + z[t]]);
          const bool lt = img < best, eq = img == best;
          cnt = (lt ? 0u : cnt) + (uint32_t)(lt | eq);
          best = lt ? img : best;
        }
      }
    }
  }
  // Wave minimum: follow lanes holding a smaller value (each step lowers m,
  // so at most 64 steps, a few on typical sets), then add the counts of the
  // lanes that hold it.
  uint64_t m = uniform64(best);
  for (;;) {
    const uint64_t lower = __ballot(best < m);
    if (lower == 0) break;
    const int l = __builtin_ctzll(lower);
    m = readlane64(best, l);
  }
  uint64_t eq = __ballot(best == m);
  uint32_t total = 0;
  while (eq != 0) {
    const int l = __builtin_ctzll(eq);
    eq &= ~(1ull << l);
    total += (uint32_t)__builtin_amdgcn_readlane(cnt, l);
  }
  // Sorted positions of the canonical image, packed as kernel_canon.py:84-104:
  // K <= 10 packs positions 0..K-1, K >= 11 skips position 0 (always 0).
  CanonOut o;
  o.stab = total;
  if (!kPack) {
    o.canon = m;
    o.key_lo = o.key_hi = 0;
    return o;
  }
  o.canon = rows_to_mask(m);
  constexpr int pack_from = (K <= 10) ? 0 : 1;
  uint64_t lo = 0, hi = 0, mm = o.canon;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const uint64_t v = (uint64_t)__builtin_ctzll(mm);
    mm &= ~(1ull << v);
    if (i < pack_from) continue;
    const int off = 6 * (i - pack_from);
    if (off + 6 <= 64) lo |= v << off;
    else if (off < 64) { lo |= v << off; hi |= v >> (64 - off); }
    else hi |= v << (off - 64);
  }
  o.key_lo = lo;
  o.key_hi = hi;
  return o;
}

constexpr int kXcds = 8;
constexpr int32_t kCanonSplitMin = 1 << 16;

// One wavefront per set, in input order.  Workgroups are dealt round-robin
// over the XCDs (workgroup b on XCD b % 8), so a pass over nwaves sets gives
// XCD x one contiguous run of them: its workgroups' waves, in order.  Its
// masks are then read, and its four outputs written, in whole cache lines of
// its own L2; with sets dealt wave by wave, neighbouring sets sat on other
// XCDs and every line was written back in pieces from several L2s (six times
// the algorithmic write traffic).  Each wave fetches the mask of its next set
// before working on the current one (a vector load, counted on vmcnt, which
// the LDS waits of the current set do not wait for).  A mask outside 1..16
// points, or with bits above 48, is flagged and its outputs are not written
// here (canon_keys_kernel may still overwrite them: they are unspecified).
// Without kPack the kernel writes |Stab| and the byte-row canonical image
// (into `canon`), and canon_keys_kernel finishes the outputs.
template <bool kPack>
__global__ __launch_bounds__(kCanonBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void canon_kernel(
    const uint64_t* __restrict__ masks, uint32_t n, uint64_t* __restrict__ key_lo,
    uint64_t* __restrict__ key_hi, uint64_t* __restrict__ canon, int32_t* __restrict__ stab,
    uint32_t* ctrl) {
  __shared__ CanonTables T;
  build_canon_tables(T);
  __syncthreads();
  const uint32_t lane = __lane_id();
  const uint64_t par_mine = T.par[lane < (uint32_t)kRows ? lane : 0u];
  constexpr uint32_t kWavesPerBlock = kCanonBlock / kWave;
  const uint32_t G = gridDim.x, nwaves = G * kWavesPerBlock;
  // This wave's place in a pass: the waves of XCDs below its own, then its
  // workgroup's rank on its XCD.
  const uint32_t b = blockIdx.x, x = b % kXcds;
  uint32_t slot = 0;
  for (uint32_t y = 0; y < x; ++y) slot += (G - y + kXcds - 1) / kXcds;
  slot = (slot + b / kXcds) * kWavesPerBlock + threadIdx.x / kWave;
  uint32_t j = slot;
  uint64_t mv = (j < n) ? masks[j] : 0ull;
  bool bad = false;
  for (; j < n; j += nwaves) {
    const uint32_t jn = j + nwaves;
    const uint64_t mv_next = (jn < n) ? masks[jn] : 0ull;
    const uint64_t mask = uniform64(mv);
    CanonOut o;
    bool ok = (mask >> kNpts) == 0;
    if (ok) {
      switch (__popcll(mask)) {
#define GT_CANON_CASE(KK) \
  case KK: o = canon_set<KK, kPack>(T, mask, lane, par_mine); break;
        GT_CANON_CASE(1) GT_CANON_CASE(2) GT_CANON_CASE(3) GT_CANON_CASE(4)
        GT_CANON_CASE(5) GT_CANON_CASE(6) GT_CANON_CASE(7) GT_CANON_CASE(8)
        GT_CANON_CASE(9) GT_CANON_CASE(10) GT_CANON_CASE(11) GT_CANON_CASE(12)
        GT_CANON_CASE(13) GT_CANON_CASE(14) GT_CANON_CASE(15) GT_CANON_CASE(16)
#undef GT_CANON_CASE
        default: ok = false; break;   // sizes outside 1..16
      }
    }
    bad |= !ok;
    if (ok && lane == 0) {
      if (kPack) {
        key_lo[j] = o.key_lo;
        key_hi[j] = o.key_hi;
      }
      canon[j] = o.canon;
      stab[j] = (int32_t)o.stab;
    }
    mv = mv_next;
  }
  if (bad && lane == 0) atomicOr(&ctrl[2], kErrBadK);
}

// (n, k) sorted index rows -> masks; a repeated or out-of-range index is flagged.
__global__ void sets_to_masks_kernel(const int32_t* __restrict__ sets, int32_t n, int32_t k,
                                     uint64_t* __restrict__ masks, uint32_t* ctrl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t m = 0;
  bool bad = false;
  for (int j = 0; j < k; ++j) {
    const int s = sets[(size_t)i * k + j];
    if (s < 0 || s >= kNpts || ((m >> s) & 1ull)) { bad = true; continue; }
    m |= 1ull << s;
  }
  if (bad) { atomicOr(&ctrl[2], kErrBadIndex); m = 0; }
  masks[i] = m;
}

hipError_t launch_sets_to_masks(const int32_t* d_sets, int32_t n, int32_t k, uint64_t* d_masks,
                                uint32_t* d_ctrl, hipStream_t s) {
  hipLaunchKernelGGL(sets_to_masks_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_sets, n, k,
                     d_masks, d_ctrl);
  return hipGetLastError();
}

// Byte-row canonical images -> 49-bit masks and packed keys, one lane per
// set (kernel_canon.py:84-104: sorted positions 6 bits each, position 0
// skipped for |S| >= 11).  A set the canonical kernel flagged left its entry
// unwritten; whatever it holds, at most 16 positions are packed.
__global__ __launch_bounds__(256) void canon_keys_kernel(uint32_t n, uint64_t* __restrict__ canon,
                                                         uint64_t* __restrict__ key_lo,
                                                         uint64_t* __restrict__ key_hi) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = rows_to_mask(canon[i]);
  const int k = __popcll(mask);
  int skip = (k >= 11) ? 1 : 0;
  uint64_t lo = 0, hi = 0, mm = mask;
  int off = 0;
  for (int t = 0; t < kMaxKCanon && mm != 0; ++t) {
    const uint64_t v = (uint64_t)__builtin_ctzll(mm);
    mm &= mm - 1;
    if (skip) { skip = 0; continue; }
    if (off < 64) lo |= v << off;
    if (off + 6 > 64) hi |= (off < 64) ? (v >> (64 - off)) : (v << (off - 64));
    off += 6;
  }
  canon[i] = mask;
  key_lo[i] = lo;
  key_hi[i] = hi;
}

hipError_t launch_canon(const uint64_t* d_masks, int32_t n, uint64_t* d_key_lo,
                        uint64_t* d_key_hi, uint64_t* d_canon, int32_t* d_stab, uint32_t* d_ctrl,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  constexpr uint32_t kWavesPerBlock = kCanonBlock / kWave;
  const int grid = (int)std::min<uint32_t>(kCanonMaxBlocks,
                                           ((uint32_t)n + kWavesPerBlock - 1) / kWavesPerBlock);
  // Large batches leave the key packing to a second, per-lane pass: about a
  // quarter of the wavefront's scalar instructions per set, and the kernel is
  // bound by issue (6.5 % of its time on 10^6 masks).  Small batches (the
  // search's generations) keep one launch.
  const bool split = n >= kCanonSplitMin;
  if (split)
    hipLaunchKernelGGL(canon_kernel<false>, dim3(grid), dim3(kCanonBlock), 0, s, d_masks,
                       (uint32_t)n, d_key_lo, d_key_hi, d_canon, d_stab, d_ctrl);
  else
    hipLaunchKernelGGL(canon_kernel<true>, dim3(grid), dim3(kCanonBlock), 0, s, d_masks,
                       (uint32_t)n, d_key_lo, d_key_hi, d_canon, d_stab, d_ctrl);
  if (split)
    hipLaunchKernelGGL(canon_keys_kernel, dim3(((uint32_t)n + 255) / 256), dim3(256), 0, s,
                       (uint32_t)n, d_canon, d_key_lo, d_key_hi);
  return hipGetLastError();
}

}  // namespace gt
