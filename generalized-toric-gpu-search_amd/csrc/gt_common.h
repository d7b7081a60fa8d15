// Shared definitions for the gfx950 exact-search core (host and device).
//
// Codes.  A code is given by k exponent points S_0..S_{k-1} in (Z/7)^2, index
// i <-> (a, b) = divmod(i, 7); its generator rows are rows S_i of the 49x49
// evaluation matrix M[i, j] = x_j^a * y_j^b over GF(8) = GF(2)[x]/(x^3+x+1)
// (precompute.py:17-43, gf8.py:9-38 in the reference).
//
// Bit planes.  A GF(8) vector of length 49 is three 49-bit planes (bit p of
// every coordinate).  It is held as six 32-bit words
//   w0 = plane0[0:32]  w1 = plane0[32:49]  w2 = plane1[0:32]
//   w3 = plane1[32:49] w4 = plane2[0:32]   w5 = plane2[32:49]
// so the Hamming weight is popc(w0|w2|w4) + popc(w1|w3|w5).  In memory a
// vector ("row") takes kRowWords = 8 words (32 bytes, two pad words) so one
// scalar load fetches it.
//
// Bit rows.  Row r = 3*i + b of a code is x^b * g_i (b = 0, 1, 2).  Multiplying
// by x is the fixed plane mix (p0, p1, p2) -> (p2, p0 ^ p2, p1), because
// x * (c0 + c1 x + c2 x^2) = c2 + (c0 + c2) x + c1 x^2 modulo x^3 + x + 1.
//
// Projective enumeration.  Scalar multiples have the same zero set, so only
// codewords whose last nonzero coefficient is 1 are visited: piece j (0 <= j <
// w) of a w-dimensional code is h_j + span_GF(2){rows 0 .. 3j-1}, 8^j
// codewords, walked in reflected Gray order so that each step adds one bit
// row.  In total (8^w - 1) / 7 codewords.
//
// Subcode through the point (1,1).  Every monomial is 1 at torus point 0 =
// (1,1), so c(1,1) = sum of the coefficients.  The torus translations
// (x,y) -> (sx, ty) act transitively on the 49 points and map C(S) onto
// itself (monomial x^a y^b scales by s^a t^b), so a codeword with z > 0 zeros
// has a translate of the same weight that vanishes at (1,1).  Hence for
// k >= 2 the maximum number of zeros over C(S) is the maximum over the
// (k-1)-dimensional subcode of codewords vanishing at (1,1), spanned by
// h_i = g_i + g_{k-1} (i < k-1); for k = 1 it is 0 (a monomial has no zero
// on the torus).  A repeated index still gives the zero codeword there
// (g_a = g_b => h_a + h_b = 0, coefficients summing to 0), as in the
// reference.  The exact walk therefore visits (8^(k-1) - 1) / 7 codewords
// per code against the reference's 8^k - 1 (kernel.py:128-139).
//
// Anchored subcodes (level m).  Write a torus point as (alpha^u, alpha^v),
// (u, v) in (Z/7)^2, alpha = x.  Besides the translations (u, v) -> (u, v) +
// t, the Frobenius map c -> c^2 applied to the coefficients preserves weight:
// f_{c^2}(P^2) = f_c(P)^2, so the codeword of the squared coefficients
// vanishes on the image of f_c's zero set under (u, v) -> 2 (u, v).  The
// group G of the maps (u, v) -> 2^i (u, v) + t (order 147) therefore maps
// zero sets of codewords of C(S) to zero sets of codewords of the same
// weight.  Fix a direction d (one of the 8 lines through 0 of (Z/7)^2); its
// 7 cosets partition the 49 points.  A codeword with z >= 7 (m - 1) + 1
// zeros has m zeros in one coset, {P + r d : r in A}, |A| = m; a map of G
// sends them to {r d : r in R}, where R is the representative of A's orbit
// under the maps r -> 2^i r + t of Z/7 (level_rep).  So when the maximum
// number of zeros z* is at least level_cert(m) = 7 (m - 1) + 1, it equals
// the maximum over the subcodes C_R = {c : c(r d) = 0 for r in R}, R over
// the level's representatives (1 for m = 1, 2, 5, 6; 3 for m = 3, 4), and
// every codeword walked is a codeword of C(S), so the maximum z_m found is
// at most z* in any case.  A found maximum certifies itself one zero
// earlier: a codeword with m zeros in some coset is carried into one of the
// subcodes, so it has at most z_m zeros, and any other codeword has at most
// m - 1 zeros in each of the 7 cosets; so z* <= max(z_m, 7 (m - 1)), and
// z_m >= found_cert(m) = 7 (m - 1) proves z_m = z*.  A code below that is
// walked again at a level whose certificate a known bound on z* meets.  Any k - 1
// points are the zeros of some nonzero codeword (k coefficients, k - 1
// conditions), so z* >= k - 1: level 2 is certified for every code of k >= 9
// and level 1 (the subcode through (1,1) above, certified whenever z* >= 1)
// for all.  C_R has dimension k - m when the m conditions are independent;
// the rows are a basis from Gaussian elimination on the codewords in the
// first direction where all representatives give m independent conditions
// (a code with none is also searched again).  At level 2 that fails only
// when every index of the set is the same point, and such a set spans the
// zero codeword (49 zeros).  A level-3 code walks 3 (8^(k-3) - 1) / 7
// codewords, 21x fewer than level 1.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GT_HD __host__ __device__ __forceinline__
#else
#define GT_HD inline
#endif

namespace gt {

constexpr int kNpts = 49;
constexpr int kRowWords = 8;
constexpr int kMaxKExact = 21;   // kernel.py:246-247, 290-291
constexpr int kMaxK = 49;        // kernel.py:288-289 (sampled)
constexpr int kMaxKCanon = 16;   // kernel_canon.py:161-162
constexpr int kWave = 64;
constexpr int kLaneBits = 6;     // 64 lanes own 6 Gray index bits

// Dimension of the exact walk of a code with k generators (see above).
GT_HD int walk_dim(int k) { return (k <= 1) ? 0 : k - 1; }

// Passes of the list search over the codes of one call (see the anchored
// subcodes above): the first at the code's level, the middle one level lower
// for the codes whose first-pass maximum certifies that level, the final one
// at second_level(k) for the rest.
constexpr int kPassFirst = 0, kPassMid = 1, kPassFinal = 2, kPasses = 3;

// Per-k launch parameters chosen by the autotuner, indexed by the code's
// generator count k (the walk has dimension k - 1).
struct MdParams {
  uint8_t L[kMaxKExact + 1];      // Gray walk bits per lane for full chunks
  uint8_t check_log2[kMaxKExact + 1];  // early-exit check every 2^x blocks
  uint8_t level[kMaxKExact + 1];  // anchor level m of the first pass
  // Per pass, the largest k whose subcodes in that pass have no piece above
  // the tiny part (walk dimension <= kSmallW): the small codes of the pass.
  uint8_t small_k[kPasses];
  // Chunks of a code of k generators at L[k] in each pass (code_level) and
  // at level 1 (the limit a call is refused at: 2^31 chunks in total),
  // saturated at 2^31; kept with L and level by md_params_set / md_level_set.
  uint32_t chunks[kPasses][kMaxKExact + 1];
  uint32_t chunks1[kMaxKExact + 1];
};

// Chunk plan of one batch, largest codes first.  Bucket j (0 <= j < kMaxKExact)
// holds the codes whose walk has dimension kMaxKExact - j (k = kMaxKExact -
// j + 1 generators), in batch order, at positions pb[j] .. pb[j+1] - 1 of the
// order array (pb[kMaxKExact] = codes in the plan).  For the claimed
// (thresholded) walk its codes own chunk ids cb[j] + q * cc[j] .. + cc[j] - 1
// for the q-th code; cb[kMaxKExact] is the total chunk count (0 when the plan
// overflowed 2^31 chunks).  Claiming ids in increasing order hands out the
// biggest chunks first, so the last chunks of a launch are the small ones of
// the small codes.  For the dealt (exact) walk the q-th code owns work units
// ub[j] + q * anchor_units(k, m) .. (ub[kMaxKExact] = total units).  A code
// of level m has level_nsub(m) subcodes of walk dimension k - m; its chunks
// (and units) cover them one after the other.
struct MdPlan {
  uint64_t ub[kMaxKExact + 1];
  uint32_t cb[kMaxKExact + 1];
  uint32_t pb[kMaxKExact + 1];
  uint32_t cc[kMaxKExact];
};

// Chunk claims are spread over kClaimShards counters, kClaimStride words
// (4 KB) apart: one device-scope atomic on one address completes about every
// 11 ns, which small chunks would otherwise queue on.  Shard s hands out the
// chunk ids s, s + kClaimShards, s + 2 kClaimShards, ... in increasing order.
constexpr int kClaimShards = 16;
constexpr int kClaimStride = 1024;

// Static unroll of the inner Gray walk: 2^kU steps per block with rows
// 0..kU-1 held in scalar registers.
constexpr int kU = 4;
// Pieces j with 3j < kLaneBits + kU are enumerated directly by the "tiny" chunk.
constexpr int kTinyTop = (kLaneBits + kU - 1) / 3;   // largest tiny j

// Functions of the walk below take its dimension w (walk_dim(k) for a code).
GT_HD int tiny_top(int w) { return (w - 1 < kTinyTop) ? (w - 1) : kTinyTop; }

// Number of chunks of Gray piece j with walk length L (L >= kU).
GT_HD uint64_t piece_chunks(int j, int L) {
  int B = 3 * j;
  return (B >= kLaneBits + L) ? (1ull << (B - kLaneBits - L)) : 1ull;
}
GT_HD int piece_L(int j, int L) {
  int B = 3 * j;
  return (B >= kLaneBits + L) ? L : (B - kLaneBits);
}

// Chunks of a walk of dimension w: Gray pieces j = w-1 .. tiny_top(w)+1, then
// one tiny chunk.
GT_HD uint64_t code_chunks(int w, int L) {
  if (w <= 0) return 0;
  uint64_t n = 1;  // tiny chunk
  for (int j = w - 1; j > tiny_top(w); --j) n += piece_chunks(j, L);
  return n;
}

// ---- anchored subcodes (see the top of this file) ------------------------
constexpr int kMaxLevel = 6;   // (a level-7 reduction spills registers; it would need z >= 43)
// Representatives of the m-subsets of Z/7 under r -> 2^i r + t, one per
// orbit, chosen to share the prefix {0, .., m-2}: representative s is that
// prefix and the point level_last(m, s).  The orbits have 7 (m = 1, 6), 21
// (m = 2, 5) and 21 + 7 + 7 (m = 3, 4: {0,1,2}, {0,1,3}, {0,1,5} and
// {0,1,2,3}, {0,1,2,4}, {0,1,2,5}) members.
constexpr GT_HD int level_nsub(int m) { return (m == 3 || m == 4) ? 3 : 1; }
constexpr GT_HD int level_last(int m, int s) {
  return (m == 3) ? (s == 0 ? 2 : s == 1 ? 3 : 5) : (m == 4) ? (s == 0 ? 3 : s == 1 ? 4 : 5) : m - 1;
}
// The representative as a bit mask over r (bit r set <=> r in R).
constexpr GT_HD uint32_t level_rep(int m, int s) {
  return ((1u << (m - 1)) - 1u) | (1u << level_last(m, s));
}
// A lower bound on z* that makes level m exact (7 (m - 1) + 1), and a
// maximum found by a level-m walk that is exact (7 (m - 1); see the top of
// this file).
GT_HD int level_cert(int m) { return 7 * m - 6; }
GT_HD int found_cert(int m) { return 7 * m - 7; }
// Level of the final pass over a code of k generators: 2 when k - 1 meets
// its certificate (k >= 9), else 1.
GT_HD int second_level(int k) { return (k >= 9) ? 2 : 1; }
// Level of the middle pass: one below the first pass's when that is above
// the final pass's, else 0 (no middle pass for k).
GT_HD int mid_level(const uint8_t* level, int k) {
  if (k <= 2 || k > kMaxKExact) return 0;
  const int m = (int)level[k] - 1;
  return (m > second_level(k)) ? m : 0;
}
// Level of a code of k generators in a pass (codes with k <= 2: level 1; k
// above the exact range, which no plan holds: 1; the middle pass of a k
// without one, which no code enters: the final pass's level).
GT_HD int code_level(const uint8_t* level, int k, int pass) {
  if (k <= 2 || k > kMaxKExact) return 1;
  if (pass == kPassFirst) return (int)level[k];
  if (pass == kPassMid && mid_level(level, k) > 0) return mid_level(level, k);
  return second_level(k);
}
// Walk dimension of one subcode of a code of k generators at level m.
GT_HD int anchor_dim(int k, int m) { return (k <= 1) ? 0 : k - m; }
// Row slots per index: a code's subcodes take nsub * 3 (k - m) <= 9k rows.
constexpr int kRowsPerIdx = 9;
// Codes of at most kFusedMaxK generators are searched one wave per code
// (rows in LDS, every level in turn); larger ones by the plan-and-deal walk.
constexpr int kFusedMaxK = 8;

// Default first-pass level per k: level m walks level_nsub(m) subcodes of
// dimension k - m and certifies a maximum of at least 7m - 6 zeros.  On the
// 4096-code mixed batch (k = 3..8) these levels certify every code: k = 4..6
// have at least 9 zeros (level 2 needs 8), k = 7, 8 at least 15 (level 3);
// k = 3 codes often have 7 (level 1).  Random codes of k = 9..12 have 19 to
// 28 zeros, mostly 25 or 28: level 4 (22) certifies about three in four
// (k = 10: 3 x 37,449 codewords against 3 x 299,593 at level 3), and the
// others meet level 3's 15 in the first pass, which sends them to the
// middle pass at level 3.
GT_HD int default_level(int k) { return (k <= 3) ? 1 : (k <= 6) ? 2 : (k <= 8) ? 3 : 4; }

// L and check interval of k, with the chunk counts the plan reads.
inline uint32_t sat31(uint64_t c) { return (c >= (1ull << 31)) ? (1u << 31) : (uint32_t)c; }
inline void md_params_refresh(MdParams& p, int k) {
  for (int pass = 0; pass < kPasses; ++pass) {
    const int m = code_level(p.level, k, pass);
    p.chunks[pass][k] = sat31((uint64_t)level_nsub(m) * code_chunks(anchor_dim(k, m), p.L[k]));
    int sk = 1;
    for (int q = 2; q <= kMaxKExact; ++q) {
      if (anchor_dim(q, code_level(p.level, q, pass)) > kTinyTop + 1) break;
      sk = q;
    }
    p.small_k[pass] = (uint8_t)sk;
  }
  p.chunks1[k] = sat31(code_chunks(walk_dim(k), p.L[k]));
}
inline void md_params_set(MdParams& p, int k, int L, int check_log2) {
  p.L[k] = (uint8_t)L;
  p.check_log2[k] = (uint8_t)check_log2;
  md_params_refresh(p, k);
}
// Levels must keep the first-pass walk dimension k - level[k] non-decreasing
// in k (the small codes are then the plan's last buckets): the caller checks
// level[k + 1] <= level[k] + 1 and 1 <= level[k] <= min(k - 1, kMaxLevel).
inline void md_level_set(MdParams& p, int k, int m) {
  p.level[k] = (uint8_t)m;
  md_params_refresh(p, k);
}

// Work units of the dealt walk.  A block is 16 Gray steps of all 64 lanes
// (1024 codewords, 174 VALU instructions).  Piece j > kTinyTop is walked
// lane-major: lane l owns the Gray indices l * 2^(3j-6) + [0, 2^(3j-6)), and
// the piece is 2^(3j-10) blocks, one unit each; a wave may walk any run of
// them.  The tiny part (pieces 0 .. tiny_top(w): 585 codewords for
// w > kTinyTop, about 160 VALU instructions in one wave) counts as
// kTinyUnits units, which also prices the start of a code (two dependent
// loads before its first row) and its fold: measured on a 4096-code mix,
// weights 6 to 8 balance best, 2 leaves the waves that hold many k = 6 or 7
// codes the last to finish.  A code's units are its pieces from the largest
// down, then its tiny part.  Codes of walk dimension <= kSmallW have no piece
// above kTinyTop; they are dealt as whole codes instead (md_main_exact).
constexpr int kBlockBits = kLaneBits + kU;
constexpr int kSmallW = kTinyTop + 1;
constexpr uint32_t kTinyUnits = 6;
GT_HD uint64_t piece_blocks(int j) { return 1ull << (3 * j - kBlockBits); }
// kTinyUnits + sum of 2^(3j-10) over j = kSmallW .. w-1 = kTinyUnits +
// 2^(3 kSmallW - 10) * (8^m - 1) / 7, m = w - kSmallW, where (8^m - 1) / 7 is
// the base-8 number of m ones: bits 0, 3, ..., 3(m-1), a mask (no division).
constexpr uint64_t kOctalOnes = 0x9249249249249249ull;   // bits 0, 3, 6, ..., 63
static_assert(3 * kSmallW >= kBlockBits, "pieces above the tiny part hold whole blocks");
GT_HD uint64_t code_units(int w) {
  if (w <= 0) return 0;
  if (w <= kSmallW) return kTinyUnits;
  return kTinyUnits + ((kOctalOnes & ((1ull << (3 * (w - kSmallW))) - 1ull))
                       << (3 * kSmallW - kBlockBits));
}

// Work units of a code of k generators at level m: its subcodes one after
// the other, each code_units(k - m).
GT_HD uint64_t anchor_units(int k, int m) {
  return (k <= 1) ? 0 : (uint64_t)level_nsub(m) * code_units(anchor_dim(k, m));
}

// Projective codewords of a w-dimensional code: (8^w - 1) / 7.
GT_HD uint64_t proj_codewords(int w) {
  return (w <= 0) ? 0 : (((1ull << (3 * w)) - 1ull) / 7ull);
}

// x * v for a bit-plane vector (six words).
GT_HD void mul_x(uint32_t w[6]) {
  uint32_t p0lo = w[0], p0hi = w[1], p1lo = w[2], p1hi = w[3], p2lo = w[4], p2hi = w[5];
  w[0] = p2lo;        w[1] = p2hi;
  w[2] = p0lo ^ p2lo; w[3] = p0hi ^ p2hi;
  w[4] = p1lo;        w[5] = p1hi;
}

// Error bits reported through the control block.
enum : uint32_t {
  kErrBadIndex = 1u,
  kErrBadK = 2u,
  kErrFull = 4u,   // hash set has no free slot
};

#if defined(__HIPCC__)
// Wave-uniform copies of 64-bit values.  The lane-read builtins return int:
// each half is converted to uint32_t before widening, or a low half with
// bit 31 set would be sign-extended over the high half.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}
#endif

}  // namespace gt
