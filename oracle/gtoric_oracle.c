/*
 * CPU oracle for the exact-search core of GF(8) generalized toric codes (n = 49).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithm; it exists to check the HIP path (tests, smoke, and the
 * cpu_baseline leg of bench.py).  Product code never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pins (what checks
 * each function independently of this restatement) are listed in
 * oracle/__init__.py and exercised by tests/test_oracle.py.
 *
 * Thresholded results.  The reference batch kernels return, on early abort,
 * "some value > tmax" chosen by thread scheduling (kernel.py:143,181-185;
 * contract in "kernel backup.py":116-125).  The oracle therefore returns the
 * deterministic value min(true_max, tmax + 1): exact when <= tmax, and exactly
 * tmax + 1 when the code fails.  Parity with the reference is defined on that
 * clamp.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NPTS 49
#define MASK49 ((1ULL << 49) - 1ULL)

/* ---------------------------------------------------------------- GF(8) -- */

/* gf8.py:11-21 — carry-less multiply modulo x^3 + x + 1 (POLY 0b1011, gf8.py:9). */
static int gf8_mul(int a, int b) {
    int r = 0;
    for (int i = 0; i < 3; ++i) {
        if (b & 1) r ^= a;
        a <<= 1;
        if (a & 8) a ^= 0xB;
        b >>= 1;
    }
    return r;
}

/* gf8.py:29-34 */
static int gf8_pow(int base, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r = gf8_mul(r, base);
    return r;
}

/* gf8.py:24-26 (MUL_TABLE, row-major a*8+b). */
void or_mul_table(int32_t out[64]) {
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) out[a * 8 + b] = gf8_mul(a, b);
}

/* precompute.py:17-43 — rows: lattice (a,b) = divmod(i,7); columns: torus
 * (x,y) = (j//7 + 1, j%7 + 1); M[i,j] = x^a * y^b. */
void or_eval_matrix(int32_t M[NPTS * NPTS]) {
    for (int i = 0; i < NPTS; ++i) {
        int a = i / 7, b = i % 7;
        for (int j = 0; j < NPTS; ++j) {
            int x = j / 7 + 1, y = j % 7 + 1;
            M[i * NPTS + j] = gf8_mul(gf8_pow(x, a), gf8_pow(y, b));
        }
    }
}

static int32_t g_M[NPTS * NPTS];
static int g_mul[64];
static int g_init = 0;

static void ensure_init(void) {
    if (g_init) return;
    or_eval_matrix(g_M);
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) g_mul[a * 8 + b] = gf8_mul(a, b);
    g_init = 1;
}

/* Call once before using the oracle from several threads. */
void or_init(void) { ensure_init(); }

/* ------------------------------------------------------------ splitmix64 -- */

/* kernel.py:10-15 — SplitMix64 finalizer (Vigna). */
uint64_t or_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* kernel.py:17-29 */
static int sampled_coeff(uint64_t work_id, int set_id, int coeff_idx, uint64_t seed) {
    uint64_t x = seed ^ ((uint64_t)(int64_t)set_id * 0x9E3779B97F4A7C15ULL) ^
                 (work_id * 0xBF58476D1CE4E5B9ULL) ^
                 ((uint64_t)(int64_t)(coeff_idx + 1) * 0x94D049BB133111EBULL);
    return (int)(or_splitmix64(x) & 7ULL);
}

/* ------------------------------------------------------------- bitpat ----- */

/* kernel.py:35-60 — bitpat[b][i][v] = 49-bit mask of bit b of GF8_MUL[v*8 + M[S_i, j]]. */
static void build_bitpat(const int32_t* S, int k, uint64_t* bp) {
    for (int b = 0; b < 3; ++b)
        for (int i = 0; i < k; ++i)
            for (int v = 0; v < 8; ++v) {
                uint64_t m = 0;
                const int32_t* row = g_M + S[i] * NPTS;
                for (int j = 0; j < NPTS; ++j)
                    if ((g_mul[v * 8 + row[j]] >> b) & 1) m |= 1ULL << j;
                bp[(b * k + i) * 8 + v] = m;
            }
}

/*
 * kernel.py:97-203 (eval_min_distance_batch_bp), one set, restated sequentially:
 * work_id = 1..work_items; poly_id = work_id (exact, 8^k - 1 polys, no scalar
 * quotient) or the sampled stream (kernel.py:136-178); zeros = popcount of the
 * 49-bit zero mask; abort when zeros > tmax.  Returns min(max_zeros, tmax + 1).
 * set_id is the index of the set inside its batch (it seeds the sampled stream).
 */
int32_t or_max_zeros_one(const int32_t* S, int32_t k, int32_t tmax,
                         int64_t sample_count, uint64_t sample_seed, int32_t set_id) {
    ensure_init();
    if (k <= 0) return 0;
    uint64_t* bp = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (size_t)k * 8);
    build_bitpat(S, k, bp);
    uint64_t total_polys = (3 * k <= 63) ? ((1ULL << (3 * k)) - 1ULL) : ~0ULL;
    uint64_t work_items = (sample_count > 0) ? (uint64_t)sample_count : total_polys;
    int use_poly_id_sampling = (sample_count > 0 && 3 * k <= 63);
    int best = 0;
    int failed = 0;
    for (uint64_t work_id = 1; work_id <= work_items; ++work_id) {
        uint64_t poly_id = work_id;
        if (use_poly_id_sampling) {
            uint64_t x = sample_seed ^ ((uint64_t)(int64_t)set_id * 0x9E3779B97F4A7C15ULL) ^
                         (work_id * 0xBF58476D1CE4E5B9ULL);
            poly_id = (or_splitmix64(x) % total_polys) + 1ULL;
        }
        int forced_idx = 0;
        if (sample_count > 0 && !use_poly_id_sampling) {
            uint64_t x = sample_seed ^ ((uint64_t)(int64_t)set_id * 0xD6E8FEB86659FD93ULL) ^
                         (work_id * 0xA5A3564E27F8861FULL);
            forced_idx = (int)(or_splitmix64(x) % (uint64_t)k);
        }
        uint64_t v0 = 0, v1 = 0, v2 = 0, tmp = poly_id;
        for (int i = 0; i < k; ++i) {
            int c;
            if (sample_count > 0 && !use_poly_id_sampling) {
                c = sampled_coeff(work_id, set_id, i, sample_seed);
                if (i == forced_idx && c == 0) c = 1;
            } else {
                c = (int)(tmp & 7ULL);
                tmp >>= 3;
            }
            v0 ^= bp[(0 * k + i) * 8 + c];
            v1 ^= bp[(1 * k + i) * 8 + c];
            v2 ^= bp[(2 * k + i) * 8 + c];
        }
        int zeros = __builtin_popcountll((~v0) & (~v1) & (~v2) & MASK49);
        if (zeros > tmax) { failed = 1; break; }
        if (zeros > best) best = zeros;
    }
    free(bp);
    return failed ? tmax + 1 : best;
}

/*
 * kernel.py:276-323 (max_zeros_batch) with the one-set-per-work-group mapping
 * (kernel.py:306-320) turned into one set per OpenMP thread.  Mixed k is
 * allowed through offsets[n+1]; tmax is per set (tmax_per_set != NULL) or the
 * scalar tmax.  set_id for the sampled stream is the index inside this call.
 */
void or_max_zeros_batch(const int32_t* idx, const int32_t* offsets, int32_t n,
                        const int32_t* tmax_per_set, int32_t tmax,
                        int64_t sample_count, uint64_t sample_seed,
                        int32_t* out, int32_t nthreads) {
    ensure_init();
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int32_t s = 0; s < n; ++s) {
        int32_t t = tmax_per_set ? tmax_per_set[s] : tmax;
        out[s] = or_max_zeros_one(idx + offsets[s], offsets[s + 1] - offsets[s], t,
                                  sample_count, sample_seed, s);
    }
}

/*
 * "kernel backup.py":203-220 (bf_max_zeros) — direct evaluation with the
 * multiplication table, no bit planes, no abort.  A second, independent
 * algorithm for the exact value.
 */
int32_t or_max_zeros_brute(const int32_t* S, int32_t k) {
    ensure_init();
    if (k <= 0) return 0;
    int best = 0;
    uint64_t total = (1ULL << (3 * k));
    int coeffs[21];
    for (uint64_t poly_id = 1; poly_id < total; ++poly_id) {
        uint64_t tmp = poly_id;
        for (int i = 0; i < k; ++i) { coeffs[i] = (int)(tmp & 7); tmp >>= 3; }
        int zeros = 0;
        for (int j = 0; j < NPTS; ++j) {
            int val = 0;
            for (int i = 0; i < k; ++i) val ^= g_mul[coeffs[i] * 8 + g_M[S[i] * NPTS + j]];
            if (val == 0) ++zeros;
        }
        if (zeros > best) best = zeros;
    }
    return best;
}

/* --------------------------------------------------------- canonical ----- */

/* champion_search_canonical.py:61-74 — GL2(F7) as point permutations, in the
 * product(range(7), repeat=4) order of (a,b,c,d) with det != 0; point
 * (x,y) = divmod(p,7) maps to ((a x + b y) % 7) * 7 + (c x + d y) % 7. */
void or_gl2_perms(int32_t perms[2016 * NPTS]) {
    int n = 0;
    for (int a = 0; a < 7; ++a)
        for (int b = 0; b < 7; ++b)
            for (int c = 0; c < 7; ++c)
                for (int d = 0; d < 7; ++d) {
                    if (((a * d - b * c) % 7 + 7) % 7 == 0) continue;
                    for (int p = 0; p < NPTS; ++p) {
                        int x = p / 7, y = p % 7;
                        perms[n * NPTS + p] = ((a * x + b * y) % 7) * 7 + (c * x + d * y) % 7;
                    }
                    ++n;
                }
}

/* champion_search_canonical.py:99-108 — rel[a,b] = index of (p_b - p_a) mod 7. */
void or_rel_table(uint8_t rel[NPTS * NPTS]) {
    for (int a = 0; a < NPTS; ++a)
        for (int b = 0; b < NPTS; ++b) {
            int da = ((b / 7 - a / 7) % 7 + 7) % 7;
            int db = ((b % 7 - a % 7) % 7 + 7) % 7;
            rel[a * NPTS + b] = (uint8_t)(da * 7 + db);
        }
}

static int32_t g_perms[2016 * NPTS];
static uint8_t g_rel[NPTS * NPTS];
static int g_canon_init = 0;

static void ensure_canon_init(void) {
    if (g_canon_init) return;
    or_gl2_perms(g_perms);
    or_rel_table(g_rel);
    g_canon_init = 1;
}

void or_canon_init(void) { ensure_init(); ensure_canon_init(); }

/*
 * One set.  champion_search_canonical.py:115-140 (k <= 10) and
 * kernel_canon.py:44-129 (k <= 16): min over the 2016 GL2 permutations and the
 * k anchors of pack(sort(rel[perm[S_anchor], perm[S_i]])), 6 bits per sorted
 * position; pack_from = 0 for k <= 10, 1 for k >= 11 (position 0 is always 0);
 * 128-bit order compares hi first.  stab counts the (perm, anchor) pairs that
 * attain the minimum; that count equals |Stab_AGL2(S)|.
 */
static void canon_one(const int32_t* S, int k, uint64_t* lo_out, uint64_t* hi_out, int32_t* stab_out) {
    uint64_t best_lo = ~0ULL, best_hi = ~0ULL;
    int32_t count = 0;
    int pack_from = (k <= 10) ? 0 : 1;
    int perm[16], t[16];
    for (int pi = 0; pi < 2016; ++pi) {
        for (int i = 0; i < k; ++i) perm[i] = g_perms[pi * NPTS + S[i]];
        for (int anc = 0; anc < k; ++anc) {
            int anchor = perm[anc];
            for (int i = 0; i < k; ++i) t[i] = g_rel[anchor * NPTS + perm[i]];
            for (int i = 1; i < k; ++i) {
                int v = t[i], j = i;
                while (j > 0 && t[j - 1] > v) { t[j] = t[j - 1]; --j; }
                t[j] = v;
            }
            uint64_t lo = 0, hi = 0;
            for (int i = pack_from; i < k; ++i) {
                int off = 6 * (i - pack_from);
                uint64_t v = (uint64_t)t[i];
                if (off + 6 <= 64) lo |= v << off;
                else if (off < 64) { lo |= v << off; hi |= v >> (64 - off); }
                else hi |= v << (off - 64);
            }
            if (hi < best_hi || (hi == best_hi && lo < best_lo)) {
                best_hi = hi; best_lo = lo; count = 1;
            } else if (hi == best_hi && lo == best_lo) {
                ++count;
            }
        }
    }
    *lo_out = best_lo;
    *hi_out = best_hi;
    if (stab_out) *stab_out = count;
}

/* kernel_canon.py:151-176 (CanonicalOracle.compute) shape: sets (n,k) int32
 * sorted rows, k <= 16; outputs lo/hi uint64 and the stabilizer order. */
int32_t or_canon_batch(const int32_t* sets, int32_t n, int32_t k,
                       uint64_t* lo, uint64_t* hi, int32_t* stab, int32_t nthreads) {
    if (k < 1 || k > 16) return -1;
    ensure_init();
    ensure_canon_init();
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int32_t s = 0; s < n; ++s)
        canon_one(sets + (size_t)s * k, k, lo + s, hi + s, stab ? stab + s : NULL);
    return 0;
}
