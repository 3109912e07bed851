// CPU oracle for the CFPQ Boolean-matrix fixpoint.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped engine links, loads or
// calls this file: only the tests, the smoke check and the benchmark's CPU
// baseline leg use it, as the checker.
//
// It restates, on the CPU, the reference's algorithm for this path:
//   * the worklist fixpoint driver of cfpq-cpp/Grammar.h:58-69 (pop order is
//     that of std::unordered_set<unsigned>, as in Grammar.h:59,93);
//   * the oracle matrix backend M4riMatrix::add_mul of m4ri/m4ri_matrix.cpp:20-28:
//     C := C | A*B over the Boolean semiring, computed here by a clean-room
//     "method of four Russians" (k = 8 Gray-free subset tables) on row-major
//     uint64 rows (column j -> word j/64, bit j%64, the m4ri convention), and
//     "changed" = result differs from the old C;
//   * the result fingerprint of Grammar::print_results, cfpq-cpp/Grammar.cpp:81-101.
// It also holds two independent checkers: a sparsity-aware semi-naive
// fixpoint (row-gather product, used at sizes where dense M4RM is too slow)
// and Hellings' worklist CFPQ over edge facts (no matrices at all).
//
// Matrix layout everywhere: n rows, wpr = ceil(n/64) words per row.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;

static inline size_t wpr_of(uint32_t n) { return (n + 63u) / 64u; }

extern "C" {

// ---------------------------------------------------------------------------
// Boolean products
// ---------------------------------------------------------------------------

// R |= A*B by Four Russians with k = 8: for each group of 8 rows of B build the
// 256-entry table of all OR-combinations, then each row of A contributes one
// table row per byte of A.  Column stripes keep the 8 tables of a 64-row group
// cache resident.  R must not alias A or B.
static void m4rm_into(u64 *R, const u64 *A, const u64 *B, uint32_t n) {
    const size_t w = wpr_of(n);
    const size_t STRIPE = 64;                 // words per column stripe
    const uint32_t kgroups = (n + 63) / 64;   // 64 rows of B = 8 tables
    std::vector<u64> tab((size_t)8 * 256 * STRIPE);
    for (size_t s0 = 0; s0 < w; s0 += STRIPE) {
        const size_t sw = std::min(STRIPE, w - s0);
        for (uint32_t g = 0; g < kgroups; ++g) {
            // build the 8 tables for B rows [64g, 64g+64), columns of this stripe
#pragma omp parallel for schedule(static)
            for (int t = 0; t < 8; ++t) {
                u64 *T = &tab[(size_t)t * 256 * STRIPE];
                std::memset(T, 0, sizeof(u64) * sw);
                for (int e = 1; e < 256; ++e) {
                    const int low = e & -e;
                    const int bit = __builtin_ctz(low);
                    const uint32_t row = 64 * g + 8 * t + bit;
                    u64 *dst = T + (size_t)e * STRIPE;
                    const u64 *prev = T + (size_t)(e ^ low) * STRIPE;
                    if (row < n) {
                        const u64 *br = B + (size_t)row * w + s0;
                        for (size_t x = 0; x < sw; ++x) dst[x] = prev[x] | br[x];
                    } else {
                        for (size_t x = 0; x < sw; ++x) dst[x] = prev[x];
                    }
                }
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                const u64 a = A[(size_t)i * w + g];
                if (!a) continue;
                u64 *r = R + (size_t)i * w + s0;
                for (int t = 0; t < 8; ++t) {
                    const unsigned key = (unsigned)(a >> (8 * t)) & 255u;
                    if (!key) continue;
                    const u64 *T = &tab[((size_t)t * 256 + key) * STRIPE];
                    for (size_t x = 0; x < sw; ++x) r[x] |= T[x];
                }
            }
        }
    }
}

// C |= A*B (reference M4riMatrix::add_mul semantics; A or B may alias C).
// Returns 1 if C gained at least one bit.
int oracle_addmul_m4rm(u64 *C, const u64 *A, const u64 *B, uint32_t n) {
    const size_t w = wpr_of(n), words = (size_t)n * w;
    std::vector<u64> R(C, C + words);
    m4rm_into(R.data(), A, B, n);
    const int changed = std::memcmp(R.data(), C, words * sizeof(u64)) != 0;
    std::memcpy(C, R.data(), words * sizeof(u64));
    return changed;
}

// C |= A*B by gathering rows of B for every set bit of A (textbook definition
// C[i] = OR_{k : A[i,k]} B[k]).  Alias-safe.  Returns 1 if C changed.
int oracle_addmul_rows(u64 *C, const u64 *A, const u64 *B, uint32_t n) {
    const size_t w = wpr_of(n), words = (size_t)n * w;
    std::vector<u64> R(C, C + words);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        u64 *r = &R[(size_t)i * w];
        const u64 *a = A + (size_t)i * w;
        for (size_t kw = 0; kw < w; ++kw) {
            u64 x = a[kw];
            while (x) {
                const uint32_t k = (uint32_t)(kw * 64 + __builtin_ctzll(x));
                x &= x - 1;
                const u64 *b = B + (size_t)k * w;
                for (size_t j = 0; j < w; ++j) r[j] |= b[j];
            }
        }
    }
    const int changed = std::memcmp(R.data(), C, words * sizeof(u64)) != 0;
    std::memcpy(C, R.data(), words * sizeof(u64));
    return changed;
}

// ---------------------------------------------------------------------------
// Worklist fixpoint (cfpq-cpp/Grammar.h:35-39, 58-69)
// ---------------------------------------------------------------------------

// rules: nrules triples (head, left, right).  mats: nnt row-major matrices,
// already holding the terminal and epsilon bits.  method 0 = M4RM,
// 1 = row gather.  Returns the number of add_mul calls made.
uint64_t oracle_fixpoint_worklist(uint32_t n, uint32_t nnt, const uint32_t *rules,
                                  uint32_t nrules, u64 **mats, int method) {
    std::vector<std::vector<uint32_t>> with_nt(nnt);
    std::unordered_set<uint32_t> todo;
    for (uint32_t r = 0; r < nrules; ++r) {
        with_nt[rules[3 * r + 1]].push_back(r);
        with_nt[rules[3 * r + 2]].push_back(r);
        todo.insert(r);
    }
    uint64_t calls = 0;
    while (!todo.empty()) {
        const uint32_t r = *todo.begin();
        todo.erase(todo.begin());
        const uint32_t c = rules[3 * r], a = rules[3 * r + 1], b = rules[3 * r + 2];
        ++calls;
        const int ch = method == 0 ? oracle_addmul_m4rm(mats[c], mats[a], mats[b], n)
                                   : oracle_addmul_rows(mats[c], mats[a], mats[b], n);
        if (ch)
            for (uint32_t q : with_nt[c]) todo.insert(q);
    }
    return calls;
}

// The same worklist, stopped after `max_seconds` of wall time (checked after
// each add_mul); *calls_done receives the number of add_mul calls made.
// Used to time a bounded sample of the reference algorithm.
double oracle_fixpoint_worklist_budget(uint32_t n, uint32_t nnt, const uint32_t *rules, uint32_t nrules,
                                       u64 **mats, int method, double max_seconds, uint64_t *calls_done) {
    std::vector<std::vector<uint32_t>> with_nt(nnt);
    std::unordered_set<uint32_t> todo;
    for (uint32_t r = 0; r < nrules; ++r) {
        with_nt[rules[3 * r + 1]].push_back(r);
        with_nt[rules[3 * r + 2]].push_back(r);
        todo.insert(r);
    }
    uint64_t calls = 0;
    const double t0 = omp_get_wtime();
    double el = 0;
    while (!todo.empty()) {
        const uint32_t r = *todo.begin();
        todo.erase(todo.begin());
        const uint32_t c = rules[3 * r], a = rules[3 * r + 1], b = rules[3 * r + 2];
        ++calls;
        const int ch = method == 0 ? oracle_addmul_m4rm(mats[c], mats[a], mats[b], n)
                                   : oracle_addmul_rows(mats[c], mats[a], mats[b], n);
        if (ch)
            for (uint32_t q : with_nt[c]) todo.insert(q);
        el = omp_get_wtime() - t0;
        if (el >= max_seconds) break;
    }
    *calls_done = calls;
    return el;
}

// ---------------------------------------------------------------------------
// Semi-naive fixpoint (independent checker, sparse-friendly)
// ---------------------------------------------------------------------------

// Per iteration, for every rule H -> A B:  new_H |= (dA * A_B) | (M_A * dB),
// masked by ~M_H;  then M |= new, d := new.  Stops when no bit is new.
// Returns the number of iterations (including the final empty one).
// Semi-naive fixpoint (Jacobi steps) whose products are the dense M4RM of
// the worklist oracle: for every rule, new_H |= (dA*M_B | M_A*dB) & ~M_H; then
// M |= new, d := new; stops at the first step with no new bit.  For workloads
// whose matrices saturate, where the row-gather checker is slow.  Returns the
// number of steps; *products receives the number of M4RM products evaluated.
uint32_t oracle_fixpoint_seminaive_m4rm(uint32_t n, uint32_t nnt, const uint32_t *rules, uint32_t nrules,
                                        u64 **mats, uint64_t *products) {
    const size_t w = wpr_of(n), words = (size_t)n * w;
    std::vector<std::vector<u64>> D(nnt), NX(nnt);
    std::vector<u64> T(words);
    std::vector<uint8_t> dnz(nnt, 0);
    for (uint32_t x = 0; x < nnt; ++x) {
        D[x].assign(mats[x], mats[x] + words);
        NX[x].assign(words, 0);
        for (size_t j = 0; j < words && !dnz[x]; ++j) dnz[x] = D[x][j] != 0;
    }
    uint64_t prods = 0;
    uint32_t iters = 0;
    for (;;) {
        ++iters;
        for (uint32_t r = 0; r < nrules; ++r) {
            const uint32_t h = rules[3 * r], a = rules[3 * r + 1], b = rules[3 * r + 2];
            for (int term = 0; term < 2; ++term) {
                if (!dnz[term == 0 ? a : b]) continue;
                const u64 *L = term == 0 ? D[a].data() : mats[a];
                const u64 *R = term == 0 ? mats[b] : D[b].data();
                std::fill(T.begin(), T.end(), 0);
                m4rm_into(T.data(), L, R, n);
                ++prods;
                const u64 *mh = mats[h];
                u64 *nh = NX[h].data();
#pragma omp parallel for schedule(static)
                for (int64_t j = 0; j < (int64_t)words; ++j) nh[j] |= T[j] & ~mh[j];
            }
        }
        bool any = false;
        for (uint32_t x = 0; x < nnt; ++x) {
            u64 *m = mats[x];
            u64 *nx = NX[x].data();
            u64 *d = D[x].data();
            uint8_t nz = 0;
            for (size_t j = 0; j < words; ++j) {
                m[j] |= nx[j];
                d[j] = nx[j];
                nz |= nx[j] != 0;
                nx[j] = 0;
            }
            dnz[x] = nz;
            any = any || nz;
        }
        if (!any) break;
    }
    if (products) *products = prods;
    return iters;
}

uint32_t oracle_fixpoint_seminaive(uint32_t n, uint32_t nnt, const uint32_t *rules,
                                   uint32_t nrules, u64 **mats) {
    const size_t w = wpr_of(n), words = (size_t)n * w;
    std::vector<std::vector<u64>> D(nnt), NX(nnt);
    std::vector<std::vector<uint8_t>> drow(nnt);
    for (uint32_t x = 0; x < nnt; ++x) {
        D[x].assign(mats[x], mats[x] + words);
        NX[x].assign(words, 0);
        drow[x].assign(n, 0);
    }
    auto mark_rows = [&](uint32_t x) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            const u64 *d = &D[x][(size_t)i * w];
            uint8_t nz = 0;
            for (size_t j = 0; j < w; ++j) nz |= d[j] != 0;
            drow[x][i] = nz;
        }
    };
    for (uint32_t x = 0; x < nnt; ++x) mark_rows(x);
    uint32_t iters = 0;
    std::vector<u64> dmask(w);
    for (;;) {
        ++iters;
        for (uint32_t r = 0; r < nrules; ++r) {
            const uint32_t h = rules[3 * r], a = rules[3 * r + 1], b = rules[3 * r + 2];
            const u64 *MA = mats[a], *MB = mats[b], *MH = mats[h];
            const u64 *DA = D[a].data(), *DB = D[b].data();
            u64 *NH = NX[h].data();
            // mask of non-empty rows of dB
            std::fill(dmask.begin(), dmask.end(), 0);
            bool any_db = false;
            for (uint32_t k = 0; k < n; ++k)
                if (drow[b][k]) { dmask[k >> 6] |= 1ull << (k & 63); any_db = true; }
#pragma omp parallel for schedule(dynamic, 64)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                u64 *nh = NH + (size_t)i * w;
                const u64 *mh = MH + (size_t)i * w;
                bool touched = false;
                std::vector<u64> acc;
                if (drow[a][i]) {  // dA row i times M_B
                    acc.assign(w, 0);
                    touched = true;
                    const u64 *da = DA + (size_t)i * w;
                    for (size_t kw = 0; kw < w; ++kw) {
                        u64 x = da[kw];
                        while (x) {
                            const size_t k = kw * 64 + __builtin_ctzll(x);
                            x &= x - 1;
                            const u64 *br = MB + k * w;
                            for (size_t j = 0; j < w; ++j) acc[j] |= br[j];
                        }
                    }
                }
                if (any_db) {  // M_A row i times dB
                    const u64 *ma = MA + (size_t)i * w;
                    for (size_t kw = 0; kw < w; ++kw) {
                        u64 x = ma[kw] & dmask[kw];
                        while (x) {
                            const size_t k = kw * 64 + __builtin_ctzll(x);
                            x &= x - 1;
                            if (!touched) { acc.assign(w, 0); touched = true; }
                            const u64 *dr = DB + k * w;
                            for (size_t j = 0; j < w; ++j) acc[j] |= dr[j];
                        }
                    }
                }
                if (touched)
                    for (size_t j = 0; j < w; ++j) nh[j] |= acc[j] & ~mh[j];
            }
        }
        uint64_t added = 0;
        for (uint32_t x = 0; x < nnt; ++x) {
            u64 *M = mats[x];
            u64 *N = NX[x].data();
            u64 *Dx = D[x].data();
#pragma omp parallel for reduction(+ : added) schedule(static)
            for (int64_t q = 0; q < (int64_t)words; ++q) {
                const u64 nw = N[q] & ~M[q];
                added += (uint64_t)__builtin_popcountll(nw);
                M[q] |= nw;
                Dx[q] = nw;
                N[q] = 0;
            }
            mark_rows(x);
        }
        if (!added) break;
    }
    return iters;
}

// ---------------------------------------------------------------------------
// Hellings' algorithm (worklist over edge facts; independent of matrix code)
// ---------------------------------------------------------------------------

// facts: in/out n*n bitsets per nonterminal (row-major, wpr words per row),
// seeded with terminal and epsilon facts.  Completes them to the least
// fixpoint of "H(i,k) <- L(i,j), R(j,k) for every rule H -> L R".
// Returns the number of facts derived in total.
uint64_t oracle_hellings(uint32_t n, uint32_t nnt, const uint32_t *rules, uint32_t nrules,
                         u64 **facts) {
    const size_t w = wpr_of(n);
    // column copies (transposes) so that "all k with X(k,i)" is a row scan
    std::vector<std::vector<u64>> colT(nnt, std::vector<u64>((size_t)n * w, 0));
    auto get = [&](uint32_t x, uint32_t i, uint32_t j) {
        return (facts[x][(size_t)i * w + (j >> 6)] >> (j & 63)) & 1ull;
    };
    struct F { uint32_t x, i, j; };
    std::deque<F> q;
    uint64_t total = 0;
    for (uint32_t x = 0; x < nnt; ++x)
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = 0; j < n; ++j)
                if (get(x, i, j)) {
                    colT[x][(size_t)j * w + (i >> 6)] |= 1ull << (i & 63);
                    q.push_back({x, i, j});
                    ++total;
                }
    std::vector<std::vector<uint32_t>> as_left(nnt), as_right(nnt);
    for (uint32_t r = 0; r < nrules; ++r) {
        as_left[rules[3 * r + 1]].push_back(r);
        as_right[rules[3 * r + 2]].push_back(r);
    }
    std::vector<F> add;
    while (!q.empty()) {
        const F f = q.front();
        q.pop_front();
        add.clear();
        for (uint32_t r : as_left[f.x]) {   // H -> f.x R : H(i,k) for R(j,k)
            const uint32_t h = rules[3 * r], rr = rules[3 * r + 2];
            const u64 *row = facts[rr] + (size_t)f.j * w;
            for (size_t kw = 0; kw < w; ++kw) {
                u64 m = row[kw];
                while (m) {
                    const uint32_t k = (uint32_t)(kw * 64 + __builtin_ctzll(m));
                    m &= m - 1;
                    add.push_back({h, f.i, k});
                }
            }
        }
        for (uint32_t r : as_right[f.x]) {  // H -> L f.x : H(k,j) for L(k,i)
            const uint32_t h = rules[3 * r], l = rules[3 * r + 1];
            const u64 *col = colT[l].data() + (size_t)f.i * w;
            for (size_t kw = 0; kw < w; ++kw) {
                u64 m = col[kw];
                while (m) {
                    const uint32_t k = (uint32_t)(kw * 64 + __builtin_ctzll(m));
                    m &= m - 1;
                    add.push_back({h, k, f.j});
                }
            }
        }
        for (const F &g : add) {
            if (get(g.x, g.i, g.j)) continue;
            facts[g.x][(size_t)g.i * w + (g.j >> 6)] |= 1ull << (g.j & 63);
            colT[g.x][(size_t)g.j * w + (g.i >> 6)] |= 1ull << (g.i & 63);
            q.push_back(g);
            ++total;
        }
    }
    return total;
}

// ---------------------------------------------------------------------------
// Fingerprint (cfpq-cpp/Grammar.cpp:81-101)
// ---------------------------------------------------------------------------

static const u64 HASH_MOD = 1000000007ull;

// Literal restatement: per row, line = line*11 (+1 if bit) mod p over columns;
// hash = hash*13 + line mod p over rows; count = number of set bits.
void oracle_fingerprint_literal(const u64 *M, uint32_t n, u64 *hash, u64 *count) {
    const size_t w = wpr_of(n);
    u64 h = 0, c = 0;
    for (uint32_t r = 0; r < n; ++r) {
        u64 line = 0;
        for (uint32_t col = 0; col < n; ++col) {
            line *= 11;
            if ((M[(size_t)r * w + (col >> 6)] >> (col & 63)) & 1ull) { ++c; ++line; }
            line %= HASH_MOD;
        }
        h = (h * 13 + line) % HASH_MOD;
    }
    *hash = h;
    *count = c;
}

// Same value from set bits only: line = sum 11^(n-1-col), hash = sum line_r * 13^(n-1-r).
void oracle_fingerprint(const u64 *M, uint32_t n, u64 *hash, u64 *count) {
    const size_t w = wpr_of(n);
    std::vector<u64> p11(n), p13(n);
    u64 a = 1, b = 1;
    for (uint32_t e = 0; e < n; ++e) {
        p11[e] = a; p13[e] = b;
        a = a * 11 % HASH_MOD; b = b * 13 % HASH_MOD;
    }
    u64 h = 0, c = 0;
#pragma omp parallel for reduction(+ : c) schedule(dynamic, 64)
    for (int64_t r = 0; r < (int64_t)n; ++r) {
        u64 line = 0;
        for (size_t kw = 0; kw < w; ++kw) {
            u64 m = M[(size_t)r * w + kw];
            c += (u64)__builtin_popcountll(m);
            while (m) {
                const uint32_t col = (uint32_t)(kw * 64 + __builtin_ctzll(m));
                m &= m - 1;
                line += p11[n - 1 - col];
            }
            line %= HASH_MOD;
        }
        const u64 term = line % HASH_MOD * p13[n - 1 - r] % HASH_MOD;
#pragma omp critical
        h = (h + term) % HASH_MOD;
    }
    *hash = h;
    *count = c;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Threads of the following OpenMP regions (the baseline's 1-thread figure).
void oracle_set_num_threads(int k) {
#ifdef _OPENMP
    omp_set_num_threads(k > 0 ? k : 1);
#else
    (void)k;
#endif
}

}  // extern "C"
