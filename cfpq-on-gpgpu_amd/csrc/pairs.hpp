// Sparse pair-list formulation of the fixpoint, for graphs too large or too
// sparse for n x n bit tiles: each nonterminal's relation is a sorted array of
// 64-bit keys i << s | j (s = bits of the vertex count, so key order is row-
// major order), kept twice, by row (M) and by column (M^T, keys j << s | i).
// A step evaluates, for every rule H -> L R,
//     cand_H = dL |><| M_R  (join on L's column = R's row)
//            | M_L |><| dR  (join on L's column = R's row, through M_L^T)
// as an expansion of every delta pair into the run of the other operand that
// shares its join vertex, then sorts the candidates of each head, drops
// duplicates and pairs already in M_H, and merges what is left into M_H and
// M_H^T.  Sorting, scanning and selection use hipCUB; the joins, the
// membership filter and the merge are the kernels below.  The reference's
// sparse backends are GraphBLAS's GrB_mxm over the LOR semiring
// (graphBlas/graphblas_matrix.cpp:19-31) and scipy CSR products on the CPU
// (python/matmul.py:18-28); both rebuild the head matrix on every product, as
// the merge does here.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cstdint>

namespace cfpq {

constexpr uint32_t PAIRS_THREADS = 256;

// Up to PAIRS_GROUP heads share one sort per step: a key then carries its
// head's index within the group above its 2s key bits (the tag).
constexpr int PAIRS_GROUP = 32;
struct PairHeads {
    const uint64_t *m[PAIRS_GROUP];  // each head's sorted pairs
    uint64_t n[PAIRS_GROUP];
};

#define PAIRS_FOR(i_, total_)                                                             \
    for (uint64_t i_ = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i_ < (total_); \
         i_ += (uint64_t)gridDim.x * blockDim.x)

// First index of a sorted array whose element is >= key.
__device__ __forceinline__ uint64_t pairs_lower_bound(const uint64_t *__restrict__ a, uint64_t n, uint64_t key) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// ABI keys (row << 32 | col) to internal keys (row << s | col), and back.
__global__ void k_pairs_pack(uint64_t *__restrict__ k, uint64_t m, uint32_t s) {
    PAIRS_FOR(e, m) k[e] = (k[e] >> 32) << s | (k[e] & 0xffffffffull);
}
__global__ void k_pairs_unpack(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t m, uint32_t s) {
    const uint64_t mask = (1ull << s) - 1;
    PAIRS_FOR(e, m) out[e] = (in[e] >> s) << 32 | (in[e] & mask);
}

// (i, j) -> tag | (j, i)
__global__ void k_pairs_transpose(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t m, uint32_t s,
                                  uint64_t tag) {
    const uint64_t mask = (1ull << s) - 1;
    PAIRS_FOR(e, m) out[e] = tag | (in[e] & mask) << s | in[e] >> s;
}

// One join of a step: the delta pairs of one operand against the other
// operand's rows.  Every part of a step is evaluated by the same launches;
// its cnt/start/incl entries sit at [at, at + m) of the step's arrays.
struct PairPart {
    const uint64_t *delta;  // side 0: dL (i, k); side 1: dR (k, j)
    const uint64_t *op;     // side 0: M_R by row; side 1: M_L by column
    uint64_t m, nop;
    uint64_t at;
    uint64_t tag;  // head tag of its outputs
    uint32_t side, pad;
};

// Tagged keys back to plain ones.
__global__ void k_pairs_untag(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t m,
                              uint32_t key_bits) {
    const uint64_t mask = (1ull << key_bits) - 1;
    PAIRS_FOR(e, m) out[e] = in[e] & mask;
}

// For every delta pair e of a part, the run of the operand holding its join
// vertex k:
//   side 0, delta of the left operand (i, k): row k of the right operand M_R;
//   side 1, delta of the right operand (k, j): column k of the left operand,
//           i.e. row k of M_L^T.
// cnt = length of the run, start = its first index in the operand.  Every
// part of a step in one launch: block row blockIdx.y is part y.
__global__ void k_pairs_count_parts(const PairPart *__restrict__ parts, uint32_t s, uint64_t *__restrict__ cnt,
                                    uint64_t *__restrict__ start) {
    const PairPart P = parts[blockIdx.y];
    const uint64_t mask = (1ull << s) - 1;
    PAIRS_FOR(e, P.m) {
        const uint64_t key = P.delta[e];
        const uint64_t k = P.side ? key >> s : key & mask;
        const uint64_t lo = pairs_lower_bound(P.op, P.nop, k << s);
        const uint64_t hi = lo + pairs_lower_bound(P.op + lo, P.nop - lo, (k + 1) << s);
        cnt[P.at + e] = hi - lo;
        start[P.at + e] = lo;
    }
}

// totals[i] = join outputs of part i; incl is the inclusive scan of the
// concatenated cnt of every part.
__global__ void k_pairs_part_totals(const PairPart *__restrict__ parts, uint32_t np, const uint64_t *__restrict__ incl,
                                    uint64_t *__restrict__ totals) {
    PAIRS_FOR(i, np) {
        const uint64_t a = parts[i].at, m = parts[i].m;
        totals[i] = incl[a + m - 1] - (a ? incl[a - 1] : 0);
    }
}

// Output p of all of a step's joins, in part order: delta pair e = the first
// with incl[e] > p over the concatenated parts, its part the last one
// starting at or before e, and the operand pair start[e] + (p - incl[e - 1]).
// One thread per output, so a hub vertex's long run is spread over as many
// threads as it produces pairs.
__global__ void k_pairs_expand_parts(const PairPart *__restrict__ parts, uint32_t np,
                                     const uint64_t *__restrict__ incl, const uint64_t *__restrict__ start,
                                     uint64_t sum_m, uint32_t s, uint64_t total, uint64_t *__restrict__ out) {
    const uint64_t mask = (1ull << s) - 1;
    PAIRS_FOR(p, total) {
        uint64_t lo = 0, hi = sum_m;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (incl[mid] <= p)
                lo = mid + 1;
            else
                hi = mid;
        }
        const uint64_t e = lo;
        uint32_t a = 0, b = np;  // parts[a].at <= e < parts[b].at
        while (b - a > 1) {
            const uint32_t mid = (a + b) >> 1;
            if (parts[mid].at <= e)
                a = mid;
            else
                b = mid;
        }
        const PairPart &P = parts[a];
        const uint64_t before = e ? incl[e - 1] : 0;
        const uint64_t o = P.op[start[e] + (p - before)];
        const uint64_t d = P.delta[e - P.at];
        out[p] = P.tag | (P.side ? (o & mask) << s | (d & mask) : (d & ~mask) | (o & mask));
    }
}

// flag[p] = c[p] is the first of its run in the sorted candidates and is not
// in the sorted set t.
__global__ void k_pairs_flag_new(const uint64_t *__restrict__ c, uint64_t nc, const uint64_t *__restrict__ t,
                                 uint64_t nt, uint8_t *__restrict__ flag) {
    PAIRS_FOR(p, nc) {
        const uint64_t x = c[p];
        bool keep = p == 0 || c[p - 1] != x;
        if (keep && nt) {
            const uint64_t at = pairs_lower_bound(t, nt, x);
            keep = at == nt || t[at] != x;
        }
        flag[p] = keep ? 1 : 0;
    }
}

// The same over tagged candidates: the set a candidate is tested against is
// its head's, named by the tag.
__global__ void k_pairs_flag_new_group(const uint64_t *__restrict__ c, uint64_t nc, PairHeads heads,
                                       uint32_t key_bits, uint8_t *__restrict__ flag) {
    const uint64_t mask = (1ull << key_bits) - 1;
    PAIRS_FOR(p, nc) {
        const uint64_t x = c[p];
        bool keep = p == 0 || c[p - 1] != x;
        if (keep) {
            const uint32_t g = (uint32_t)(x >> key_bits);
            const uint64_t *t = heads.m[g];
            const uint64_t nt = heads.n[g], key = x & mask;
            if (nt) {
                const uint64_t at = pairs_lower_bound(t, nt, key);
                keep = at == nt || t[at] != key;
            }
        }
        flag[p] = keep ? 1 : 0;
    }
}

// bounds[g] = first index of tag g in the sorted tagged keys k[0, *count) for
// g < groups, bounds[groups] = *count.  One block of at least groups + 1 threads.
__global__ void k_pairs_group_bounds(const uint64_t *__restrict__ k, const uint64_t *__restrict__ count,
                                     uint32_t key_bits, uint32_t groups, uint64_t *__restrict__ bounds) {
    const uint32_t g = threadIdx.x;
    if (g > groups) return;
    const uint64_t m = *count;
    bounds[g] = g == groups ? m : pairs_lower_bound(k, m, (uint64_t)g << key_bits);
}

// Merge of two sorted arrays with no key in common: an element's place in the
// output is its index in its own array plus the number of smaller keys in the
// other one.
__global__ void k_pairs_merge(const uint64_t *__restrict__ a, uint64_t na, const uint64_t *__restrict__ b, uint64_t nb,
                              uint64_t *__restrict__ out) {
    PAIRS_FOR(q, na + nb) {
        if (q < na)
            out[q + pairs_lower_bound(b, nb, a[q])] = a[q];
        else
            out[(q - na) + pairs_lower_bound(a, na, b[q - na])] = b[q - na];
    }
}

#undef PAIRS_FOR

}  // namespace cfpq
