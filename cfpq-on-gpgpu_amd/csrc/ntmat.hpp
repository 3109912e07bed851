// Single-matrix formulation of the fixpoint: one n x n matrix whose element
// (i, j) is the set of nonterminals (bit x = nonterminal x, up to 32) that
// derive a path i -> j, closed under the binary rules.  The product over the
// grammar semiring is
//     (A (x) B)[i][j] = OR_k OR_{rules H -> X Y : X in A[i][k], Y in B[k][j]} {H},
// the element operation of cuda/cutlass/CutlassMatrix.cu (grammar in constant
// memory, one GEMM covers every rule), evaluated here semi-naively:
//     step 0:  N = (T (x) T) \ T
//     step s:  N = (D (x) T  |  T (x) D) \ T,   T |= N,  D := N
// until N is empty.  One n^2 array of uint32 per buffer (T, D, N), row-major.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cfpq {

constexpr int NT_BLOCK = 64;   // output block of a work-group: 64 x 64 elements
constexpr int NT_K = 32;       // k-chunk staged in LDS per operand
constexpr int NT_MAX_RULES = 256;
constexpr int NT_THREADS = 256;

// rule r packed as head | left << 8 | right << 16
struct NtRules {
    const uint32_t *packed;
    uint32_t count;
};

// Sets derived from one pair of operand elements a (left) and b (right).
__device__ __forceinline__ uint32_t nt_pair(uint32_t a, uint32_t b, const uint32_t *rules, uint32_t nr) {
    uint32_t out = 0;
    for (uint32_t r = 0; r < nr; ++r) {
        const uint32_t x = rules[r];
        const uint32_t hit = (a >> ((x >> 8) & 31u)) & (b >> ((x >> 16) & 31u)) & 1u;
        out |= hit << (x & 31u);
    }
    return out;
}

// N[i][j] = (product over k) & ~T[i][j] for one 64 x 64 output block per
// work-group; thread (ty, tx) of a 16 x 16 layout owns rows ty + 16 a and
// columns tx + 16 b (a, b < 4).  `first`: T (x) T; else D (x) T | T (x) D.
// Counts the non-empty new elements into *count.
__global__ __launch_bounds__(NT_THREADS) void k_nt_product(const uint32_t *__restrict__ T,
                                                           const uint32_t *__restrict__ D,
                                                           uint32_t *__restrict__ N, uint32_t n, NtRules rules,
                                                           int first, unsigned long long *__restrict__ count) {
    __shared__ uint32_t s_rules[NT_MAX_RULES];
    __shared__ uint32_t s_ta[NT_BLOCK][NT_K + 1], s_da[NT_BLOCK][NT_K + 1];  // A rows of T and D
    __shared__ uint32_t s_tb[NT_K][NT_BLOCK + 1], s_db[NT_K][NT_BLOCK + 1];  // B columns of T and D
    const uint32_t tid = threadIdx.x, tx = tid & 15u, ty = tid >> 4;
    const uint32_t i0 = blockIdx.y * NT_BLOCK, j0 = blockIdx.x * NT_BLOCK;
    const uint32_t nr = rules.count;
    for (uint32_t r = tid; r < nr; r += NT_THREADS) s_rules[r] = rules.packed[r];
    uint32_t acc[4][4] = {};
    for (uint32_t k0 = 0; k0 < n; k0 += NT_K) {
        __syncthreads();
        // stage the 64 x 32 A blocks and the 32 x 64 B blocks (zero outside the matrix)
        for (uint32_t e = tid; e < NT_BLOCK * NT_K; e += NT_THREADS) {
            const uint32_t r = e / NT_K, c = e % NT_K;
            const uint32_t gi = i0 + r, gk = k0 + c;
            const bool in = gi < n && gk < n;
            const size_t off = (size_t)gi * n + gk;
            s_ta[r][c] = in ? T[off] : 0u;
            s_da[r][c] = in && !first ? D[off] : 0u;
            const uint32_t kr = e / NT_BLOCK, jc = e % NT_BLOCK;
            const uint32_t gk2 = k0 + kr, gj = j0 + jc;
            const bool in2 = gk2 < n && gj < n;
            const size_t off2 = (size_t)gk2 * n + gj;
            s_tb[kr][jc] = in2 ? T[off2] : 0u;
            s_db[kr][jc] = in2 && !first ? D[off2] : 0u;
        }
        __syncthreads();
        for (uint32_t kk = 0; kk < NT_K; ++kk) {
            uint32_t ta[4], da[4], tb[4], db[4];
            uint32_t any_a = 0, any_b = 0;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                ta[a] = s_ta[ty + 16 * a][kk];
                da[a] = s_da[ty + 16 * a][kk];
                tb[a] = s_tb[kk][tx + 16 * a];
                db[a] = s_db[kk][tx + 16 * a];
                any_a |= ta[a] | da[a];
                any_b |= tb[a] | db[a];
            }
            if (!any_a || !any_b) continue;  // sparse matrices: most k give no pair
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (first) {
                        if (ta[a] && tb[b]) acc[a][b] |= nt_pair(ta[a], tb[b], s_rules, nr);
                    } else {
                        if (da[a] && tb[b]) acc[a][b] |= nt_pair(da[a], tb[b], s_rules, nr);
                        if (ta[a] && db[b]) acc[a][b] |= nt_pair(ta[a], db[b], s_rules, nr);
                    }
                }
        }
    }
    uint32_t nonzero = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t gi = i0 + ty + 16 * a, gj = j0 + tx + 16 * b;
            if (gi < n && gj < n) {
                const size_t off = (size_t)gi * n + gj;
                const uint32_t nw = acc[a][b] & ~T[off];
                N[off] = nw;
                nonzero += nw != 0;
            }
        }
    // one atomic per wave
    for (int o = 32; o > 0; o >>= 1) nonzero += __shfl_xor(nonzero, o, 64);
    if ((tid & 63u) == 0 && nonzero) atomicAdd(count, (unsigned long long)nonzero);
}

// T |= N over n^2 elements (grid-strided).
__global__ void k_nt_apply(uint32_t *__restrict__ T, const uint32_t *__restrict__ N, size_t total) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x)
        T[x] |= N[x];
}

}  // namespace cfpq
