// gfx950 kernels of the delta fixpoint over tiled bit-packed matrices.
//
// One step of the fixpoint (all rules at once, Jacobi style) is
//   k_begin    zero this step's counters
//   k_build    list the (head, C tile) items whose product can be non-empty
//   k_product  persistent: for each item, OR the tile-products of every term
//              into registers, keep only bits not yet in the head, write them
//              to the head's "next delta" buffer
//   k_apply    persistent: M[head] |= next delta, tile metadata updated
//   k_clear    persistent: zero the tiles of the delta consumed by this step
// Work groups are 256 threads; thread r owns row r of the current C tile.
#pragma once
#include <hip/hip_runtime.h>
#include "tiles.hpp"

namespace cfpq {

struct alignas(32) Row4 {
    uint64_t w[4];
};

__device__ __forceinline__ Row4 load_row(const uint64_t *p) {
    return *reinterpret_cast<const Row4 *>(p);
}
__device__ __forceinline__ void store_row(uint64_t *p, const Row4 &v) {
    *reinterpret_cast<Row4 *>(p) = v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Work-group-wide dynamic work cursor for the persistent kernels: thread 0
// takes the next index and broadcasts it through LDS.  The barriers on both
// sides are load-bearing.  They keep every divergent block of the caller's
// loop body apart from the "thread 0" block here; without them the compiler
// may fuse an "r == 0" block at the end of the body with the "r == 0" fetch
// of the next iteration, turning the loop into a divergent inner loop in
// which 255 threads re-read a stale index while thread 0 waits outside it --
// a livelock on SIMT hardware.
__device__ __forceinline__ uint32_t wg_next(uint32_t *cursor, uint32_t *s_idx) {
    __syncthreads();
    if (threadIdx.x == 0) *s_idx = atomicAdd(cursor, 1u);
    __syncthreads();
    return *s_idx;
}

// ---------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_begin(StepCounters *ring, uint32_t slot) {
    if (threadIdx.x == 0) {
        StepCounters z = {};
        ring[slot] = z;
    }
}

// Item = (head << 40) | tile index.  A C tile is listed when some term has
// a k with both operand tiles non-empty and the head tile is not already full.
__global__ __launch_bounds__(BLOCK) void k_build(const MatDev *__restrict__ mats,
                                                 const HeadDev *__restrict__ heads,
                                                 const TermDev *__restrict__ terms, int parity,
                                                 StepCounters *__restrict__ c,
                                                 uint64_t *__restrict__ items) {
    const uint32_t h = blockIdx.y;
    const HeadDev hd = heads[h];
    const uint32_t T = hd.T;
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    bool want = false;
    if (p < T * T) {
        const uint32_t I = p / T, J = p - I * T;
        if (mats[hd.target].cnt[p] < TILE_BITS) {
            for (uint32_t t = hd.t0; t < hd.t1 && !want; ++t) {
                const MatDev L = mats[terms[t].left[parity]];
                const MatDev R = mats[terms[t].right[parity]];
                for (uint32_t w = 0; w < L.ow; ++w)
                    if (L.occR[I * L.ow + w] & R.occC[J * R.ow + w]) {
                        want = true;
                        break;
                    }
            }
        }
    }
    const uint64_t bal = __ballot(want);
    if (!bal) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&c->n_items, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    if (want) items[base + __popcll(bal & ((1ull << lane) - 1ull))] = ((uint64_t)h << 40) | p;
}

// ---------------------------------------------------------------------------
// tile-products:  acc (row r of the C tile) |= A tile row r * B tile
// ---------------------------------------------------------------------------

// Sparse A tile: gather the B rows named by A's set bits straight from HBM/L2.
__device__ __forceinline__ void tp_gather(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                          uint64_t rm, int r, Row4 &acc, uint32_t &arows, uint32_t &brows) {
    if (!((rm >> (r & 63)) & 1ull)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint64_t x = a.w[w];
        while (x) {
            const int k = w * 64 + __builtin_ctzll(x);
            x &= x - 1;
            const Row4 b = load_row(Bt + k * TROW_WORDS);
            acc.w[0] |= b.w[0];
            acc.w[1] |= b.w[1];
            acc.w[2] |= b.w[2];
            acc.w[3] |= b.w[3];
            ++brows;
        }
    }
}

// Dense A tile: stage the B tile in LDS (8 KiB, one coalesced row per
// thread), then OR rows from LDS.  Must be reached by the whole work-group.
__device__ __forceinline__ void tp_lds(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                       uint64_t rm, int r, Row4 *Bs, Row4 &acc, uint32_t &arows,
                                       uint32_t &brows_lds) {
    __syncthreads();
    Bs[r] = load_row(Bt + r * TROW_WORDS);
    __syncthreads();
    if (!((rm >> (r & 63)) & 1ull)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint64_t x = a.w[w];
        while (x) {
            const int k = w * 64 + __builtin_ctzll(x);
            x &= x - 1;
            const Row4 b = Bs[k];
            acc.w[0] |= b.w[0];
            acc.w[1] |= b.w[1];
            acc.w[2] |= b.w[2];
            acc.w[3] |= b.w[3];
            ++brows_lds;
        }
    }
}

// Very dense A tile: "method of four Russians" with 4-bit keys.  The B tile
// is staged in LDS, then all 64 tables of 16 OR-combinations (B rows
// 4g..4g+3, 32 KiB) are built once; each row of A then costs one table
// lookup per non-zero nibble instead of one row-OR per set bit.
// Must be reached by the whole work-group.
constexpr int M4R_GROUPS = TILE / 4;  // 64 tables of 16 entries
__device__ __forceinline__ void tp_m4r(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                       uint64_t rm, int r, Row4 *Bs, Row4 *Tab, Row4 &acc, uint32_t &arows,
                                       uint32_t &lookups) {
    __syncthreads();
    Bs[r] = load_row(Bt + r * TROW_WORDS);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (M4R_GROUPS * 16) / BLOCK; ++q) {
        const int e = q * BLOCK + r;     // entry index = g * 16 + key
        const int g = e >> 4, key = e & 15;
        Row4 v = {{0, 0, 0, 0}};
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((key >> b) & 1) {
                const Row4 x = Bs[4 * g + b];
                v.w[0] |= x.w[0];
                v.w[1] |= x.w[1];
                v.w[2] |= x.w[2];
                v.w[3] |= x.w[3];
            }
        Tab[e] = v;
    }
    __syncthreads();
    if (!((rm >> (r & 63)) & 1ull)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint64_t x = a.w[w];
        while (x) {
            const int nib = __builtin_ctzll(x) >> 2;   // lowest non-zero nibble
            const int key = (int)((x >> (4 * nib)) & 15u);
            x &= ~(15ull << (4 * nib));
            const Row4 t = Tab[(w * 16 + nib) * 16 + key];
            acc.w[0] |= t.w[0];
            acc.w[1] |= t.w[1];
            acc.w[2] |= t.w[2];
            acc.w[3] |= t.w[3];
            ++lookups;
        }
    }
}

// ---------------------------------------------------------------------------
// product kernel (persistent, dynamic item cursor)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_product(const MatDev *__restrict__ mats,
                                                   const HeadDev *__restrict__ heads,
                                                   const TermDev *__restrict__ terms, int parity,
                                                   StepCounters *__restrict__ c,
                                                   const uint64_t *__restrict__ items,
                                                   uint64_t *__restrict__ touched, uint32_t dense_bits,
                                                   uint32_t m4r_bits) {
    __shared__ Row4 Bs[TILE];
    __shared__ Row4 Tab[M4R_GROUPS * 16];
    __shared__ uint32_t s_idx;
    __shared__ uint32_t s_sum[4];
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const uint32_t n_items = c->n_items;
    uint32_t tp = 0, tpd = 0, arows = 0, brows = 0, brows_lds = 0, out_rows = 0, lookups = 0;
    for (uint32_t idx = wg_next(&c->item_next, &s_idx); idx < n_items; idx = wg_next(&c->item_next, &s_idx)) {
        const uint64_t it = items[idx];
        const uint32_t h = uniform((uint32_t)(it >> 40));
        const uint32_t p = uniform((uint32_t)(it & 0xffffffffffull));
        const HeadDev hd = heads[h];
        const uint32_t T = hd.T, I = p / T, J = p - (p / T) * T;
        Row4 acc = {{0, 0, 0, 0}};
        for (uint32_t t = hd.t0; t < hd.t1; ++t) {
            const MatDev L = mats[terms[t].left[parity]];
            const MatDev R = mats[terms[t].right[parity]];
            for (uint32_t w = 0; w < L.ow; ++w) {
                uint32_t kb = uniform(L.occR[I * L.ow + w] & R.occC[J * R.ow + w]);
                while (kb) {
                    const uint32_t K = w * 32 + __builtin_ctz(kb);
                    kb &= kb - 1;
                    const uint32_t lt = I * T + K, rt = K * T + J;
                    const uint64_t *At = L.bits + (size_t)lt * TILE_WORDS;
                    const uint64_t *Bt = R.bits + (size_t)rt * TILE_WORDS;
                    const uint64_t rm = L.rmask[(size_t)lt * 4 + wv];
                    ++tp;
                    const uint32_t acnt = uniform(L.cnt[lt]);
                    if (acnt >= m4r_bits) {
                        ++tpd;
                        tp_m4r(At, Bt, rm, r, Bs, Tab, acc, arows, lookups);
                    } else if (acnt >= dense_bits) {
                        ++tpd;
                        tp_lds(At, Bt, rm, r, Bs, acc, arows, brows_lds);
                    } else {
                        tp_gather(At, Bt, rm, r, acc, arows, brows);
                    }
                }
            }
        }
        const MatDev M = mats[hd.target];
        const MatDev O = mats[hd.out[parity]];
        const size_t row_off = (size_t)p * TILE_WORDS + r * TROW_WORDS;
        const Row4 m = load_row(M.bits + row_off);
        Row4 nw;
        nw.w[0] = acc.w[0] & ~m.w[0];
        nw.w[1] = acc.w[1] & ~m.w[1];
        nw.w[2] = acc.w[2] & ~m.w[2];
        nw.w[3] = acc.w[3] & ~m.w[3];
        const bool nz = (nw.w[0] | nw.w[1] | nw.w[2] | nw.w[3]) != 0;
        if (nz) store_row(O.bits + row_off, nw);
        const uint64_t bal = __ballot(nz);
        out_rows += nz;
        const uint32_t pc = wave_sum((uint32_t)(__popcll(nw.w[0]) + __popcll(nw.w[1]) + __popcll(nw.w[2]) +
                                                __popcll(nw.w[3])));
        if (lane == 0) {
            O.rmask[(size_t)p * 4 + wv] = bal;
            s_sum[wv] = pc;
        }
        __syncthreads();
        if (r == 0) {
            const uint32_t total = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
            if (total) {
                O.cnt[p] = total;
                atomicOr(&O.occR[I * O.ow + (J >> 5)], 1u << (J & 31));
                atomicOr(&O.occC[J * O.ow + (I >> 5)], 1u << (I & 31));
                const uint32_t pos = atomicAdd(&c->n_touched, 1u);
                touched[pos] = it;
                atomicAdd((unsigned long long *)&c->new_bits, (unsigned long long)total);
            }
        }
    }
    const uint32_t ar = wave_sum(arows), br = wave_sum(brows), orow = wave_sum(out_rows);
    const uint32_t brl = wave_sum(brows_lds), lk = wave_sum(lookups);
    if (lane == 0) {
        if (lk) atomicAdd((unsigned long long *)&c->m4r_lookups, (unsigned long long)lk);
        if (brl) atomicAdd((unsigned long long *)&c->b_rows_lds, (unsigned long long)brl);
        if (ar) atomicAdd((unsigned long long *)&c->a_rows, (unsigned long long)ar);
        if (br) atomicAdd((unsigned long long *)&c->b_rows, (unsigned long long)br);
        if (orow) atomicAdd((unsigned long long *)&c->out_rows, (unsigned long long)orow);
    }
    if (r == 0 && tp) {
        atomicAdd((unsigned long long *)&c->tile_products, (unsigned long long)tp);
        atomicAdd((unsigned long long *)&c->tp_dense, (unsigned long long)tpd);
    }
}

// M[head] |= next-delta tile, for every tile the product kernel wrote.
__global__ __launch_bounds__(BLOCK) void k_apply(const MatDev *__restrict__ mats,
                                                 const HeadDev *__restrict__ heads, int parity,
                                                 StepCounters *__restrict__ c,
                                                 const uint64_t *__restrict__ touched) {
    __shared__ uint32_t s_idx;
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const uint32_t n = c->n_touched;
    for (uint32_t idx = wg_next(&c->apply_next, &s_idx); idx < n; idx = wg_next(&c->apply_next, &s_idx)) {
        const uint64_t it = touched[idx];
        const uint32_t h = uniform((uint32_t)(it >> 40));
        const uint32_t p = uniform((uint32_t)(it & 0xffffffffffull));
        const HeadDev hd = heads[h];
        const MatDev M = mats[hd.target];
        const MatDev O = mats[hd.out[parity]];
        const uint64_t rm = O.rmask[(size_t)p * 4 + wv];
        if ((rm >> lane) & 1ull) {
            const size_t off = (size_t)p * TILE_WORDS + r * TROW_WORDS;
            Row4 m = load_row(M.bits + off);
            const Row4 o = load_row(O.bits + off);
            m.w[0] |= o.w[0];
            m.w[1] |= o.w[1];
            m.w[2] |= o.w[2];
            m.w[3] |= o.w[3];
            store_row(M.bits + off, m);
        }
        if (lane == 0) M.rmask[(size_t)p * 4 + wv] |= rm;
        if (r == 0) {
            const uint32_t T = hd.T, I = p / T, J = p - (p / T) * T;
            const uint32_t old = M.cnt[p];
            M.cnt[p] = old + O.cnt[p];
            if (!old) {
                atomicOr(&M.occR[I * M.ow + (J >> 5)], 1u << (J & 31));
                atomicOr(&M.occC[J * M.ow + (I >> 5)], 1u << (I & 31));
            }
        }
    }
}

// Zero the delta tiles listed in `list` (written by the step of parity
// `parity`): rows named by rmask, then rmask and cnt.
__global__ __launch_bounds__(BLOCK) void k_clear(const MatDev *__restrict__ mats,
                                                 const HeadDev *__restrict__ heads, int parity,
                                                 const StepCounters *__restrict__ list_ctr,
                                                 StepCounters *__restrict__ c,
                                                 const uint64_t *__restrict__ list) {
    __shared__ uint32_t s_idx;
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const uint32_t n = list_ctr->n_touched;
    for (uint32_t idx = wg_next(&c->clear_next, &s_idx); idx < n; idx = wg_next(&c->clear_next, &s_idx)) {
        const uint64_t it = list[idx];
        const uint32_t h = uniform((uint32_t)(it >> 40));
        const uint32_t p = uniform((uint32_t)(it & 0xffffffffffull));
        const MatDev O = mats[heads[h].out[parity]];
        const uint64_t rm = O.rmask[(size_t)p * 4 + wv];
        if ((rm >> lane) & 1ull) store_row(O.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS, Row4{{0, 0, 0, 0}});
        __syncthreads();
        if (lane == 0) O.rmask[(size_t)p * 4 + wv] = 0;
        if (r == 0) O.cnt[p] = 0;
    }
}

// Zero every non-empty tile and the metadata of matrices ids[0..nid).
__global__ __launch_bounds__(BLOCK) void k_clear_full(const MatDev *__restrict__ mats,
                                                      const uint32_t *__restrict__ ids, uint32_t nid) {
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    for (uint32_t q = blockIdx.y; q < nid; q += gridDim.y) {
        const MatDev O = mats[ids[q]];
        const uint32_t tiles = O.T * O.T;
        for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
            if (uniform(O.cnt[p])) {
                const uint64_t rm = O.rmask[(size_t)p * 4 + wv];
                if ((rm >> lane) & 1ull) store_row(O.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS, Row4{{0, 0, 0, 0}});
                __syncthreads();
                if (lane == 0) O.rmask[(size_t)p * 4 + wv] = 0;
                if (r == 0) O.cnt[p] = 0;
                __syncthreads();  // keeps the divergent tail apart from the next tile (see wg_next)
            }
        }
        const uint32_t ow_words = O.T * O.ow;
        for (uint32_t x = blockIdx.x * BLOCK + r; x < ow_words; x += gridDim.x * BLOCK) {
            O.occR[x] = 0;
            O.occC[x] = 0;
        }
    }
}

// Zero the occupancy bitsets of matrices ids[0..nid).
__global__ void k_clear_occ(const MatDev *__restrict__ mats, const uint32_t *__restrict__ ids, uint32_t nid) {
    for (uint32_t q = blockIdx.y; q < nid; q += gridDim.y) {
        const MatDev O = mats[ids[q]];
        const uint32_t words = O.T * O.ow;
        for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < words; x += gridDim.x * blockDim.x) {
            O.occR[x] = 0;
            O.occC[x] = 0;
        }
    }
}

// dst := src (same T) for the matrices pairs (dst_ids[q], src_ids[q]); only
// non-empty tiles are copied, dst must be all-zero on entry.
__global__ __launch_bounds__(BLOCK) void k_copy(const MatDev *__restrict__ mats, const uint32_t *__restrict__ dst_ids,
                                                const uint32_t *__restrict__ src_ids, uint32_t nid) {
    const int r = threadIdx.x;
    for (uint32_t q = blockIdx.y; q < nid; q += gridDim.y) {
        const MatDev D = mats[dst_ids[q]];
        const MatDev S = mats[src_ids[q]];
        const uint32_t tiles = S.T * S.T;
        for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
            const uint32_t cn = uniform(S.cnt[p]);
            if (!cn) continue;
            const size_t off = (size_t)p * TILE_WORDS + r * TROW_WORDS;
            store_row(D.bits + off, load_row(S.bits + off));
            if (r < 4) D.rmask[(size_t)p * 4 + r] = S.rmask[(size_t)p * 4 + r];
            if (r == 0) D.cnt[p] = cn;
        }
        const uint32_t ow_words = S.T * S.ow;
        for (uint32_t x = blockIdx.x * BLOCK + r; x < ow_words; x += gridDim.x * BLOCK) {
            D.occR[x] = S.occR[x];
            D.occC[x] = S.occC[x];
        }
    }
}

// ---------------------------------------------------------------------------
// layout conversion and metadata
// ---------------------------------------------------------------------------

// Recompute cnt / rmask / occupancy of one tile from its bits (whole WG).
__device__ __forceinline__ void tile_meta(const MatDev &M, uint32_t p, const Row4 &v, uint32_t *s_sum) {
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const bool nz = (v.w[0] | v.w[1] | v.w[2] | v.w[3]) != 0;
    const uint64_t bal = __ballot(nz);
    const uint32_t pc = wave_sum((uint32_t)(__popcll(v.w[0]) + __popcll(v.w[1]) + __popcll(v.w[2]) + __popcll(v.w[3])));
    if (lane == 0) {
        M.rmask[(size_t)p * 4 + wv] = bal;
        s_sum[wv] = pc;
    }
    __syncthreads();
    if (r == 0) {
        const uint32_t total = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
        M.cnt[p] = total;
        if (total) {
            const uint32_t I = p / M.T, J = p - (p / M.T) * M.T;
            atomicOr(&M.occR[I * M.ow + (J >> 5)], 1u << (J & 31));
            atomicOr(&M.occC[J * M.ow + (I >> 5)], 1u << (I & 31));
        }
    }
    __syncthreads();
}

// Row-major host layout (n rows x wpr words) -> tiles, metadata included.
// Occupancy bitsets must be zero on entry.
__global__ __launch_bounds__(BLOCK) void k_rows_to_tiles(MatDev M, const uint64_t *__restrict__ rows, uint32_t n,
                                                         uint32_t wpr) {
    __shared__ uint32_t s_sum[4];
    const int r = threadIdx.x;
    const uint32_t tiles = M.T * M.T;
    for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
        const uint32_t I = p / M.T, J = p - (p / M.T) * M.T;
        const uint32_t i = I * TILE + r;
        Row4 v = {{0, 0, 0, 0}};
        if (i < n) {
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const uint32_t cw = J * 4 + w;
                if (cw < wpr) {
                    uint64_t x = rows[(size_t)i * wpr + cw];
                    const uint32_t c0 = cw * 64;
                    if (c0 >= n) x = 0;
                    else if (n - c0 < 64) x &= (1ull << (n - c0)) - 1ull;
                    v.w[w] = x;
                }
            }
        }
        store_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS, v);
        tile_meta(M, p, v, s_sum);
    }
}

// Tiles -> row-major (n rows x wpr words, wpr = ceil(n/64)).
__global__ __launch_bounds__(BLOCK) void k_tiles_to_rows(MatDev M, uint64_t *__restrict__ rows, uint32_t n,
                                                         uint32_t wpr) {
    const int r = threadIdx.x;
    const uint32_t tiles = M.T * M.T;
    for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
        const uint32_t I = p / M.T, J = p - (p / M.T) * M.T;
        const uint32_t i = I * TILE + r;
        if (i >= n) continue;
        const Row4 v = load_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const uint32_t cw = J * 4 + w;
            if (cw < wpr) rows[(size_t)i * wpr + cw] = v.w[w];
        }
    }
}

__global__ void k_set_bits(MatDev M, const uint32_t *__restrict__ ri, const uint32_t *__restrict__ ci, size_t count) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t i = ri[q], j = ci[q];
        const uint32_t p = (i / TILE) * M.T + (j / TILE);
        const size_t off = (size_t)p * TILE_WORDS + (i % TILE) * TROW_WORDS + ((j % TILE) >> 6);
        atomicOr((unsigned long long *)(M.bits + off), 1ull << (j & 63));
    }
}

// Recompute all tile metadata from the bits (occupancy must be zero on entry).
__global__ __launch_bounds__(BLOCK) void k_recount(MatDev M) {
    __shared__ uint32_t s_sum[4];
    const int r = threadIdx.x;
    const uint32_t tiles = M.T * M.T;
    for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
        const Row4 v = load_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS);
        tile_meta(M, p, v, s_sum);
    }
}

__global__ void k_popcount(MatDev M, unsigned long long *out) {
    const uint32_t tiles = M.T * M.T;
    unsigned long long s = 0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < tiles; p += gridDim.x * blockDim.x) s += M.cnt[p];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

// Fingerprint of cfpq-cpp/Grammar.cpp:81-101 in closed form:
//   line_i = sum_{j : M[i][j]} 11^(n-1-j) mod P,  hash = sum_i line_i * 13^(n-1-i) mod P.
// One thread per row; out[0] += hash terms (reduced mod P on the host), out[1] += popcount.
__global__ __launch_bounds__(BLOCK) void k_fingerprint(MatDev M, uint32_t n, const uint64_t *__restrict__ p11,
                                                       const uint64_t *__restrict__ p13,
                                                       unsigned long long *out) {
    const uint64_t P = 1000000007ull;
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long term = 0, cnt = 0;
    if (i < n) {
        const uint32_t I = i / TILE, r = i % TILE;
        uint64_t line = 0;
        for (uint32_t J = 0; J < M.T; ++J) {
            const uint32_t p = I * M.T + J;
            if (!((M.rmask[(size_t)p * 4 + (r >> 6)] >> (r & 63)) & 1ull)) continue;
            const Row4 v = load_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS);
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                uint64_t x = v.w[w];
                cnt += __popcll(x);
                while (x) {
                    const uint32_t j = J * TILE + w * 64 + __builtin_ctzll(x);
                    x &= x - 1;
                    if (j < n) line += p11[n - 1 - j];
                }
                line %= P;
            }
        }
        term = (line % P) * p13[n - 1 - i] % P;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        term += __shfl_xor(term, o, 64);
        cnt += __shfl_xor(cnt, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], term);
        atomicAdd(&out[1], cnt);
    }
}

}  // namespace cfpq
