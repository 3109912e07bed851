// gfx950 kernels of the delta fixpoint over tiled bit-packed matrices.
//
// One step of the fixpoint (all rules at once, Jacobi style) is
//   k_begin           zero this step's counters
//   k_dense_occ       mark the operand tiles heavy enough for the LDS / MFMA paths
//   k_build           list the work: a (head, C tile) item goes to the MFMA
//                     list if some term has an MFMA-heavy A tile for it, else
//                     to the dense list if some A tile is heavy, otherwise
//                     each non-empty 64-row quarter becomes a sparse unit
//   k_product_sparse  one wavefront per unit, no LDS, no barriers: gathers the
//                     B rows named by the set bits of its A rows
//   k_product_dense   one 256-thread work-group per item: B tiles staged in
//                     LDS, OR-of-rows or Four Russians per k-tile
//   k_product_mfma    one 512-thread work-group per item: exact int8 0/1
//                     matrix-core products with int32 accumulation, > 0
//   k_apply           M[head] |= next delta, tile metadata updated
//   k_clear           zero the delta tiles consumed by this step
// Both product kernels keep only bits not yet in the head and write them to
// the head's "next delta" buffer; the first writer of a tile lists it.
#pragma once
#include <hip/hip_runtime.h>
#include "tiles.hpp"

namespace cfpq {

struct alignas(32) Row4 {
    uint64_t w[4];
};

// Tile rows and metadata live in device memory (global address space).  The
// pointers reach the kernels inside descriptor tables, so the compiler cannot
// tell their address space and would emit flat instructions, which count
// against both the vector-memory and the LDS wait counters: every LDS access
// of a wave (shuffles, the binary searches, the row accumulator) would then
// also wait for its outstanding loads.  These helpers state the global
// address space.  Never pass them an LDS pointer.
#define GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T *p) {
    return *(const GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ void gst(T *p, T v) {
    *(GLOBAL T *)p = v;
}
__device__ __forceinline__ uint32_t g_or(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_or((GLOBAL uint32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t g_or(uint64_t *p, uint64_t v) {
    return __hip_atomic_fetch_or((GLOBAL uint64_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t g_add(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_add((GLOBAL uint32_t *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (through a vector type: a struct copy would be lowered as a flat memcpy)
typedef uint64_t u64x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ Row4 load_row(const uint64_t *p) {
    const u64x4 v = *(const GLOBAL u64x4 *)p;
    return Row4{{v.x, v.y, v.z, v.w}};
}
__device__ __forceinline__ void store_row(uint64_t *p, const Row4 &v) {
    *(GLOBAL u64x4 *)p = u64x4{v.w[0], v.w[1], v.w[2], v.w[3]};
}
__device__ __forceinline__ void or_row(Row4 &a, const Row4 &b) {
    a.w[0] |= b.w[0];
    a.w[1] |= b.w[1];
    a.w[2] |= b.w[2];
    a.w[3] |= b.w[3];
}
__device__ __forceinline__ bool row_nonzero(const Row4 &a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) != 0; }
__device__ __forceinline__ uint32_t row_popc(const Row4 &a) {
    return (uint32_t)(__popcll(a.w[0]) + __popcll(a.w[1]) + __popcll(a.w[2]) + __popcll(a.w[3]));
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | uniform((uint32_t)v);
}

// Wave-uniform read of memory that only the host writes (the plan's
// descriptor tables, uploaded before the kernels that read them): going
// through the constant address space lets the compiler use scalar loads
// through the scalar cache.  Never use it for data that kernels write (work
// lists, counts, row masks, occupancy): scalar loads of such data were seen
// to return values older than an earlier launch's stores (the retire
// launch's reads of the head row masks made the deferral path lose bits
// intermittently until they went through the vector path).  ld_dev is for
// those.
template <class T>
__device__ __forceinline__ T ld_const(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}
// Read of memory that kernels write, always through the vector path (an
// atomic load at work-group scope is a plain vector load that the compiler
// never turns into a scalar one).
template <class T>
__device__ __forceinline__ T ld_dev(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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

// Work-group-uniform copy of a value read from memory that other
// work-groups may be updating during this kernel (the head matrices'
// occupancy and counts, with the head update fused into the products): every
// wave must see the same value, or the waves take different paths and their
// barriers no longer pair up.  Thread 0's read is broadcast through one of two
// LDS slots, alternating with `n` (a per-thread call count), which lets one
// barrier per call suffice: a slot is rewritten only after every thread has
// passed the next call's barrier, i.e. after it has read the slot.
__device__ __forceinline__ uint32_t wg_bcast(uint32_t v, uint32_t *slots, uint32_t &n) {
    uint32_t *slot = slots + (n++ & 1u);
    if (threadIdx.x == 0) *slot = v;
    __syncthreads();
    return *slot;
}

// Wavefront-wide cursor (single-wave work-groups; no barrier involved):
// claims `chunk` consecutive indices per atomic.
__device__ __forceinline__ uint32_t wave_next(uint32_t *cursor, uint32_t chunk) {
    uint32_t v = 0;
    if ((threadIdx.x & 63) == 0) v = atomicAdd(cursor, chunk);
    return uniform(v);
}

// OR of a row over the 64 lanes of the wavefront (all lanes get the result).
__device__ __forceinline__ Row4 wave_or_row(Row4 v) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        uint32_t lo = (uint32_t)v.w[w], hi = (uint32_t)(v.w[w] >> 32);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo |= __shfl_xor(lo, o, 64);
            hi |= __shfl_xor(hi, o, 64);
        }
        v.w[w] = ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// Inclusive prefix sum over the lanes of the wave.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Value of lane l (l wave-uniform), through the scalar unit: no LDS traffic.
__device__ __forceinline__ uint32_t lane_read(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint64_t lane_read64(uint64_t v, uint32_t l) {
    return ((uint64_t)lane_read((uint32_t)(v >> 32), l) << 32) | lane_read((uint32_t)v, l);
}

// Lanes at or below which a ballot is read lane by lane (v_readlane per set
// bit) instead of by log-step shuffles through the LDS crossbar: the output
// rows and filter hits of a sparse unit are usually a handful.
constexpr int FEW_LANES = 8;

// Column union of the rows v of the lanes in `bal` (a ballot), and the sum
// of `cnt` over those lanes.  Whole wave, uniform results.
struct RowsUnion {
    Row4 cu;
    uint32_t sum;
};
__device__ __forceinline__ RowsUnion rows_union(const Row4 &v, uint32_t cnt, uint64_t bal) {
    RowsUnion r = {{{0, 0, 0, 0}}, 0};
    if (__popcll(bal) <= FEW_LANES) {
        for (uint64_t b = bal; b; b &= b - 1) {
            const uint32_t l = (uint32_t)__builtin_ctzll(b);
#pragma unroll
            for (int w = 0; w < 4; ++w) r.cu.w[w] |= lane_read64(v.w[w], l);
            r.sum += lane_read(cnt, l);
        }
    } else {
        r.cu = wave_or_row(v);
        r.sum = wave_sum(cnt);
    }
    return r;
}

// Inclusive prefix sum of v over the lanes, where only lanes in `bal` (a
// ballot) may hold non-zero values; `total` gets the wave's sum.
__device__ __forceinline__ uint32_t scan_ballot(uint32_t v, uint64_t bal, uint32_t &total) {
    const int lane = threadIdx.x & 63;
    if (__popcll(bal) <= FEW_LANES) {
        uint32_t run = 0, incl = 0;
        for (uint64_t b = bal; b; b &= b - 1) {
            const uint32_t l = (uint32_t)__builtin_ctzll(b);
            run += lane_read(v, l);
            if ((uint32_t)lane >= l) incl = run;
        }
        total = run;
        return incl;
    }
    const uint32_t incl = wave_scan_incl(v);
    total = __shfl(incl, 63, 64);
    return incl;
}

__device__ __forceinline__ uint64_t *colq_of(const MatDev &M, uint32_t p, uint32_t q) {
    return M.colq + ((size_t)p * QUARTERS + q) * TROW_WORDS;
}
__device__ __forceinline__ uint64_t *colT_of(const MatDev &M, uint32_t p, uint32_t q) {
    return M.colT + ((size_t)p * QUARTERS + q) * TILE;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// 64 x 64 bit transpose across the wave: on entry lane r holds row r (bit c =
// element (r, c)), on exit lane c holds column c (bit r = element (r, c)).
// Stage j swaps the off-diagonal j x j sub-blocks of every 2j x 2j block:
// of a lane pair (a, a + j), a keeps its low j-bit groups (mask keep) and
// takes b's low groups shifted up, b keeps its high groups and takes a's high
// groups shifted down.  Each lane sends its value already shifted the way its
// partner needs it, so every lane runs the same bitwise selects (no branch on
// the lane's side of the pair).
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t v) {
    const int lane = threadIdx.x & 63;
    const uint64_t keep[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                              0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int j = 32 >> s;
        const uint64_t hm = (lane & j) ? ~0ull : 0ull;  // all ones on the pair's upper lane
        const uint64_t send = ((v << j) & hm) | ((v >> j) & ~hm);
        const uint64_t recv = shfl_xor64(send, j);
        const uint64_t k = keep[s] ^ hm;  // the bits this lane keeps
        v = (v & k) | (recv & ~k);
    }
    return v;
}

// Column-major form of one 64-row quarter, whole wave: lane r holds row r in
// v.  Word k of dst becomes the mask of the rows holding column k, for every
// column k of cu (the caller's choice of columns: those of the quarter, or
// those that changed).  Words of other columns are left as they are.  Lane l
// computes the words of columns l, 64 + l, 128 + l and 192 + l.
__device__ __forceinline__ void quarter_transpose(const Row4 &v, const Row4 &cu, uint64_t *dst) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (uniform64(cu.w[w]) == 0) continue;
        const uint64_t col = wave_transpose64(v.w[w]);
        if ((cu.w[w] >> lane) & 1ull) gst(dst + (uint32_t)w * 64u + (uint32_t)lane, col);
    }
}

// A wave-uniform 32-byte row of kernel-written data (ld_dev), in scalar registers.
__device__ __forceinline__ Row4 ld_dev_row_uniform(const uint64_t *p) {
    return Row4{{uniform64(ld_dev(p)), uniform64(ld_dev(p + 1)), uniform64(ld_dev(p + 2)), uniform64(ld_dev(p + 3))}};
}

__device__ __forceinline__ Row4 ld_const_row(const uint64_t *p) {
    const __attribute__((address_space(4))) uint64_t *c = (const __attribute__((address_space(4))) uint64_t *)p;
    Row4 r;
    r.w[0] = c[0];
    r.w[1] = c[1];
    r.w[2] = c[2];
    r.w[3] = c[3];
    return r;
}

__device__ __forceinline__ void set_occ(const MatDev &M, uint32_t I, uint32_t J) {
    g_or(&M.occR[I * M.ow + (J >> 5)], 1u << (J & 31));
    g_or(&M.occC[J * M.ow + (I >> 5)], 1u << (I & 31));
}
__device__ __forceinline__ void set_occq(const MatDev &M, uint32_t q, uint32_t I, uint32_t J) {
    g_or(&M.occQ[(q * M.T + I) * M.ow + (J >> 5)], 1u << (J & 31));
}

// Epilogue of one 64-row quarter q of C tile p = (I, J) of a head: the new
// rows (acc & ~M) go to the output delta O and, with `fuse`, are OR-ed into
// the head matrix M right away (fixpoint steps only: an operand read later in
// the same step may then already hold them, which only adds facts that are
// true, while every new row still reaches the next step through O).  Each
// (head, tile, quarter) has exactly one writer per step, so the row and the
// quarter's metadata are updated without atomics; the tile count is shared
// by the quarters.  Whole wave; lane r owns row 64q + r.  Returns the
// quarter's row mask and new-bit count (all lanes).
struct QuarterOut {
    uint64_t bal;
    uint32_t pc;
};
__device__ __forceinline__ QuarterOut quarter_epilogue(const MatDev &O, const MatDev &M, uint32_t p, uint32_t q,
                                                       uint32_t I, uint32_t J, const Row4 &acc, int fuse,
                                                       uint32_t &out_rows, bool shared = false) {
    const int lane = threadIdx.x & 63;
    const size_t off = (size_t)p * TILE_WORDS + (q * WAVE + lane) * TROW_WORDS;
    Row4 nw = {{0, 0, 0, 0}};
    if (row_nonzero(acc)) {
        Row4 m = load_row(M.bits + off);
        nw.w[0] = acc.w[0] & ~m.w[0];
        nw.w[1] = acc.w[1] & ~m.w[1];
        nw.w[2] = acc.w[2] & ~m.w[2];
        nw.w[3] = acc.w[3] & ~m.w[3];
        if (row_nonzero(nw)) {
            store_row(O.bits + off, nw);
            if (fuse) {
                or_row(m, nw);
                store_row(M.bits + off, m);
            }
            ++out_rows;
        }
    }
    const uint64_t bal = __ballot(row_nonzero(nw));
    const RowsUnion ru = rows_union(nw, row_popc(nw), bal);
    const uint32_t pc = ru.sum;
    if (bal) {
        const Row4 cu = ru.cu;
        if (lane == 0) {
            if (shared) {
                // other sub-units write other rows of this quarter: OR, not store
                g_or(&O.rmask[(size_t)p * 4 + q], bal);
                uint64_t *oc = colq_of(O, p, q);
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    if (cu.w[w]) g_or(oc + w, cu.w[w]);
            } else {
                gst(&O.rmask[(size_t)p * 4 + q], bal);
                store_row(colq_of(O, p, q), cu);
            }
            set_occq(O, q, I, J);
            if (fuse) {
                const uint64_t old = gld(&M.rmask[(size_t)p * 4 + q]);
                gst(&M.rmask[(size_t)p * 4 + q], old | bal);
                Row4 mc = load_row(colq_of(M, p, q));
                or_row(mc, cu);
                store_row(colq_of(M, p, q), mc);
                if (!old) set_occq(M, q, I, J);
                if (g_add(&M.cnt[p], pc) == 0) set_occ(M, I, J);
            }
        }
    }
    return QuarterOut{bal, pc};
}

// ---------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_begin(StepCounters *ring, uint32_t slot) {
    uint64_t *w = reinterpret_cast<uint64_t *>(ring + slot);
    for (uint32_t i = threadIdx.x; i < sizeof(StepCounters) / 8; i += blockDim.x) gst(w + i, (uint64_t)0);
}

// occD / occX of matrices ids[0..nid): one thread per 32-tile word of a tile row.
__global__ void k_dense_occ(const MatDev *__restrict__ mats, const uint32_t *__restrict__ ids, uint32_t nid,
                            uint32_t dense_bits, uint32_t mfma_bits) {
    for (uint32_t q = blockIdx.y; q < nid; q += gridDim.y) {
        const MatDev M = mats[ids[q]];
        const uint32_t words = M.T * M.ow;
        for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < words; x += gridDim.x * blockDim.x) {
            const uint32_t I = x / M.ow, w = x - (x / M.ow) * M.ow;
            uint32_t bits = 0, xbits = 0;
            const uint32_t k0 = w * 32, k1 = (k0 + 32 < M.T) ? k0 + 32 : M.T;
            for (uint32_t K = k0; K < k1; ++K) {
                const uint32_t cn = M.cnt[I * M.T + K];
                if (cn >= dense_bits) bits |= 1u << (K - k0);
                if (cn >= mfma_bits) xbits |= 1u << (K - k0);
            }
            M.occD[x] = bits;
            M.occX[x] = xbits;
        }
    }
}

// Wave-aggregated append of `val` for the lanes with `want`.
__device__ __forceinline__ void wave_append(bool want, uint64_t val, uint32_t *count, uint64_t *list) {
    const uint64_t bal = __ballot(want);
    if (!bal) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    if (want) list[base + __popcll(bal & ((1ull << lane) - 1ull))] = val;
}

// Candidate scan of the terms [t0, t1) for C tile (I, J): which quarters have
// candidates, whether some A tile is heavy enough for the dense / MFMA paths
// (only with `heavy`: the heavy-tile bitsets are not maintained when both
// paths are off).
__device__ __forceinline__ void scan_terms(const MatDev *__restrict__ mats, const TermDev *__restrict__ terms,
                                           uint32_t t0, uint32_t t1, int parity, uint32_t T, uint32_t I, uint32_t J,
                                           uint32_t &qmask, bool &dense, bool &mfma, uint32_t *qcnt, int heavy,
                                           bool live = true, bool all_lanes = false) {
    // When a tile row spans whole waves (T a multiple of 64), the wave's
    // tiles share the tile row I: the A operand's words of row I (occupancy,
    // per-quarter, heavy-tile bitsets) are the same for every lane, eight
    // words at a time and together with the B column's words, instead of a
    // chain of dependent loads per word.  all_lanes: every lane of the wave
    // runs this (the tiles to scan are the `live` ones), and the row's 40
    // words come in one vector load, a word per lane, read lane by lane;
    // otherwise (callers under divergent conditions) every active lane reads
    // each word at one address.  (Kernels write these words: never through
    // the scalar cache.)
    const bool row_uniform = (T % WAVE) == 0;
    if (row_uniform && all_lanes && !__ballot(live)) return;
    for (uint32_t t = t0; t < t1; ++t) {
        const MatDev L = mats[terms[t].left[parity]];
        const MatDev R = mats[terms[t].right[parity]];
        if (row_uniform) {
            const uint32_t Iu = uniform(I);
            for (uint32_t w0 = 0; w0 < L.ow; w0 += 8) {
                uint32_t col[8], qw[QUARTERS][8];
                if (all_lanes) {
                    const int lane = threadIdx.x & 63;
                    const uint32_t lw = w0 + (uint32_t)(lane & 7);
                    const int lq = (lane >> 3) - 1;  // -1: occupancy word, 0..3: quarter lq
                    uint32_t rowv = 0;
                    if (lane < 40 && lw < L.ow)
                        rowv = lq < 0 ? gld(&L.occR[Iu * L.ow + lw]) : gld(&L.occQ[((uint32_t)lq * T + Iu) * L.ow + lw]);
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint32_t w = w0 + (uint32_t)i;
                        col[i] = (live && w < L.ow) ? gld(&R.occC[J * R.ow + w]) & lane_read(rowv, (uint32_t)i) : 0u;
#pragma unroll
                        for (uint32_t q = 0; q < QUARTERS; ++q) qw[q][i] = lane_read(rowv, 8u + q * 8u + (uint32_t)i);
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint32_t w = w0 + (uint32_t)i;
                        col[i] = w < L.ow ? gld(&R.occC[J * R.ow + w]) & uniform(ld_dev(&L.occR[Iu * L.ow + w])) : 0u;
#pragma unroll
                        for (uint32_t q = 0; q < QUARTERS; ++q)
                            qw[q][i] = w < L.ow ? uniform(ld_dev(&L.occQ[(q * T + Iu) * L.ow + w])) : 0u;
                    }
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t w = w0 + (uint32_t)i;
                    if (heavy && col[i]) {
                        if (ld_dev(&L.occD[Iu * L.ow + w]) & col[i]) dense = true;
                        if (ld_dev(&L.occX[Iu * L.ow + w]) & col[i]) mfma = true;
                    }
#pragma unroll
                    for (uint32_t q = 0; q < QUARTERS; ++q) {
                        const uint32_t k = qw[q][i] & col[i];
                        if (k) qmask |= 1u << q;
                        qcnt[q] += __popc(k);
                    }
                }
            }
            continue;
        }
        if (!live) continue;
        for (uint32_t w = 0; w < L.ow; ++w) {
            uint32_t col = gld(&R.occC[J * R.ow + w]);
            if (!col) continue;
            // the per-quarter and heavy-tile bitsets are subsets of the
            // tile-row occupancy: one load rules most words out
            col &= gld(&L.occR[I * L.ow + w]);
            if (!col) continue;
            if (heavy) {
                if (gld(&L.occD[I * L.ow + w]) & col) dense = true;
                if (gld(&L.occX[I * L.ow + w]) & col) mfma = true;
            }
#pragma unroll
            for (uint32_t q = 0; q < QUARTERS; ++q) {
                const uint32_t k = gld(&L.occQ[(q * T + I) * L.ow + w]) & col;
                if (k) qmask |= 1u << q;
                qcnt[q] += __popc(k);
            }
        }
    }
}

// Append the sparse units of one (head, C tile): quarters in qmask, a heavy
// quarter (more than `split` candidates; split = 0: none) as up to SPLIT_MAX
// sub-units of about `piece` candidates each
// while the split budget lasts.  Sub-unit s of a split quarter takes the
// candidates g = s (mod nsub) of the quarter and leaves its rows in a slot
// of the partial-row buffer (slot = its index in the sub-unit list); the
// split quarter is listed in `splits` (the base item with nsub, and the
// first slot) for k_merge_split, which ORs the partial rows and writes the
// quarter.  Whole wave: the single units of all four
// quarters take one atomic on the list count per wave (the count is one
// address that every wave of the build hits).  Sub-units go to their own list
// at units + heavy_base (count n_heavy), which the sparse kernel claims
// first: the longest units start early instead of ending the launch late.
__device__ __forceinline__ void append_units(bool want, uint32_t qmask, const uint32_t *qcnt, uint64_t base,
                                             uint32_t split, uint32_t piece, uint32_t split_budget,
                                             StepCounters *__restrict__ c,
                                             uint64_t *__restrict__ units, uint32_t heavy_base, int tile_units,
                                             uint64_t *__restrict__ splits, bool prio = false,
                                             uint64_t *__restrict__ prio_list = nullptr, int split_rows = 0,
                                             bool block_agg = false) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    if (prio_list) {
        // items the scheduler ranks first: their units (whole, unsplit) go to
        // the list the first phase of the step evaluates
        const bool pr = want && prio && qmask;
        if (tile_units) {
            wave_append(pr, base | ITEM_QSET | ((uint64_t)qmask << ITEM_QSET_SHIFT), &c->n_prio, prio_list);
        } else {
#pragma unroll
            for (uint32_t q = 0; q < QUARTERS; ++q) wave_append(pr && ((qmask >> q) & 1u), base | (uint64_t)q, &c->n_prio, prio_list);
        }
        want = want && !pr;
    }
    // the work-unit shape: tile_units 1 = one unit for every quarter of the
    // tile with candidates (256 output rows), 2 = one per half (quarters 0-1
    // and 2-3, 128 rows), 0 = one per quarter (64 rows).  A group holding a
    // quarter heavy enough to be split is left to the quarter units.
    uint32_t rest = want ? qmask : 0u;
    if (tile_units) {
        const uint32_t groups = tile_units == 1 ? 1u : 2u;
        for (uint32_t gi = 0; gi < groups; ++gi) {
            const uint32_t gm = tile_units == 1 ? 0xFu : (3u << (2 * gi));
            const uint32_t m = rest & gm;
            bool heavyq = false;
#pragma unroll
            for (uint32_t q = 0; q < QUARTERS; ++q) heavyq |= split && ((m >> q) & 1u) && qcnt[q] > split;
            const bool tu = m && !heavyq;
            wave_append(tu, base | ITEM_QSET | ((uint64_t)m << ITEM_QSET_SHIFT), &c->n_units, units);
            if (tu) rest &= ~gm;
        }
    }
    // hv[q]: the quarter's entries in the heavy list -- its sub-units (more
    // than one), or 1 for a quarter too light to split but of more than
    // HEAVY_UNSPLIT_16THS / 16 of the split threshold, which the heavy list
    // only moves to the front of the claim order, so that it starts with the
    // step instead of ending it late; 0: the quarter is one unit of the
    // ordinary list
    uint32_t hv[QUARTERS];
    uint64_t single[QUARTERS];
    uint32_t total = 0;
    const uint32_t big = (uint32_t)((uint64_t)split * HEAVY_UNSPLIT_16THS / 16u);
#pragma unroll
    for (uint32_t q = 0; q < QUARTERS; ++q) {
        const bool wq = ((rest >> q) & 1u) != 0;
        hv[q] = 0;
        uint32_t want = 0;
        if (wq && split && qcnt[q] > split) {
            uint32_t ns = (uint32_t)(((uint64_t)qcnt[q] + piece - 1) / piece);
            // a quarter of at least SPLIT_MAX / 2 pieces takes all SPLIT_MAX
            // sub-units: on the forest such quarters' candidates also carry
            // the most (hit, column) items each, so their sub-units were the
            // longest of their step
            if (ns >= SPLIT_MAX / 2) ns = SPLIT_MAX;
            ns = ns > SPLIT_MAX ? SPLIT_MAX : (ns < 1 ? 1 : ns);
            want = ns;
        } else if (wq && split && qcnt[q] > big) {
            want = 1;
        }
        // the budget counts heavy-list entries (a sub-unit's is also its
        // partial-row slot), reserved for the wave's heavy quarters with one
        // atomic (lane i's entries end at the prefix sum of the requests); a
        // refused reservation only leaves budget unused
        const uint64_t wb = __ballot(want != 0);
        if (wb) {
            uint32_t tot = 0;
            const uint32_t incl = scan_ballot(want, wb, tot);
            uint32_t old = 0;
            if (lane == 0) old = atomicAdd(&c->split_extra, tot);
            old = uniform(old);
            if (want) hv[q] = old + incl > split_budget ? 0u : want;
        }
        single[q] = __ballot(wq && hv[q] == 0);
        total += (uint32_t)__popcll(single[q]);
    }
    // The single units' list positions: one atomic per work-group (block_agg:
    // every thread of the group calls this once) or per wave.  Every wave of
    // the build hits the one count, so fewer atomics on it shorten the build.
    uint32_t at = 0;
    if (block_agg) {
        __shared__ uint32_t s_wcnt[BLOCK / WAVE];
        __shared__ uint32_t s_base;
        const uint32_t wv = threadIdx.x / WAVE;
        if (lane == 0) s_wcnt[wv] = total;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t sum = 0;
            for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) sum += s_wcnt[w];
            s_base = sum ? atomicAdd(&c->n_units, sum) : 0u;
        }
        __syncthreads();
        at = s_base;
        for (uint32_t w = 0; w < wv; ++w) at += s_wcnt[w];
    } else if (total) {
        if (lane == 0) at = atomicAdd(&c->n_units, total);
        at = uniform(at);
    }
    if (total) {
#pragma unroll
        for (uint32_t q = 0; q < QUARTERS; ++q) {
            if ((single[q] >> lane) & 1u) units[at + __popcll(single[q] & below)] = base | (uint64_t)q;
            at += (uint32_t)__popcll(single[q]);
        }
    }
    // heavy quarters: their heavy-list entries, and the split entries of the
    // split ones, one atomic per wave each
#pragma unroll
    for (uint32_t q = 0; q < QUARTERS; ++q) {
        const uint32_t n = hv[q];
        const uint64_t hb = __ballot(n != 0);
        if (!hb) continue;
        const uint64_t sb = __ballot(n > 1);  // split quarters
        uint32_t tot = 0;
        const uint32_t incl = scan_ballot(n, hb, tot);
        uint32_t hbase = 0, sbase = 0;
        if (lane == 0) {
            hbase = atomicAdd(&c->n_heavy, tot);
            if (!split_rows && sb) sbase = atomicAdd(&c->n_split, (uint32_t)__popcll(sb));
        }
        hbase = uniform(hbase);
        sbase = uniform(sbase);
        if (n) {
            const uint32_t hat = hbase + incl - n;
            const uint64_t u = base | (uint64_t)q | ((uint64_t)(n - 1) << 59) | (split_rows && n > 1 ? ITEM_ROWSPLIT : 0ull);
            for (uint32_t sidx = 0; sidx < n; ++sidx) units[heavy_base + hat + sidx] = u | ((uint64_t)sidx << 55);
            if (!split_rows && n > 1) {  // candidate split: the quarter is finished by k_merge_split
                const uint32_t sq = sbase + (uint32_t)__popcll(sb & below);
                splits[2 * (size_t)sq] = u;
                splits[2 * (size_t)sq + 1] = hat;
            }
        }
    }
}

// One thread per (head h0 + blockIdx.y, C tile).  Skips full head tiles.
__global__ __launch_bounds__(BLOCK) void k_build(const MatDev *__restrict__ mats,
                                                 const HeadDev *__restrict__ heads,
                                                 const TermDev *__restrict__ terms, int parity, int full_pass,
                                                 StepCounters *__restrict__ c, uint64_t *__restrict__ units,
                                                 uint64_t *__restrict__ dense_items,
                                                 uint64_t *__restrict__ mfma_items, uint32_t split, uint32_t piece,
                                                 uint32_t split_budget, int heavy, uint32_t heavy_base,
                                                 uint32_t h0, const StepCounters *__restrict__ prevc,
                                                 int tile_units, uint64_t *__restrict__ splits,
                                                 int split_rows) {
    // a delta step after one that added no bit has an empty delta: no work
    // (the steps the host enqueues past convergence end here)
    if (prevc) {
        uint64_t nb = prevc->new_bits;
#pragma unroll
        for (int s = 0; s < SHARDS; ++s) nb += prevc->shard[s].new_bits;
        if (uniform((uint32_t)(nb != 0)) == 0) return;
    }
    const uint32_t h = h0 + blockIdx.y;
    const uint32_t T = uniform(heads[h].T), target = uniform(heads[h].target);
    const uint32_t t0 = uniform(full_pass ? heads[h].f0 : heads[h].t0);
    const uint32_t t1 = uniform(full_pass ? heads[h].f1 : heads[h].t1);
    const uint64_t flag = full_pass ? ITEM_FULL : 0;
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    bool dense = false, mfma = false;
    uint32_t qmask = 0;
    uint32_t qcnt[QUARTERS] = {0, 0, 0, 0};
    {
        // every lane scans (the live tiles: inside the head, not full)
        const bool in = p < T * T;
        const uint32_t I = in ? p / T : 0u, J = in ? p - (p / T) * T : 0u;
        const bool live = in && gld(&mats[target].cnt[p]) < TILE_BITS;
        scan_terms(mats, terms, t0, t1, parity, T, I, J, qmask, dense, mfma, qcnt, heavy, live, true);
    }
    wave_append(mfma, flag | ((uint64_t)h << 40) | p, &c->n_mfma, mfma_items);
    wave_append(dense && !mfma, flag | ((uint64_t)h << 40) | p, &c->n_dense, dense_items);
    dense = dense || mfma;
    append_units(!dense, qmask, qcnt, flag | ((uint64_t)h << 40) | ((uint64_t)p << 2), split, piece, split_budget, c,
                 units, heavy_base, tile_units, splits, false, nullptr, split_rows, true);
}

// ---------------------------------------------------------------------------
// sparse product: one wavefront per (head, C tile, quarter)
// ---------------------------------------------------------------------------

// SPARSE_BATCH: candidate k-tiles whose loads are issued together (1, 2 or 4;
// more hides more latency per wave but costs registers, hence occupancy).
template <int SPARSE_BATCH>
__global__ __launch_bounds__(WAVE) void k_product_sparse(const MatDev *__restrict__ mats,
                                                         const HeadDev *__restrict__ heads,
                                                         const TermDev *__restrict__ terms, int parity,
                                                         StepCounters *__restrict__ c,
                                                         const uint64_t *__restrict__ units,
                                                         uint64_t *__restrict__ touched, uint32_t chunk,
                                                         int fuse) {
    const int lane = threadIdx.x & 63;
    const uint32_t n_units = c->n_units;
    uint32_t tp = 0, arows = 0, brows = 0, out_rows = 0, hits = 0, max_cands = 0;
    uint64_t new_bits = 0;  // lane 0: new bits of this wave's units
    for (uint32_t u0 = wave_next(&c->unit_next, chunk); u0 < n_units; u0 = wave_next(&c->unit_next, chunk)) {
    const uint32_t u1 = (u0 + chunk < n_units) ? u0 + chunk : n_units;
    for (uint32_t u = u0; u < u1; ++u) {
        const uint64_t it = uniform64(ld_dev(units + u));
        const uint32_t h = uniform(item_head(it));
        const uint32_t p = uniform((uint32_t)((it & ITEM_TILE_MASK) >> 2));
        const uint32_t q = uniform((uint32_t)(it & 3u));
        const bool full = (it & ITEM_FULL) != 0;
        const uint32_t T = ld_const(&heads[h].T);
        const uint32_t t0 = full ? ld_const(&heads[h].f0) : ld_const(&heads[h].t0);
        const uint32_t t1 = full ? ld_const(&heads[h].f1) : ld_const(&heads[h].t1);
        const uint32_t I = p / T, J = p - (p / T) * T;
        const uint32_t row = q * WAVE + lane;
        const uint32_t tp_unit0 = tp;
        Row4 acc = {{0, 0, 0, 0}};
        for (uint32_t t = t0; t < t1; ++t) {
            const MatDev L = mats[ld_const(&terms[t].left[parity])];
            const MatDev R = mats[ld_const(&terms[t].right[parity])];
            for (uint32_t w = 0; w < L.ow; ++w) {
                uint32_t kb = uniform(ld_dev(&L.occQ[(q * T + I) * L.ow + w]) & ld_dev(&R.occC[J * R.ow + w]));
                while (kb) {
                    // Up to SPARSE_BATCH candidate k-tiles at once, so that their
                    // dependent loads (row masks, then A rows, then B rows) are
                    // in flight together instead of one latency chain each.
                    uint32_t lts[SPARSE_BATCH], rts[SPARSE_BATCH];
                    uint32_t nk = 0;
#pragma unroll
                    for (int j = 0; j < SPARSE_BATCH; ++j) {
                        const uint32_t K = w * 32 + (kb ? __builtin_ctz(kb) : 0);
                        lts[j] = I * T + K;
                        rts[j] = K * T + J;
                        if (kb) {
                            kb &= kb - 1;
                            nk = j + 1;
                        }
                    }
                    uint64_t rm[SPARSE_BATCH];
                    Row4 brm[SPARSE_BATCH];
#pragma unroll
                    for (int j = 0; j < SPARSE_BATCH; ++j) {
                        rm[j] = 0;
                        brm[j] = Row4{{0, 0, 0, 0}};
                        if ((uint32_t)j < nk) {
                            // all three loads are independent and go out together
                            brm[j] = ld_dev_row_uniform(R.rmask + (size_t)rts[j] * 4);  // non-empty rows of the B tile
                            const Row4 cu = ld_dev_row_uniform(colq_of(L, lts[j], q));  // columns used by this A quarter
                            const uint64_t arm = uniform64(ld_dev(&L.rmask[(size_t)lts[j] * 4 + q]));
                            const bool hit = ((cu.w[0] & brm[j].w[0]) | (cu.w[1] & brm[j].w[1]) |
                                              (cu.w[2] & brm[j].w[2]) | (cu.w[3] & brm[j].w[3])) != 0;
                            rm[j] = hit ? arm : 0;
                            hits += rm[j] != 0;
                        }
                    }
                    Row4 a[SPARSE_BATCH];
#pragma unroll
                    for (int j = 0; j < SPARSE_BATCH; ++j) {
                        a[j] = Row4{{0, 0, 0, 0}};
                        if ((rm[j] >> lane) & 1ull) {
                            ++arows;
                            a[j] = load_row(L.bits + (size_t)lts[j] * TILE_WORDS + row * TROW_WORDS);
                        }
                    }
                    tp += nk;
#pragma unroll
                    for (int j = 0; j < SPARSE_BATCH; ++j) {
                        const uint64_t *Bt = R.bits + (size_t)rts[j] * TILE_WORDS;
#pragma unroll
                        for (int ww = 0; ww < 4; ++ww) {
                            uint64_t x = a[j].w[ww] & brm[j].w[ww];  // only B rows that hold bits
                            while (x) {
                                const int k = ww * 64 + __builtin_ctzll(x);
                                x &= x - 1;
                                or_row(acc, load_row(Bt + k * TROW_WORDS));
                                ++brows;
                            }
                        }
                    }
                }
            }
        }
        max_cands = max(max_cands, tp - tp_unit0);
        const MatDev O = mats[uniform(heads[h].out[parity])];
        const MatDev M = mats[uniform(heads[h].target)];
        const QuarterOut qo = quarter_epilogue(O, M, p, q, I, J, acc, fuse, out_rows);
        if (qo.bal && lane == 0) {
            const uint32_t old = g_add(&O.cnt[p], qo.pc);
            if (old == 0) {
                set_occ(O, I, J);
                touched[atomicAdd(&c->n_touched, 1u)] = ((uint64_t)h << 40) | p;
            }
            new_bits += qo.pc;
        }
    }
    }
    const uint32_t ar = wave_sum(arows), br = wave_sum(brows), orow = wave_sum(out_rows);
    if (lane == 0) {
        if (tp) atomicAdd((unsigned long long *)&c->tile_products, (unsigned long long)tp);
        if (ar) atomicAdd((unsigned long long *)&c->a_rows, (unsigned long long)ar);
        if (br) atomicAdd((unsigned long long *)&c->b_rows, (unsigned long long)br);
        if (orow) atomicAdd((unsigned long long *)&c->out_rows, (unsigned long long)orow);
        if (hits) atomicAdd((unsigned long long *)&c->cand_hits, (unsigned long long)hits);
        if (max_cands) atomicMax(&c->max_unit_cands, max_cands);
        if (new_bits) atomicAdd((unsigned long long *)&c->new_bits, (unsigned long long)new_bits);
    }
}

// Inclusive prefix sum over the 64 lanes.

// Position of the j-th (0-based) set bit of m (j < popcount(m)), by halving.
__device__ __forceinline__ int select_bit(uint64_t m, uint32_t j) {
    int r = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << s) - 1ull));
        if (j >= c) {
            j -= c;
            m >>= s;
            r += s;
        }
    }
    return r;
}

// (hit, shared column) items a lane of the lane-parallel sparse kernel takes
// per round in its column-major gathers: their loads are in flight together.
constexpr int COL_ITEMS = 4;

// Rows [64 s / n, 64 (s + 1) / n) of a quarter, as a mask (n <= 64).
__device__ __forceinline__ uint64_t rows_of_part(uint32_t s, uint32_t n) {
    const uint32_t a = 64u * s / n, b = 64u * (s + 1u) / n;
    const uint64_t hi = b >= 64u ? ~0ull : (1ull << b) - 1ull;
    return hi & ~((1ull << a) - 1ull);
}

// Column of the j-th (0-based) set bit of a 256-bit row mask (j < popcount).
__device__ __forceinline__ uint32_t select_bit256(const Row4 &m, uint32_t j) {
    uint32_t base = 0, k = 0;
    bool found = false;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t c = (uint32_t)__popcll(m.w[w]);
        if (!found && j < c) {
            k = base + (uint32_t)select_bit(m.w[w], j);
            found = true;
        }
        if (!found) j -= c;
        base += 64;
    }
    return k;
}

// m with its lowest j set bits cleared (j <= popcount(m)).
__device__ __forceinline__ Row4 bits_from(Row4 m, uint32_t j) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t c = (uint32_t)__popcll(m.w[w]);
        if (j >= c) {
            j -= c;
            m.w[w] = 0;
        } else {
            if (j) m.w[w] &= ~0ull << select_bit(m.w[w], j);
            j = 0;
        }
    }
    return m;
}

// Hub rows: a pair round whose pairs name at most GATHER_HEAVY B rows each
// is gathered lane by lane; otherwise its rows are dealt out over the wave
// in chunks of GATHER_CH rows (loads in flight per lane).
constexpr int GATHER_CH = 8;
// B-row loads in flight per lane per round in the common (non-hub) gather
// (8 measured 7% slower on the ontology query: the extra rows cost registers).
constexpr int GATHER_ROUND = 4;
constexpr uint32_t GATHER_HEAVY = 32;

// OR the B rows named by the lowest GATHER_CH set bits of x (tile row k is
// bit k) into the LDS row accumulator *acc; returns the rows loaded.
__device__ __forceinline__ uint32_t gather_chunk(Row4 x, const uint64_t *Bt, Row4 *acc) {
    int ks[GATHER_CH];
    int nk = 0;
#pragma unroll
    for (int jj = 0; jj < GATHER_CH; ++jj) {
        int k = 0;
        bool have = true;
        if (x.w[0]) {
            k = __builtin_ctzll(x.w[0]);
            x.w[0] &= x.w[0] - 1;
        } else if (x.w[1]) {
            k = 64 + __builtin_ctzll(x.w[1]);
            x.w[1] &= x.w[1] - 1;
        } else if (x.w[2]) {
            k = 128 + __builtin_ctzll(x.w[2]);
            x.w[2] &= x.w[2] - 1;
        } else if (x.w[3]) {
            k = 192 + __builtin_ctzll(x.w[3]);
            x.w[3] &= x.w[3] - 1;
        } else {
            have = false;
        }
        ks[jj] = k;
        if (have) nk = jj + 1;
    }
    Row4 b[GATHER_CH];
#pragma unroll
    for (int jj = 0; jj < GATHER_CH; ++jj) b[jj] = jj < nk ? load_row(Bt + ks[jj] * TROW_WORDS) : Row4{{0, 0, 0, 0}};
    Row4 v = {{0, 0, 0, 0}};
#pragma unroll
    for (int jj = 0; jj < GATHER_CH; ++jj) or_row(v, b[jj]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (v.w[i]) atomicOr((unsigned long long *)&acc->w[i], (unsigned long long)v.w[i]);
    return (uint32_t)nk;
}

// The operand arrays of one term, staged per scan lane in LDS for the filter
// and gather stages (a subset of TermPtr).
struct LaneOps {
    const uint64_t *lrmask, *lcolq, *lcolT, *lbits, *rrmask, *rbits;
    const uint32_t *rcnt;
};

// Lane-parallel variant: the same work unit (head, C tile, 64-row quarter),
// but the wave's lanes work on different candidate k-tiles at once.
//   scan    lane l reads the candidate mask of (term, occupancy word) pair l
//           (the A quarter's tile-row occupancy AND the B tile-column
//           occupancy); a prefix sum of their popcounts numbers the
//           candidates of all terms and words of the unit;
//   filter  lane g takes candidate g (binary search for its word, bit
//           selection within it) and tests it with vector loads of the B
//           tile's row mask and the A quarter's column union and row mask; a
//           ballot gives the hits;
//   gather  column-major: the hits' (candidate, shared column k) items are
//           numbered by a prefix sum over the lanes and dealt out 64 at a
//           time; a lane finds its item's candidate by binary search and the
//           column by bit selection, loads the A quarter's column word k (the
//           rows holding k) and B row k together, and ORs B row k into each
//           of those rows of the unit's LDS row accumulator (ds_or_b64);
//   epilogue  lane r owns row r of the quarter, as in the batched kernel.
// 64 candidates are filtered per load round and COL_ITEMS x 64 items are in
// flight per wave, however they are spread over the candidates.
// QSET: the units may cover several quarters of a tile (tile units, a
// 256-row accumulator); otherwise every unit is one quarter and the
// accumulator has 64 rows (a smaller LDS footprint, more waves per CU).
// RSPLIT: the step may hold row-split sub-units (ITEM_ROWSPLIT); the kernel
// without that support is launched otherwise (its per-item row selection
// measured a few per cent of the kernel's time).
template <bool QSET, bool RSPLIT>
__global__ __launch_bounds__(WAVE) void k_product_sparse_lp(const MatDev *__restrict__ mats,
                                                            const HeadDev *__restrict__ heads,
                                                            const TermPtr *__restrict__ tptr, int parity,
                                                            StepCounters *__restrict__ c,
                                                            const uint64_t *__restrict__ units,
                                                            uint32_t heavy_base, TouchedLists tl, uint32_t chunk,
                                                            int fuse, uint64_t *__restrict__ partials,
                                                            uint64_t *__restrict__ pmask, int phase,
                                                            const uint64_t *__restrict__ prio_list) {
    constexpr uint32_t ACC_ROWS = QSET ? TILE : WAVE;
    __shared__ Row4 acc_s[ACC_ROWS];  // rows of the unit's quarters (row 64 q + r with QSET, else r)
    __shared__ Row4 s_brm[WAVE];
    __shared__ uint64_t s_arm[WAVE];
    // with QSET: the quarter of each staged hit, and the candidate masks per
    // (term, word) lane and quarter (without, s_kb holds the one mask)
    __shared__ uint32_t s_hq[QSET ? WAVE : 1];
    __shared__ uint32_t s_kq[QSET ? WAVE : 1][QUARTERS];
    __shared__ const uint64_t *s_ab[WAVE];
    __shared__ const uint64_t *s_bb[WAVE];
    __shared__ uint32_t s_end[WAVE], s_kb[WAVE], s_cend[WAVE];
    __shared__ LaneOps s_ops[WAVE];
    const int lane = threadIdx.x & 63;
    const uint64_t clk0 = __builtin_amdgcn_s_memtime();
    // phase 1: the units the scheduler ranked first (prio_list); otherwise
    // list entries [0, n_heavy) are the heavy list (sub-units of split units and
    // quarters of more than half the split threshold), then the others
    const uint32_t n_heavy = phase == 1 ? 0u : c->n_heavy;
    const uint32_t n_units = phase == 1 ? c->n_prio : c->n_units + n_heavy;
    // Claims x are mapped to list entries by reversing the low six bits of x:
    // neighbours in the list (the quarters of one tile, the sub-units of one
    // heavy quarter) go to claims far apart within their block of 64, i.e. to
    // different waves, while blocks -- and the locality of the list order --
    // are kept.  x runs over whole blocks; entries past the list end are skipped.
    // Block b belongs to shard b mod SHARDS; a wave claims `chunk` (a divisor
    // of 64) claims at a time from its own shard's cursor, then from the
    // other shards' in turn once its own is used up.
    const uint32_t n_claim = (n_units + 63u) & ~63u;
    // A step with no more units than waves (a small problem, the last steps
    // of a fixpoint): wave b takes unit b, with no claim at all.
    const bool direct = n_units <= gridDim.x;
    if (direct ? blockIdx.x >= n_units : (uint64_t)blockIdx.x * chunk >= n_claim) return;  // more waves than claims
    const uint32_t nblk = n_claim / 64u;
    const uint32_t home = blockIdx.x % SHARDS;
    CounterShard *shards = c->shard;
    uint32_t tp = 0, arows = 0, acols = 0, brows = 0, out_rows = 0, hits_n = 0, max_cands = 0;
    uint64_t new_bits = 0;  // lane 0: new bits of this wave's units
    uint64_t unit_max = 0;  // longest unit of this wave (clocks)
    uint64_t unit_max_a = 0, unit_max_b = 0, unit_max_c = 0, unit_max_d = 0;  // its shape (CounterShard)
    uint64_t unit_clk = 0, idle_clk = 0;  // lane 0: clocks of all units / of units that added no bit
    uint32_t idle_units = 0, idle_brows = 0, units_done = 0;
    uint32_t sh = home;
    // The claim of the next chunk is issued when a chunk starts and read when
    // it ends, so the atomic's round trip overlaps the chunk's own loads.  A
    // claim is always taken by the wave that issued it (cursors only grow, so
    // a claim past the end of a shard only says the shard is used up).
    auto claim = [&](uint32_t s) -> uint32_t {
        uint32_t k = 0xffffffffu;
        if (lane == 0) k = atomicAdd(phase == 1 ? &shards[s].prio_next : &shards[s].unit_next, chunk);
        return k;
    };
    uint32_t pending = direct ? 0u : claim(sh);
    bool first = true;
    for (;;) {
    uint32_t u0, u1;
    if (direct) {
        if (!first) break;
        first = false;
        u0 = blockIdx.x;
        u1 = u0 + 1u;
    } else {
    const uint32_t kc = uniform(pending);
    const uint32_t blk = (kc / 64u) * SHARDS + sh;
    if (blk >= nblk) {
        // this shard's blocks are all claimed: mark it in the step's mask of
        // used-up shards, which the same atomic returns, and go on with the
        // next shard not marked, or end (one round trip, instead of a claim on
        // every other shard in turn at the end of every wave)
        uint32_t m = 0;
        if (lane == 0) m = atomicOr(phase == 1 ? &c->prio_shards_done : &c->shards_done, 1u << sh);
        m = uniform(m) | (1u << sh);
        const uint32_t left = ~m & ((1u << SHARDS) - 1u);
        if (!left) break;
        const uint32_t after = left & ~((2u << sh) - 1u);  // shards past sh
        sh = (uint32_t)__ffs(after ? after : left) - 1u;
        pending = claim(sh);
        continue;
    }
    pending = claim(sh);
    u0 = blk * 64u + kc % 64u;
    u1 = (u0 + chunk < n_claim) ? u0 + chunk : n_claim;
    }
    for (uint32_t x = u0; x < u1; ++x) {
        const uint32_t u = direct ? x : (x & ~63u) | (__builtin_bitreverse32(x & 63u) >> 26);
        if (u >= n_units) continue;
        const uint64_t tu0 = __builtin_amdgcn_s_memtime();
        // (the list entry is written by this step's build, so it is read through
        // the vector path; the head's fields are host-written and come through
        // the scalar cache.  Loading a chunk's entries one per lane in one
        // load measured no faster.)
        const uint64_t it = uniform64(phase == 1 ? ld_dev(prio_list + u) : ld_dev(units + (u < n_heavy ? heavy_base + u : u - n_heavy)));
        const uint32_t h = uniform(item_head(it));
        const uint32_t p = uniform(unit_tile(it));
        // the unit's quarters: one, or every quarter of the tile with candidates
        const uint32_t qset = uniform(unit_quarters(it));
        const bool full = (it & ITEM_FULL) != 0;
        const uint32_t sub = uniform(unit_sub(it)), nsub = uniform(unit_nsub(it));
        // a row-split sub-unit: its rows of the quarter
        const bool rsplit = RSPLIT && nsub > 1 && (it & ITEM_ROWSPLIT) != 0;
        const uint64_t rowsel = rsplit ? rows_of_part(sub, nsub) : ~0ull;
        const uint32_t T = ld_const(&heads[h].T);
        const uint32_t t0 = full ? ld_const(&heads[h].f0) : ld_const(&heads[h].t0);
        const uint32_t t1 = full ? ld_const(&heads[h].f1) : ld_const(&heads[h].t1);
        const uint32_t I = p / T, J = p - (p / T) * T;
        // the output descriptors are read now, while the unit's first loads are in flight
        const MatDev O = mats[uniform(heads[h].out[parity])];
        const MatDev M = mats[uniform(heads[h].target)];
        const uint32_t tp_unit0 = tp, br_unit0 = brows;
        uint32_t unit_pairs = 0;  // (hit, shared column) items of the unit (its shape, for the step log)
        // a split quarter's sub-unit `sub` of `nsub` takes the candidates
        // g = sub (mod nsub), so that a heavy quarter's gathers are spread over
        // several waves; its rows go to a partial-row slot (k_merge_split)
#pragma unroll
        for (uint32_t qq = 0; qq < QUARTERS; ++qq)
            if ((qset >> qq) & 1u) acc_s[(QSET ? qq * WAVE : 0u) + lane] = Row4{{0, 0, 0, 0}};
        __syncthreads();
        const uint32_t ow = (T + 31) / 32;
        const uint32_t ntw = (t1 - t0) * ow;  // (term, occupancy word) pairs of the unit
        uint64_t c_scan = 0, c_filter = 0, c_pairs = 0;  // stage clocks of the unit (step log)
        for (uint32_t tw0 = 0; tw0 < ntw; tw0 += WAVE) {
            // 1. candidate masks of 64 (term, word) pairs at once, compacted by
            //    a prefix sum of their popcounts
            const uint64_t t_scan = __builtin_amdgcn_s_memtime();
            const uint32_t j = tw0 + lane;
            uint32_t kq[QUARTERS] = {0, 0, 0, 0};
            LaneOps ops = {};
            if (j < ntw) {
                const uint32_t t = t0 + j / ow, w = j % ow;
                const TermPtr e = tptr[t];
                const uint32_t col = gld(&e.occC[J * ow + w]);
#pragma unroll
                for (uint32_t qq = 0; qq < QUARTERS; ++qq)
                    if ((qset >> qq) & 1u) kq[qq] = gld(&e.occQ[(qq * T + I) * ow + w]) & col;
                ops = LaneOps{e.lrmask, e.lcolq, e.lcolT, e.lbits, e.rrmask, e.rbits, e.rcnt};
            }
            const uint32_t kbw = (uint32_t)(__popc(kq[0]) + __popc(kq[1]) + __popc(kq[2]) + __popc(kq[3]));
            uint32_t total;
            const uint32_t incw = scan_ballot(kbw, __ballot(kbw != 0), total);
            c_scan += __builtin_amdgcn_s_memtime() - t_scan;
            if (!total) continue;
#pragma unroll
            for (uint32_t qq = 0; qq < QUARTERS; ++qq)
                if (QSET) s_kq[lane][qq] = kq[qq];
            s_kb[lane] = QSET ? kbw : (kq[0] | kq[1] | kq[2] | kq[3]);
            s_cend[lane] = incw;
            s_ops[lane] = ops;
            __syncthreads();
            for (uint32_t cb = 0; cb < total; cb += WAVE) {
                const uint32_t nc = (total - cb < WAVE) ? total - cb : WAVE;
                const uint64_t t_filter = __builtin_amdgcn_s_memtime();
                // 2. filter up to 64 candidates: lane g takes the g-th candidate
                //    (of a sub-unit: the candidates of its residue class)
                const uint32_t g = cb + lane;
                const bool mine = (uint32_t)lane < nc && (nsub == 1 || rsplit || g % nsub == sub);
                // (the sub-units of a row-split quarter share its candidates: counted once)
                if (!rsplit || sub == 0) tp += (uint32_t)__popcll(__ballot(mine));
                uint64_t arm = 0;  // (a hit on a full B tile: the A quarter's non-empty rows)
                bool hit = false;
                Row4 brm = {{0, 0, 0, 0}};
                const uint64_t *abase = nullptr, *bbase = nullptr;
                uint32_t cq = 0;  // the candidate's quarter
                if (mine) {
                    uint32_t lo = 0, hi = WAVE - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (s_cend[mid] > g) hi = mid;
                        else lo = mid + 1;
                    }
                    // the owner's candidates, quarter by quarter: find this one's quarter and bit
                    uint32_t q, kb, idx;
                    if (QSET) {
                        idx = g - (s_cend[lo] - s_kb[lo]);
                        q = 0;
                        kb = s_kq[lo][0];
                        while ((uint32_t)__popc(kb) <= idx) {
                            idx -= (uint32_t)__popc(kb);
                            kb = s_kq[lo][++q];
                        }
                    } else {
                        q = (uint32_t)__builtin_ctz(qset);
                        kb = s_kb[lo];
                        idx = g - (s_cend[lo] - (uint32_t)__popc(kb));
                    }
                    cq = q;
                    const int bit = select_bit(kb, idx);
                    const uint32_t jj = tw0 + lo;
                    const uint32_t K = (jj % ow) * 32 + (uint32_t)bit;
                    const LaneOps o = s_ops[lo];
                    const size_t lt = (size_t)I * T + K, rt = (size_t)K * T + J;
                    // the B tile's row mask, the A quarter's column union and
                    // row mask and the B tile's count, loaded together (reading
                    // the last two only for hits on B tiles whose rows are all
                    // non-empty costs a second round trip where such tiles are
                    // common, as in saturating queries)
                    brm = load_row(o.rrmask + rt * 4);
                    const Row4 cu = load_row(o.lcolq + (lt * QUARTERS + q) * TROW_WORDS);
                    const uint64_t qrm = gld(&o.lrmask[lt * 4 + q]);
                    const uint32_t bcnt = gld(&o.rcnt[rt]);
                    hit = ((cu.w[0] & brm.w[0]) | (cu.w[1] & brm.w[1]) | (cu.w[2] & brm.w[2]) |
                           (cu.w[3] & brm.w[3])) != 0;
                    // column-major gathers: the columns the A quarter and the B
                    // tile's non-empty rows share, and the quarter's column words
                    brm = Row4{{cu.w[0] & brm.w[0], cu.w[1] & brm.w[1], cu.w[2] & brm.w[2], cu.w[3] & brm.w[3]}};
                    abase = o.lcolT + (lt * QUARTERS + q) * TILE;
                    bbase = o.rbits + rt * TILE_WORDS;
                    // a full B tile needs no gathers (every row of the quarter
                    // becomes all ones): bit 0 of the (8-byte aligned) pointer
                    // marks it, and the quarter's row mask names the rows
                    if (hit && bcnt == TILE_BITS) {
                        bbase = (const uint64_t *)((uintptr_t)bbase | 1u);
                        arm = qrm;
                    }
                }
                // 3. stage the hits in lane order and spread their (candidate,
                //    A row) pairs over the lanes
                const uint64_t hmask = __ballot(hit);
                if (!hmask) continue;
                const uint32_t nh = __popcll(hmask);
                // work items of a hit: the columns it shares with the B tile
                // (one item for a full B tile, whose rows are all ones)
                const uint32_t cnt = !hit ? 0u : ((uintptr_t)bbase & 1u) ? 1u : row_popc(brm);
                uint32_t pairs;
                const uint32_t incl = scan_ballot(cnt, hmask, pairs);
                unit_pairs += pairs;
                if (hit) {
                    const uint32_t slot = __popcll(hmask & ((1ull << lane) - 1ull));
                    s_arm[slot] = arm;
                    s_end[slot] = incl;
                    if (QSET) s_hq[slot] = cq;
                    s_brm[slot] = brm;
                    s_ab[slot] = abase;
                    s_bb[slot] = bbase;
                }
                if (!rsplit || sub == 0) hits_n += hit;
                acols += ((uintptr_t)bbase & 1u) ? 0u : cnt;
                __syncthreads();
                const uint64_t t_pairs = __builtin_amdgcn_s_memtime();
                c_filter += t_pairs - t_filter;
                {
                    // 4. column-major gathers: work item gp is (hit, shared
                    //     column k).  The quarter's column word k (the rows
                    //     holding k) and B row k are loaded together, and B row
                    //     k is OR-ed into each of those rows.  No A row is
                    //     loaded for nothing, and no B row twice per hit.
                    for (uint32_t base = 0; base < pairs; base += WAVE * COL_ITEMS) {
                        // COL_ITEMS items per lane: all their loads first
                        uint64_t rows[COL_ITEMS];
                        Row4 b[COL_ITEMS];
                        uint32_t r0[COL_ITEMS];
#pragma unroll
                        for (int j = 0; j < COL_ITEMS; ++j) {
                            const uint32_t gp = base + (uint32_t)j * WAVE + lane;
                            rows[j] = 0;
                            b[j] = Row4{{0, 0, 0, 0}};
                            r0[j] = 0;
                            if (gp < pairs) {
                                uint32_t lo = 0, hi = nh - 1;  // owner: first slot whose range ends past gp
                                while (lo < hi) {
                                    const uint32_t mid = (lo + hi) >> 1;
                                    if (s_end[mid] > gp) hi = mid;
                                    else lo = mid + 1;
                                }
                                r0[j] = QSET ? s_hq[lo] * WAVE : 0u;
                                const uint64_t *Bt = s_bb[lo];
                                if ((uintptr_t)Bt & 1u) {  // full B tile: every row of the quarter becomes all ones
                                    for (uint64_t rr = s_arm[lo] & rowsel; rr; rr &= rr - 1) {
                                        const uint32_t r = r0[j] + (uint32_t)__builtin_ctzll(rr);
#pragma unroll
                                        for (int i = 0; i < 4; ++i) atomicOr((unsigned long long *)&acc_s[r].w[i], ~0ull);
                                    }
                                } else {
                                    const Row4 rel = s_brm[lo];
                                    const uint32_t k = select_bit256(rel, gp - (s_end[lo] - row_popc(rel)));
                                    rows[j] = gld(s_ab[lo] + k) & rowsel;
                                    b[j] = load_row(Bt + k * TROW_WORDS);
                                    ++brows;
                                }
                            }
                        }
#pragma unroll
                        for (int j = 0; j < COL_ITEMS; ++j) {
                            // one LDS atomic per row and word.  Quarter-set units
                            // (small graphs: B rows of a few bits) skip the zero
                            // words; single-quarter units OR them too, since the
                            // exec-mask branch around each atomic costs more there
                            // than the OR of a zero (both measured)
                            for (uint64_t rr = rows[j]; rr; rr &= rr - 1) {
                                const uint32_t r = r0[j] + (uint32_t)__builtin_ctzll(rr);
#pragma unroll
                                for (int i = 0; i < 4; ++i)
                                    if (!QSET || b[j].w[i])
                                        atomicOr((unsigned long long *)&acc_s[r].w[i], (unsigned long long)b[j].w[i]);
                            }
                        }
                    }
                }
                c_pairs += __builtin_amdgcn_s_memtime() - t_pairs;
                __syncthreads();  // the hit staging is refilled by the next candidate round
            }
            __syncthreads();  // the candidate staging is refilled by the next (term, word) chunk
        }
        __syncthreads();
        const uint64_t t_epi = __builtin_amdgcn_s_memtime();
        max_cands = max(max_cands, tp - tp_unit0);
        uint32_t pc = 0;  // new bits of the unit's quarters
        if (nsub > 1 && !rsplit) {
            // a sub-unit of a split quarter: its rows go to its partial-row
            // slot (its index u in the sub-unit list), non-empty ones only,
            // with their mask; k_merge_split finishes the quarter
            const uint32_t qq = (uint32_t)__builtin_ctz(qset);
            const Row4 acc = acc_s[(QSET ? qq * WAVE : 0u) + lane];
            const bool nz = row_nonzero(acc);
            if (nz) store_row(partials + ((size_t)u * WAVE + lane) * TROW_WORDS, acc);
            const uint64_t bal = __ballot(nz);
            if (lane == 0) gst(&pmask[u], bal);
        } else {
#pragma unroll
            for (uint32_t qq = 0; qq < QUARTERS; ++qq) {
                if (!((qset >> qq) & 1u)) continue;
                const Row4 acc = acc_s[(QSET ? qq * WAVE : 0u) + lane];
                pc += quarter_epilogue(O, M, p, qq, I, J, acc, fuse, out_rows, rsplit).pc;
            }
        }
        if (pc && lane == 0) {
            if (g_add(&O.cnt[p], pc) == 0) {
                set_occ(O, I, J);
                const uint32_t at = atomicAdd(&shards[home].n_touched, 1u);
                gst(&tl.list[(size_t)(1 + home) * tl.seg_cap + at], ((uint64_t)h << 40) | p);
            }
            new_bits += pc;
        }
        const uint64_t t_end = __builtin_amdgcn_s_memtime();
        const uint64_t tu = t_end - tu0;
        const uint32_t ucand = min(tp - tp_unit0, 511u), upairs = min(unit_pairs / 16u, 1023u);
        const uint32_t ugath_all = wave_sum(brows - br_unit0);
        const uint32_t ugath = min(ugath_all, 0xffffffu);
        unit_clk += tu;
        ++units_done;
        if (nsub == 1 && pc == 0) {  // (a split quarter's sub-units count as productive)
            idle_clk += tu;
            ++idle_units;
            idle_brows += ugath_all;
        }
        if (tu > unit_max) {
            unit_max = tu;
            unit_max_a = (tu << 24) | min(nsub, 31u) | (ucand << 5) | (upairs << 14);
            unit_max_b = (tu << 24) | ugath;
            auto cap12 = [](uint64_t v) -> uint64_t { return (v >> 8) < 4095u ? (v >> 8) : 4095u; };
            unit_max_c = (tu << 24) | (cap12(c_filter) << 12) | cap12(c_pairs);
            unit_max_d = (tu << 24) | (cap12(c_scan) << 12) | cap12(t_end - t_epi);
        }
        __syncthreads();  // the accumulator is reset by the next unit
    }
    }
    const uint32_t ar = wave_sum(arows), br = wave_sum(brows), orow = wave_sum(out_rows), hs = wave_sum(hits_n);
    const uint32_t ac = wave_sum(acols);
    if (lane == 0) {
        CounterShard &S = shards[home];
        if (tp) atomicAdd((unsigned long long *)&S.tile_products, (unsigned long long)tp);
        if (ar) atomicAdd((unsigned long long *)&S.a_rows, (unsigned long long)ar);
        if (ac) atomicAdd((unsigned long long *)&S.a_cols, (unsigned long long)ac);
        if (br) atomicAdd((unsigned long long *)&S.b_rows, (unsigned long long)br);
        if (orow) atomicAdd((unsigned long long *)&S.out_rows, (unsigned long long)orow);
        if (hs) atomicAdd((unsigned long long *)&S.cand_hits, (unsigned long long)hs);
        if (max_cands) atomicMax(&S.max_unit_cands, max_cands);
        if (new_bits) atomicAdd((unsigned long long *)&S.new_bits, (unsigned long long)new_bits);
        const uint64_t busy = __builtin_amdgcn_s_memtime() - clk0;
        atomicAdd((unsigned long long *)&S.busy, (unsigned long long)busy);
        atomicMax((unsigned long long *)&S.busy_max, (unsigned long long)busy);
        if (unit_max) atomicMax((unsigned long long *)&S.unit_max, (unsigned long long)unit_max);
        if (unit_max_a) atomicMax((unsigned long long *)&S.unit_max_a, (unsigned long long)unit_max_a);
        if (unit_max_b) atomicMax((unsigned long long *)&S.unit_max_b, (unsigned long long)unit_max_b);
        if (unit_max_c) atomicMax((unsigned long long *)&S.unit_max_c, (unsigned long long)unit_max_c);
        if (unit_max_d) atomicMax((unsigned long long *)&S.unit_max_d, (unsigned long long)unit_max_d);
        if (unit_clk) atomicAdd((unsigned long long *)&S.unit_clk, (unsigned long long)unit_clk);
        if (idle_clk) atomicAdd((unsigned long long *)&S.idle_clk, (unsigned long long)idle_clk);
        if (idle_units) atomicAdd((unsigned long long *)&S.idle_units, (unsigned long long)idle_units);
        if (idle_brows) atomicAdd((unsigned long long *)&S.idle_brows, (unsigned long long)idle_brows);
        if (units_done) atomicAdd((unsigned long long *)&S.units_done, (unsigned long long)units_done);
    }
}

// The split quarters of a step (k_build's list `splits`: base item with
// nsub, first partial-row slot): one wave per quarter ORs the partial rows of
// its sub-units (written by the sparse product, launched before) and writes
// the quarter as a single owner, like the unsplit units' epilogue.  Split
// units exist only without the fused head update.
__global__ __launch_bounds__(WAVE) void k_merge_split(const MatDev *__restrict__ mats,
                                                      const HeadDev *__restrict__ heads, int parity,
                                                      StepCounters *__restrict__ c, const uint64_t *__restrict__ splits,
                                                      const uint64_t *__restrict__ partials,
                                                      const uint64_t *__restrict__ pmask, TouchedLists tl) {
    const int lane = threadIdx.x & 63;
    const uint32_t n_split = c->n_split;
    // counts and touched-list entries go to this XCD's shard (6,000 split
    // quarters in a step would otherwise queue on one address)
    const uint32_t home = blockIdx.x % SHARDS;
    CounterShard *shard = &c->shard[home];
    uint32_t out_rows = 0;
    uint64_t new_bits = 0;
    for (uint32_t sq = blockIdx.x; sq < n_split; sq += gridDim.x) {
        const uint64_t it = uniform64(ld_dev(splits + 2 * (size_t)sq));
        const uint32_t at = (uint32_t)uniform64(ld_dev(splits + 2 * (size_t)sq + 1));
        const uint32_t h = uniform(item_head(it)), p = uniform(unit_tile(it)), q = uniform((uint32_t)(it & 3u));
        const uint32_t nsub = uniform(unit_nsub(it));
        // the sub-units' partial rows, four sub-units' loads in flight at a time
        Row4 acc = {{0, 0, 0, 0}};
        for (uint32_t s0 = 0; s0 < nsub; s0 += 4) {
            Row4 part[4];
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t sidx = s0 + j;
                const uint64_t m = sidx < nsub ? ld_dev(pmask + at + sidx) : 0ull;
                part[j] = ((m >> lane) & 1ull) ? load_row(partials + ((size_t)(at + sidx) * WAVE + lane) * TROW_WORDS)
                                               : Row4{{0, 0, 0, 0}};
            }
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) or_row(acc, part[j]);
        }
        const uint32_t T = ld_const(&heads[h].T), I = p / T, J = p - (p / T) * T;
        const MatDev O = mats[uniform(heads[h].out[parity])];
        const MatDev M = mats[uniform(heads[h].target)];
        const QuarterOut qo = quarter_epilogue(O, M, p, q, I, J, acc, 0, out_rows);
        if (qo.bal && lane == 0) {
            const uint32_t old = g_add(&O.cnt[p], qo.pc);
            if (old == 0) {
                set_occ(O, I, J);
                const uint32_t at = atomicAdd(&shard->n_touched, 1u);
                gst(&tl.list[(size_t)(1 + home) * tl.seg_cap + at], ((uint64_t)h << 40) | p);
            }
            new_bits += qo.pc;
        }
    }
    const uint32_t orow = wave_sum(out_rows);
    if (lane == 0) {
        if (orow) atomicAdd((unsigned long long *)&shard->out_rows, (unsigned long long)orow);
        if (new_bits) atomicAdd((unsigned long long *)&shard->new_bits, (unsigned long long)new_bits);
    }
}

// ---------------------------------------------------------------------------
// dense product: one 256-thread work-group per (head, C tile); thread r owns
// row r of the tile
// ---------------------------------------------------------------------------

// A tile with few bits: gather the B rows named by A's set bits from HBM/L2.
// A row whose accumulated bits together with the head row's known bits
// (`known`) are all ones can gain nothing more: its remaining lookups and
// gathers are skipped (the new bits are acc & ~M).
__device__ __forceinline__ bool row_saturated(const Row4 &acc, const Row4 &known) {
    return ((acc.w[0] | known.w[0]) & (acc.w[1] | known.w[1]) & (acc.w[2] | known.w[2]) & (acc.w[3] | known.w[3])) ==
           ~0ull;
}

__device__ __forceinline__ void tp_gather(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                          uint64_t rm, int r, Row4 &acc, const Row4 &known, uint32_t &arows,
                                          uint32_t &brows) {
    if (!((rm >> (r & 63)) & 1ull) || row_saturated(acc, known)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (row_saturated(acc, known)) break;
        uint64_t x = a.w[w];
        while (x) {
            const int k = w * 64 + __builtin_ctzll(x);
            x &= x - 1;
            or_row(acc, load_row(Bt + k * TROW_WORDS));
            ++brows;
        }
    }
}

// Heavier A tile: stage the B tile in LDS (8 KiB, one coalesced row per
// thread), then OR rows from LDS.  Must be reached by the whole work-group.
__device__ __forceinline__ void tp_lds(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                       uint64_t rm, int r, Row4 *Bs, Row4 &acc, const Row4 &known, uint32_t &arows,
                                       uint32_t &brows_lds) {
    // (the caller's broadcast before this call is the barrier that protects Bs)
    Bs[r] = load_row(Bt + r * TROW_WORDS);
    __syncthreads();
    if (!((rm >> (r & 63)) & 1ull) || row_saturated(acc, known)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (row_saturated(acc, known)) break;
        uint64_t x = a.w[w];
        while (x) {
            const int k = w * 64 + __builtin_ctzll(x);
            x &= x - 1;
            or_row(acc, Bs[k]);
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
                                       uint64_t rm, int r, Row4 *Bs, Row4 *Tab, Row4 &acc, const Row4 &known,
                                       uint32_t &arows, uint32_t &lookups) {
    // (no barrier before the B tile is staged: the caller's broadcast of this
    // tile-product's counts is one, after every lane's reads of the previous
    // tile-product's Bs and tables)
    Bs[r] = load_row(Bt + r * TROW_WORDS);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (M4R_GROUPS * 16) / BLOCK; ++q) {
        const int e = q * BLOCK + r;     // entry index = g * 16 + key
        const int g = e >> 4, key = e & 15;
        Row4 v = {{0, 0, 0, 0}};
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((key >> b) & 1) or_row(v, Bs[4 * g + b]);
        Tab[e] = v;
    }
    __syncthreads();
    if (!((rm >> (r & 63)) & 1ull) || row_saturated(acc, known)) return;
    ++arows;
    const Row4 a = load_row(At + r * TROW_WORDS);
    // Nibble positions in the same order in every lane: at each step the
    // lanes read from one 16-entry table (512 contiguous bytes, every bank
    // once, identical keys broadcast) instead of from tables scattered over
    // 32 KiB.  Zero nibbles are skipped per lane.
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        if (row_saturated(acc, known)) break;
        const uint64_t x = a.w[w];
#pragma unroll
        for (int nib = 0; nib < 16; ++nib) {
            const int key = (int)((x >> (4 * nib)) & 15u);
            if (key) {
                or_row(acc, Tab[(w * 16 + nib) * 16 + key]);
                ++lookups;
            }
        }
    }
}

// Very dense A tile, 8-bit keys, one table per wavefront: wave w owns the
// 64-column slice w of the C tile for all 256 rows (lane l: rows l + 64 j,
// accumulated in acc8[j]).  The B tile is staged in LDS; then, four groups of
// 8 B rows at a time, each wave builds the 256 OR-combinations of its slice of
// those rows (2 KiB per group, 8 KiB per wave, in the shared table space) and
// every row costs one 8-byte lookup per non-zero byte of A.  Must be reached
// by the whole work-group.  rmp: the A tile's four quarter row masks.
__device__ __forceinline__ void tp_m4r8(const uint64_t *__restrict__ At, const uint64_t *__restrict__ Bt,
                                        const uint64_t *__restrict__ rmp, int r, Row4 *Bs, uint64_t *T8,
                                        uint64_t *acc8, uint32_t &arows, uint32_t &lookups) {
    const int wv = r >> 6, lane = r & 63;
    __syncthreads();
    Bs[r] = load_row(Bt + r * TROW_WORDS);
    Row4 a[QUARTERS];
#pragma unroll
    for (int j = 0; j < QUARTERS; ++j) {
        a[j] = Row4{{0, 0, 0, 0}};
        if ((rmp[j] >> lane) & 1ull) a[j] = load_row(At + (j * WAVE + lane) * TROW_WORDS);
    }
    if (wv == 0) {
#pragma unroll
        for (int j = 0; j < QUARTERS; ++j) arows += (rmp[j] >> lane) & 1ull;
    }
    uint64_t *tw = T8 + (size_t)wv * 4 * 256;  // this wave's four 256-entry tables
    __syncthreads();
    for (int g0 = 0; g0 < TILE / 8; g0 += 4) {
#pragma unroll
        for (int gi = 0; gi < 4; ++gi) {
            const int g = g0 + gi;
            uint64_t b[8];
#pragma unroll
            for (int bb = 0; bb < 8; ++bb) b[bb] = Bs[8 * g + bb].w[wv];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int key = e * WAVE + lane;
                uint64_t v = 0;
#pragma unroll
                for (int bb = 0; bb < 8; ++bb)
                    if ((key >> bb) & 1) v |= b[bb];
                tw[gi * 256 + key] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < QUARTERS; ++j) {
#pragma unroll
            for (int gi = 0; gi < 4; ++gi) {
                const int g = g0 + gi;
                const uint32_t key = (uint32_t)(a[j].w[g >> 3] >> ((g & 7) * 8)) & 255u;
                if (key) {
                    acc8[j] |= tw[gi * 256 + key];
                    ++lookups;
                }
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(BLOCK) void k_product_dense(const MatDev *__restrict__ mats,
                                                         const HeadDev *__restrict__ heads,
                                                         const TermDev *__restrict__ terms, int parity,
                                                         StepCounters *__restrict__ c,
                                                         const uint64_t *__restrict__ items,
                                                         uint64_t *__restrict__ touched, uint32_t dense_bits,
                                                         uint32_t m4r_bits, int fuse, int m4r8,
                                                         const uint32_t *__restrict__ last_eval) {
    __shared__ Row4 Bs[TILE];
    __shared__ Row4 Tab[M4R_GROUPS * 16];
    __shared__ uint32_t s_idx;
    __shared__ uint32_t s_sum[4];
    __shared__ uint32_t s_bc[2];
    uint32_t nbc = 0;
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const uint32_t n_items = c->n_dense;
    if (blockIdx.x >= n_items) return;  // surplus work-groups would only queue on the cursor
    uint32_t tp = 0, tpd = 0, tpu = 0, arows = 0, brows = 0, brows_lds = 0, out_rows = 0, lookups = 0;
    uint64_t new_bits = 0;  // thread 0: new bits of this work-group's items
    for (uint32_t idx = wg_next(&c->dense_next, &s_idx); idx < n_items; idx = wg_next(&c->dense_next, &s_idx)) {
        const uint64_t it = items[idx];
        const uint32_t h = uniform(item_head(it));
        const uint32_t p = uniform((uint32_t)(it & ITEM_TILE_MASK));
        const bool full = (it & ITEM_FULL) != 0;
        const uint32_t T = uniform(heads[h].T);
        const uint32_t t0 = uniform(full ? heads[h].f0 : heads[h].t0), t1 = uniform(full ? heads[h].f1 : heads[h].t1);
        const uint32_t I = p / T, J = p - (p / T) * T;
        // sweeps: the step that last evaluated this entry (0: none in this
        // call); a tile-product whose two operand tiles have not changed since
        // then was evaluated then, with the same operands, and is skipped
        const uint32_t last = last_eval ? uniform(ld_dev(&last_eval[h])) : 0u;
        const MatDev M = mats[uniform(heads[h].target)];
        // the head's row r as it is now (a subset of it later, with the head
        // update fused): the bits this item cannot add
        const Row4 known = load_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS);
        Row4 acc = {{0, 0, 0, 0}};
        uint64_t acc8[QUARTERS] = {0, 0, 0, 0};  // 8-bit Four Russians: slice wv of rows lane + 64 j
        bool used8 = false;
        for (uint32_t t = t0; t < t1; ++t) {
            const MatDev L = mats[uniform(terms[t].left[parity])];
            const MatDev R = mats[uniform(terms[t].right[parity])];
            for (uint32_t w = 0; w < L.ow; ++w) {
                uint32_t kb = wg_bcast(L.occR[I * L.ow + w] & R.occC[J * R.ow + w], s_bc, nbc);
                while (kb) {
                    const uint32_t K = w * 32 + __builtin_ctz(kb);
                    kb &= kb - 1;
                    const uint32_t lt = I * T + K, rt = K * T + J;
                    const uint64_t *At = L.bits + (size_t)lt * TILE_WORDS;
                    const uint64_t *Bt = R.bits + (size_t)rt * TILE_WORDS;
                    const uint64_t rm = L.rmask[(size_t)lt * 4 + wv];
                    // A-tile count, bit 31 = "B tile full", bit 30 = "both
                    // operand tiles unchanged since the last evaluation" (counts are <= 2^16)
                    const uint32_t ab = wg_bcast(L.cnt[lt] | (R.cnt[rt] == TILE_BITS ? 0x80000000u : 0u) |
                                                     (last && L.stamp[lt] < last && R.stamp[rt] < last ? 0x40000000u : 0u),
                                                 s_bc, nbc);
                    if (ab & 0x40000000u) {
                        ++tpu;
                        continue;
                    }
                    ++tp;
                    const uint32_t acnt = ab & 0x3fffffffu;
                    if (ab >> 31) {
                        // B all ones: every non-empty A row gives an all-ones row
                        if ((rm >> (r & 63)) & 1ull) acc = Row4{{~0ull, ~0ull, ~0ull, ~0ull}};
                        ++tpd;
                        continue;
                    }
                    if (acnt == TILE_BITS) {
                        // A all ones: every row gets the OR of all B rows
#pragma unroll
                        for (int qq = 0; qq < QUARTERS; ++qq) or_row(acc, load_row(colq_of(R, rt, qq)));
                        ++tpd;
                        continue;
                    }
                    if (acnt >= m4r_bits && m4r8) {
                        ++tpd;
                        used8 = true;
                        tp_m4r8(At, Bt, L.rmask + (size_t)lt * 4, r, Bs, (uint64_t *)Tab, acc8, arows, lookups);
                    } else if (acnt >= m4r_bits) {
                        ++tpd;
                        tp_m4r(At, Bt, rm, r, Bs, Tab, acc, known, arows, lookups);
                    } else if (acnt >= dense_bits) {
                        ++tpd;
                        tp_lds(At, Bt, rm, r, Bs, acc, known, arows, brows_lds);
                    } else {
                        tp_gather(At, Bt, rm, r, acc, known, arows, brows);
                    }
                }
            }
        }
        if (used8) {  // (uniform: every thread of the group takes the same paths)
            // the slice accumulators back into rows: wave w holds word w of rows lane + 64 j
            __syncthreads();
#pragma unroll
            for (int j = 0; j < QUARTERS; ++j) Bs[j * WAVE + lane].w[wv] = acc8[j];
            __syncthreads();
            or_row(acc, Bs[r]);
        }
        const MatDev O = mats[uniform(heads[h].out[parity])];
        const QuarterOut qo = quarter_epilogue(O, M, p, wv, I, J, acc, fuse, out_rows);
        if (lane == 0) s_sum[wv] = qo.pc;
        __syncthreads();
        if (r == 0) {
            const uint32_t total = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
            if (total) {
                O.cnt[p] = total;
                set_occ(O, I, J);
                touched[atomicAdd(&c->n_touched, 1u)] = it;
                new_bits += total;
            }
        }
    }
    const uint32_t ar = wave_sum(arows), br = wave_sum(brows), orow = wave_sum(out_rows);
    const uint32_t brl = wave_sum(brows_lds), lk = wave_sum(lookups);
    if (lane == 0) {
        if (ar) atomicAdd((unsigned long long *)&c->a_rows_dense, (unsigned long long)ar);
        if (br) atomicAdd((unsigned long long *)&c->b_rows_dense, (unsigned long long)br);
        if (orow) atomicAdd((unsigned long long *)&c->out_rows_dense, (unsigned long long)orow);
        if (brl) atomicAdd((unsigned long long *)&c->b_rows_lds, (unsigned long long)brl);
        if (lk) atomicAdd((unsigned long long *)&c->m4r_lookups, (unsigned long long)lk);
    }
    if (r == 0 && tp) {
        atomicAdd((unsigned long long *)&c->tile_products, (unsigned long long)tp);
        atomicAdd((unsigned long long *)&c->tp_dense, (unsigned long long)tpd);
    }
    if (r == 0 && tpu) atomicAdd((unsigned long long *)&c->tp_unchanged, (unsigned long long)tpu);
    if (r == 0 && new_bits) atomicAdd((unsigned long long *)&c->new_bits, (unsigned long long)new_bits);
}

// ---------------------------------------------------------------------------
// MFMA product: one 512-thread work-group (8 waves) per (head, C tile)
// ---------------------------------------------------------------------------
//
// C tile (256 x 256) = sum over terms and k-tiles of A tile * B tile, as an
// exact integer product of 0/1 int8 operands with int32 accumulation
// (v_mfma_i32_32x32x32_i8), then "> 0".  Wave w owns rows 64*(w%4) .. +63 and
// columns 128*(w/4) .. +127 of the C tile: 2 x 4 blocks of 32 x 32, 128
// accumulator registers per lane, kept across all k-tiles of the item.
//
// Operands are expanded from the bit-packed tiles on the fly:
//   * A: lane l of a 32-row block holds row l%32, and the 16 consecutive k of
//     half (l/32) of the current 32-k block, one byte per bit;
//   * B: the B tile is staged in LDS, each wave transposes two 64 x 64 bit
//     blocks across its lanes (six butterfly stages), and writes the columns
//     as bytes to BX[column][k] (64 KiB; 16-byte chunks XOR-swizzled by the
//     column so that a 32-lane fragment read touches distinct banks).
// A and B bytes are placed by the same (lane/32, byte) -> k rule, which is all
// the matrix core needs for the contraction; the result layout is the
// documented 32 x 32 one: lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4,
// column l%32.
constexpr int MFMA_BLOCK = 512;
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));

// bit e of x (e < 4) -> byte e (0 or 1): the four shifted copies x << 7e do
// not overlap, so the product has no carries.
__device__ __forceinline__ int32_t bytes_of_nibble(uint32_t x) {
    return (int32_t)(((x & 15u) * 0x00204081u) & 0x01010101u);
}
__device__ __forceinline__ i32x4 bytes_of_16(uint32_t x) {
    i32x4 r;
    r[0] = bytes_of_nibble(x);
    r[1] = bytes_of_nibble(x >> 4);
    r[2] = bytes_of_nibble(x >> 8);
    r[3] = bytes_of_nibble(x >> 12);
    return r;
}

// word index of the 16-byte chunk `chunk` (0..15, along k) of column n in BX
__device__ __forceinline__ int bx_index(int n, int chunk) { return n * 64 + ((chunk ^ (n & 15)) << 2); }

__global__ __launch_bounds__(MFMA_BLOCK) void k_product_mfma(const MatDev *__restrict__ mats,
                                                             const HeadDev *__restrict__ heads,
                                                             const TermDev *__restrict__ terms, int parity,
                                                             StepCounters *__restrict__ c,
                                                             const uint64_t *__restrict__ items,
                                                             uint64_t *__restrict__ touched, int fuse) {
    __shared__ __attribute__((aligned(16))) uint32_t BX[TILE * TILE / 4];  // 64 KiB
    __shared__ Row4 Bs[TILE];  // B tile bits; the result bits in the epilogue
    __shared__ uint32_t s_cand[WAVE * 32];  // candidate k-tiles: (pair index << 16) | K
    __shared__ uint32_t s_ncand;
    __shared__ uint32_t s_idx;
    __shared__ uint32_t s_sum[4];
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
    const int rs = wv & 3, ch = wv >> 2;
    const uint32_t n_items = c->n_mfma;
    if (blockIdx.x >= n_items) return;  // surplus work-groups would only queue on the cursor
    uint32_t tp = 0, instr = 0, out_rows = 0;
    uint64_t new_bits = 0;  // thread 0: new bits of this work-group's items
    for (uint32_t idx = wg_next(&c->mfma_next, &s_idx); idx < n_items; idx = wg_next(&c->mfma_next, &s_idx)) {
        const uint64_t it = items[idx];
        const uint32_t h = uniform(item_head(it));
        const uint32_t p = uniform((uint32_t)(it & ITEM_TILE_MASK));
        const bool full = (it & ITEM_FULL) != 0;
        const uint32_t T = uniform(heads[h].T);
        const uint32_t t0 = uniform(full ? heads[h].f0 : heads[h].t0), t1 = uniform(full ? heads[h].f1 : heads[h].t1);
        const uint32_t I = p / T, J = p - (p / T) * T;
        i32x16 acc[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) acc[rb][cb] = i32x16{};
        // The item's candidate k-tiles, 64 (term, occupancy word) pairs at a
        // time: wave 0 reads the masks once and lays the candidates out in LDS,
        // so every wave walks the same list (the head matrices may be updated
        // by other work-groups meanwhile), and the next tile's A and B rows can
        // be fetched while the current one is transposed and multiplied.
        const uint32_t ow = (T + 31) / 32, ntw = (t1 - t0) * ow;
        for (uint32_t tw0 = 0; tw0 < ntw; tw0 += WAVE) {
            __syncthreads();  // the previous chunk's list is no longer read
            if (wv == 0) {
                const uint32_t j = tw0 + lane;
                uint32_t kbw = 0;
                if (j < ntw) {
                    const uint32_t tt = t0 + j / ow, w = j % ow;
                    kbw = mats[terms[tt].left[parity]].occR[I * ow + w] & mats[terms[tt].right[parity]].occC[J * ow + w];
                }
                const uint32_t incl = wave_scan_incl((uint32_t)__popc(kbw));
                uint32_t pos = incl - (uint32_t)__popc(kbw);
                for (uint32_t x = kbw; x; x &= x - 1)
                    s_cand[pos++] = ((j - tw0) << 16) | ((j % ow) * 32 + (uint32_t)__builtin_ctz(x));
                if (lane == 63) s_ncand = incl;
            }
            __syncthreads();
            const uint32_t ncand = s_ncand;
            if (!ncand) continue;
            auto tile_ptrs = [&](uint32_t e, const uint64_t *&At, const uint64_t *&Bt) {
                const uint32_t jj = tw0 + (e >> 16), tt = t0 + jj / ow, K = e & 0xffffu;
                At = mats[terms[tt].left[parity]].bits + (size_t)(I * T + K) * TILE_WORDS;
                Bt = mats[terms[tt].right[parity]].bits + (size_t)(K * T + J) * TILE_WORDS;
            };
            const uint64_t *At, *Bt;
            tile_ptrs(s_cand[0], At, Bt);
            Row4 bcur = t < TILE ? load_row(Bt + t * TROW_WORDS) : Row4{{0, 0, 0, 0}};
            Row4 a0 = load_row(At + (rs * 64 + (lane & 31)) * TROW_WORDS);
            Row4 a1 = load_row(At + (rs * 64 + 32 + (lane & 31)) * TROW_WORDS);
            for (uint32_t ci = 0; ci < ncand; ++ci) {
                ++tp;
                __syncthreads();  // the previous k-tile's readers of Bs / BX are done
                if (t < TILE) Bs[t] = bcur;
                __syncthreads();
                // fetch the next tile while this one is transposed and multiplied
                Row4 bnext = {{0, 0, 0, 0}}, a0n = {{0, 0, 0, 0}}, a1n = {{0, 0, 0, 0}};
                if (ci + 1 < ncand) {
                    tile_ptrs(s_cand[ci + 1], At, Bt);
                    if (t < TILE) bnext = load_row(Bt + t * TROW_WORDS);
                    a0n = load_row(At + (rs * 64 + (lane & 31)) * TROW_WORDS);
                    a1n = load_row(At + (rs * 64 + 32 + (lane & 31)) * TROW_WORDS);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int blk = 2 * wv + i, kb64 = blk >> 2, cb64 = blk & 3;
                    const uint64_t col = wave_transpose64(Bs[kb64 * 64 + lane].w[cb64]);
                    const int n = cb64 * 64 + lane;  // this lane's B column; bits = k in kb64*64 ..
#pragma unroll
                    for (int cq = 0; cq < 4; ++cq)
                        *reinterpret_cast<i32x4 *>(&BX[bx_index(n, kb64 * 4 + cq)]) =
                            bytes_of_16((uint32_t)(col >> (16 * cq)) & 0xFFFFu);
                }
                __syncthreads();
#pragma unroll
                for (int k32 = 0; k32 < 8; ++k32) {
                    const int sh = (k32 & 1) * 32 + (lane >> 5) * 16;
                    const uint32_t x0 = (uint32_t)(a0.w[k32 >> 1] >> sh) & 0xFFFFu;
                    const uint32_t x1 = (uint32_t)(a1.w[k32 >> 1] >> sh) & 0xFFFFu;
                    const bool any0 = __ballot(x0 != 0) != 0, any1 = __ballot(x1 != 0) != 0;
                    if (!any0 && !any1) continue;
                    const i32x4 fa0 = bytes_of_16(x0), fa1 = bytes_of_16(x1);
                    const int chunk = k32 * 2 + (lane >> 5);
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        const int n = ch * 128 + cb * 32 + (lane & 31);
                        const i32x4 fb = *reinterpret_cast<const i32x4 *>(&BX[bx_index(n, chunk)]);
                        if (any0) acc[0][cb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa0, fb, acc[0][cb], 0, 0, 0);
                        if (any1) acc[1][cb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa1, fb, acc[1][cb], 0, 0, 0);
                    }
                    instr += 4 * ((uint32_t)any0 + (uint32_t)any1);
                }
                bcur = bnext;
                a0 = a0n;
                a1 = a1n;
            }
        }
        // "> 0" per accumulator register, packed by ballots: register r of block
        // (rb, cb) gives columns cb*32 .. +31 of rows m = 8*(r/4) + r%4 (lanes
        // 0..31) and m + 4 (lanes 32..63).  Piece 2k / 2k+1 of ballot k goes to
        // lane (2k)%64 / (2k+1)%64 of pieces[(2k)/64].
        uint32_t pieces[4] = {0, 0, 0, 0};
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint64_t b = __ballot(acc[rb][cb][r] > 0);
                    const int k = (rb * 4 + cb) * 16 + r;
                    if (lane == ((2 * k) & 63)) pieces[(2 * k) >> 6] = (uint32_t)b;
                    if (lane == ((2 * k + 1) & 63)) pieces[(2 * k + 1) >> 6] = (uint32_t)(b >> 32);
                }
        __syncthreads();  // every wave is past its last read of Bs
        uint32_t *Cs = reinterpret_cast<uint32_t *>(Bs);  // [256 rows][8 dwords]
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int P = 64 * j + lane, k = P >> 1, half = P & 1;
            const int rb = k >> 6, cb = (k >> 4) & 3, r = k & 15;
            const int row = rs * 64 + rb * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
            Cs[row * 8 + ch * 4 + cb] = pieces[j];
        }
        __syncthreads();
        if (t < TILE) {
            const Row4 accb = Bs[t];
            const MatDev O = mats[uniform(heads[h].out[parity])];
            const MatDev M = mats[uniform(heads[h].target)];
            const QuarterOut qo = quarter_epilogue(O, M, p, wv, I, J, accb, fuse, out_rows);
            if (lane == 0) s_sum[wv] = qo.pc;
        }
        __syncthreads();
        if (t == 0) {
            const uint32_t total = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
            if (total) {
                const MatDev O = mats[heads[h].out[parity]];
                O.cnt[p] = total;
                set_occ(O, I, J);
                touched[atomicAdd(&c->n_touched, 1u)] = it;
                new_bits += total;
            }
        }
    }
    const uint32_t orow = wave_sum(out_rows);
    if (lane == 0) {
        if (orow) atomicAdd((unsigned long long *)&c->out_rows_dense, (unsigned long long)orow);
        if (instr) atomicAdd((unsigned long long *)&c->mfma_instr, (unsigned long long)instr);
    }
    if (t == 0 && tp) {
        atomicAdd((unsigned long long *)&c->tile_products, (unsigned long long)tp);
        atomicAdd((unsigned long long *)&c->tp_mfma, (unsigned long long)tp);
    }
    if (t == 0 && new_bits) atomicAdd((unsigned long long *)&c->new_bits, (unsigned long long)new_bits);
}

// ---------------------------------------------------------------------------
// apply / clear
// ---------------------------------------------------------------------------

// Merge the new rows of one touched tile (head h, tile p: entry `it`) into the
// head, whole wave, all four quarters at once: the tile's row masks and column
// unions (delta and head) come in one round of scalar loads, then every row
// either quarter needs in one round of vector loads (lane r: row 64 q + r of
// quarter q), so a tile costs a few dependent round trips instead of a few per
// quarter.  The column-major form of every quarter with new rows is computed
// here for the delta (from its new rows) and OR-ed into the head's, with or
// without `merge` (merge = 0: the product kernels already merged rows and
// metadata into the head).  Each (head, tile) is touched once per step, so
// nothing else in the launch writes the head tile.
__device__ __forceinline__ void apply_tile(const MatDev *__restrict__ mats, const HeadDev *__restrict__ heads,
                                           int parity, uint64_t it, int merge, uint32_t stamp_step) {
    const int lane = threadIdx.x & 63;
    const uint32_t h = uniform(item_head(it));
    const uint32_t p = uniform((uint32_t)(it & ITEM_TILE_MASK));
    const MatDev M = mats[uniform(heads[h].target)];
    const MatDev O = mats[uniform(heads[h].out[parity])];
    const uint32_t T = uniform(heads[h].T), I = p / T, J = p - (p / T) * T;
    const bool keepT = uniform(heads[h].colT) != 0;
    // the tile's row masks and column unions, delta and head: 40 words,
    // one per lane in a single vector load (per-lane addresses, so never the
    // scalar cache, which may hold them from an earlier launch), then read
    // into scalar registers lane by lane
    const uint64_t *src = lane < 4    ? O.rmask + (size_t)p * 4 + lane
                          : lane < 8  ? M.rmask + (size_t)p * 4 + (lane - 4)
                          : lane < 24 ? O.colq + (size_t)p * QUARTERS * TROW_WORDS + (lane - 8)
                                      : M.colq + (size_t)p * QUARTERS * TROW_WORDS + (lane - 24);
    const uint64_t meta = lane < 40 ? gld(src) : 0ull;
    Row4 rmO, rmM, ocu[QUARTERS], mcu[QUARTERS];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        rmO.w[w] = lane_read64(meta, (uint32_t)w);
        rmM.w[w] = lane_read64(meta, 4u + (uint32_t)w);
    }
#pragma unroll
    for (int q = 0; q < QUARTERS; ++q)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            ocu[q].w[w] = lane_read64(meta, 8u + (uint32_t)(q * 4 + w));
            mcu[q].w[w] = lane_read64(meta, 24u + (uint32_t)(q * 4 + w));
        }
    if (merge && lane == 0) {
        const uint32_t old = gld(&M.cnt[p]);
        gst(&M.cnt[p], old + gld(&O.cnt[p]));
        if (!old) set_occ(M, I, J);
    }
    // sweeps: the head tile gained bits in this step
    if (stamp_step && lane == 0) gst(&M.stamp[p], stamp_step);
    // the head's rows: with merge, its non-empty rows before the merge and the
    // new ones; without, the product kernels merged them already
    Row4 nw[QUARTERS], m[QUARTERS];
#pragma unroll
    for (int q = 0; q < QUARTERS; ++q) {
        const uint64_t rm = rmO.w[q];
        const uint64_t need = rm ? (keepT ? (rmM.w[q] | rm) : 0ull) | (merge ? rm : 0ull) : 0ull;
        const size_t off = (size_t)p * TILE_WORDS + (q * WAVE + lane) * TROW_WORDS;
        nw[q] = ((rm >> lane) & 1ull) ? load_row(O.bits + off) : Row4{{0, 0, 0, 0}};
        m[q] = ((need >> lane) & 1ull) ? load_row(M.bits + off) : Row4{{0, 0, 0, 0}};
    }
#pragma unroll
    for (int q = 0; q < QUARTERS; ++q) {
        const uint64_t rm = rmO.w[q];
        if (!rm) continue;
        if (merge && ((rm >> lane) & 1ull)) {
            or_row(m[q], nw[q]);
            store_row(M.bits + (size_t)p * TILE_WORDS + (q * WAVE + lane) * TROW_WORDS, m[q]);
        }
        if (keepT) {
            // column words of the columns that gained bits: the delta's from its
            // rows, the head's recomputed from all of its rows (a head that was
            // cleared and filled again may hold stale words for columns that
            // left its colq, so they are rewritten, never OR-ed into)
            quarter_transpose(nw[q], ocu[q], colT_of(O, p, q));
            quarter_transpose(m[q], ocu[q], colT_of(M, p, q));
        }
        if (merge && lane == 0) {
            gst(&M.rmask[(size_t)p * 4 + q], rmM.w[q] | rm);
            Row4 cu = mcu[q];
            or_row(cu, ocu[q]);
            store_row(colq_of(M, p, q), cu);
            if (!rmM.w[q]) set_occq(M, q, I, J);
        }
    }
}

// Zero one delta tile (entry `it` of the list written by the step of parity
// `parity`), whole wave: its rows named by the four row masks, then its
// metadata and count.
__device__ __forceinline__ void clear_tile(const MatDev *__restrict__ mats, const HeadDev *__restrict__ heads,
                                           int parity, uint64_t it) {
    const int lane = threadIdx.x & 63;
    const uint32_t h = uniform(item_head(it));
    const uint32_t p = uniform((uint32_t)(it & ITEM_TILE_MASK));
    const MatDev O = mats[uniform(heads[h].out[parity])];
    const uint64_t rml = lane < 4 ? gld(O.rmask + (size_t)p * 4 + lane) : 0ull;  // (per-lane: vector path)
    const Row4 rm = Row4{{lane_read64(rml, 0), lane_read64(rml, 1), lane_read64(rml, 2), lane_read64(rml, 3)}};
#pragma unroll
    for (int q = 0; q < QUARTERS; ++q) {
        if ((rm.w[q] >> lane) & 1ull)
            store_row(O.bits + (size_t)p * TILE_WORDS + (q * WAVE + lane) * TROW_WORDS, Row4{{0, 0, 0, 0}});
        if (lane == 0 && rm.w[q]) {
            gst(&O.rmask[(size_t)p * 4 + q], (uint64_t)0);
            store_row(colq_of(O, p, q), Row4{{0, 0, 0, 0}});
        }
    }
    if (lane == 0) gst(&O.cnt[p], 0u);
}

// Zero the occupancy bitsets (all four kinds) of one matrix, grid-strided.
__device__ __forceinline__ void zero_occ(const MatDev &O, uint32_t x0, uint32_t stride) {
    const uint32_t words = O.T * O.ow;
    for (uint32_t x = x0; x < words; x += stride) {
        O.occR[x] = 0;
        O.occC[x] = 0;
        O.occD[x] = 0;
        O.occX[x] = 0;
#pragma unroll
        for (int q = 0; q < QUARTERS; ++q) O.occQ[q * words + x] = 0;
    }
}

// Sweep stamps of matrices ids[0..nid) to 0 ("changed before this call").
__global__ void k_zero_stamps(const MatDev *__restrict__ mats, const uint32_t *__restrict__ ids, uint32_t nid) {
    for (uint32_t m = blockIdx.y; m < nid; m += gridDim.y) {
        const MatDev M = mats[ids[m]];
        const uint32_t tiles = M.T * M.T;
        for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < tiles; p += gridDim.x * blockDim.x) M.stamp[p] = 0;
    }
}

__global__ void k_clear_occ(const MatDev *__restrict__ mats, const uint32_t *__restrict__ ids, uint32_t nid) {
    for (uint32_t q = blockIdx.y; q < nid; q += gridDim.y)
        zero_occ(mats[ids[q]], blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The end of a step in one launch, its parts on disjoint arrays:
//   blocks [0, nw)             one wave per tile, strided over this step's
//                              touched tiles (parity `parity`, list `tl`,
//                              counts in c; with `apply`: their new rows merged
//                              into the heads unless the products did so
//                              already (merge = 0), and the column-major form
//                              of every new quarter; sweeps stamp the head
//                              tiles with stamp_step) and then the tiles of the
//                              delta the step consumed (with `clear`: parity
//                              1 - parity, the previous step's list `ptl`,
//                              counts in prev; rows and metadata zeroed);
//   blocks [nw, nw + nid * per) zero that delta's occupancy bitsets (ids[0..nid)),
//                              `per` blocks per matrix;
//   the last block             zero the counters of the next ring slot, so that
//                              the next step starts without a launch of its own.
// Block 0 also records stamp_step as plan entry `entry`'s last evaluation (sweeps).
__global__ __launch_bounds__(WAVE) void k_retire(const MatDev *__restrict__ mats,
                                                 const HeadDev *__restrict__ heads, int parity,
                                                 const StepCounters *__restrict__ c, TouchedLists tl,
                                                 const StepCounters *__restrict__ prev, TouchedLists ptl,
                                                 const uint32_t *__restrict__ ids, uint32_t nid, uint32_t per,
                                                 uint32_t nw, int apply, int clear, StepCounters *__restrict__ next,
                                                 int merge, uint32_t stamp_step, int entry,
                                                 uint32_t *__restrict__ last_eval) {
    const uint32_t b = blockIdx.x;
    // sweeps: this step evaluated plan entry `entry` (its products have run)
    if (b == 0 && threadIdx.x == 0 && last_eval && entry >= 0) gst(&last_eval[entry], stamp_step);
    if (b < nw) {
        const uint32_t n_a = apply ? touched_total(c) : 0u;
        const uint32_t n_c = clear ? touched_total(prev) : 0u;
        for (uint32_t i = b; i < n_a + n_c; i += nw) {  // static: the tiles are small and alike
            if (i < n_a) apply_tile(mats, heads, parity, touched_entry(c, tl, i), merge, stamp_step);
            else clear_tile(mats, heads, 1 - parity, touched_entry(prev, ptl, i - n_a));
        }
    } else if (b < nw + nid * per) {
        const uint32_t o = b - nw, m = o / per;
        zero_occ(mats[ids[m]], (o - m * per) * WAVE + threadIdx.x, per * WAVE);
    } else if (next) {
        uint64_t *w = reinterpret_cast<uint64_t *>(next);
        for (uint32_t i = threadIdx.x; i < sizeof(StepCounters) / 8; i += blockDim.x) gst(w + i, (uint64_t)0);
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
                // the column words of the quarter's columns go too: a word is
                // never more than its column's rows, also after a clear
                const Row4 cq = load_row(colq_of(O, p, wv));
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    if ((cq.w[w] >> lane) & 1ull) colT_of(O, p, wv)[w * 64 + lane] = 0;
                __syncthreads();
                if (lane == 0) {
                    O.rmask[(size_t)p * 4 + wv] = 0;
                    store_row(colq_of(O, p, wv), Row4{{0, 0, 0, 0}});
                }
                if (r == 0) O.cnt[p] = 0;
                __syncthreads();  // keeps the divergent tail apart from the next tile (see wg_next)
            }
        }
        zero_occ(O, blockIdx.x * BLOCK + r, gridDim.x * BLOCK);
    }
}

// dst := src (same T) for the matrix pairs (dst_ids[q], src_ids[q]); only
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
            if (r < QUARTERS * TROW_WORDS) D.colq[(size_t)p * QUARTERS * TROW_WORDS + r] = S.colq[(size_t)p * QUARTERS * TROW_WORDS + r];
            if (r == 0) D.cnt[p] = cn;
            // the column-major words of the columns each quarter holds (thread r: column r)
#pragma unroll
            for (uint32_t qq = 0; qq < QUARTERS; ++qq)
                if ((S.colq[((size_t)p * QUARTERS + qq) * TROW_WORDS + (r >> 6)] >> (r & 63)) & 1ull)
                    colT_of(D, p, qq)[r] = colT_of(S, p, qq)[r];
        }
        const uint32_t words = S.T * S.ow;
        for (uint32_t x = blockIdx.x * BLOCK + r; x < words; x += gridDim.x * BLOCK) {
            D.occR[x] = S.occR[x];
            D.occC[x] = S.occC[x];
            D.occD[x] = S.occD[x];
            D.occX[x] = S.occX[x];
#pragma unroll
            for (int qq = 0; qq < QUARTERS; ++qq) D.occQ[qq * words + x] = S.occQ[qq * words + x];
        }
    }
}

// ---------------------------------------------------------------------------
// layout conversion and metadata
// ---------------------------------------------------------------------------

// Recompute cnt / rmask / occupancy of one tile from its bits (whole WG).
__device__ __forceinline__ void tile_meta(const MatDev &M, uint32_t p, const Row4 &v, uint32_t *s_sum) {
    const int r = threadIdx.x, wv = r >> 6, lane = r & 63;
    const uint64_t bal = __ballot(row_nonzero(v));
    const uint32_t pc = wave_sum(row_popc(v));
    const Row4 cu = wave_or_row(v);
    const uint32_t I = p / M.T, J = p - (p / M.T) * M.T;
    if (lane == 0) {
        M.rmask[(size_t)p * 4 + wv] = bal;
        store_row(colq_of(M, p, wv), cu);
        if (bal) set_occq(M, wv, I, J);
        s_sum[wv] = pc;
    }
    // every column word of the quarter, zeros included (the tile may have held
    // other bits before)
    quarter_transpose(v, Row4{{~0ull, ~0ull, ~0ull, ~0ull}}, colT_of(M, p, wv));
    __syncthreads();
    if (r == 0) {
        const uint32_t total = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
        M.cnt[p] = total;
        if (total) set_occ(M, I, J);
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

// Rebuild the column-major form of every quarter of M, every word (one wave
// per quarter, four per work-group).
__global__ __launch_bounds__(BLOCK) void k_rebuild_colT(MatDev M) {
    const int r = threadIdx.x, wv = r >> 6;
    const uint32_t tiles = M.T * M.T;
    for (uint32_t p = blockIdx.x; p < tiles; p += gridDim.x) {
        const Row4 v = uniform64(M.rmask[(size_t)p * 4 + wv]) ? load_row(M.bits + (size_t)p * TILE_WORDS + r * TROW_WORDS)
                                                               : Row4{{0, 0, 0, 0}};
        quarter_transpose(v, Row4{{~0ull, ~0ull, ~0ull, ~0ull}}, colT_of(M, p, wv));
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
