// Device-side layout of a bit-packed Boolean matrix (gfx950).
//
// An n x n matrix is padded to T = ceil(n/256) tiles per side.  Tile (I, J)
// covers rows [256I, 256I+256) and columns [256J, 256J+256) and is stored as
// 256 rows x 4 uint64 words (8 KiB, contiguous): word w of tile row r holds
// columns 256J + 64w .. 256J + 64w + 63, bit b = column 256J + 64w + b.
// Tiles are stored row-major by (I, J).  A tile's rows fall in four
// "quarters" of 64 rows, one per wavefront.  Per tile the engine keeps
//   cnt[I*T+J]          popcount of the tile,
//   rmask[(I*T+J)*4+q]  bit r%64 of word q = r/64 set iff tile row r != 0,
//   colq[((I*T+J)*4+q)*4 .. +4]  OR of the rows of quarter q (256 column bits),
//   colT[((I*T+J)*4+q)*256 + k]  the quarter in column-major form: bit r set iff
//                       row 64q + r holds column k.  A word never holds more
//                       than its column's rows (a cleared quarter's words are
//                       zeroed), and between fixpoint steps the words of the
//                       columns in colq are exact; readers read only those,
// and occupancy bitsets over tile rows / columns (ow = ceil(T/32) words each):
//   occR[I*ow + K/32]           bit K%32 <=> tile (I, K) non-empty,
//   occC[J*ow + I/32]           bit I%32 <=> tile (I, J) non-empty,
//   occQ[(q*T + I)*ow + K/32]   bit K%32 <=> quarter q of tile (I, K) non-empty,
//   occD[I*ow + K/32]           bit K%32 <=> tile (I, K) holds at least the
//                               dense-path threshold of bits (rebuilt per step),
//   occX[I*ow + K/32]           the same for the MFMA threshold.
// Bits at rows or columns >= n are always zero.
#pragma once
#include <stdint.h>

namespace cfpq {

constexpr int TILE = 256;                 // tile edge in bits
constexpr int TROW_WORDS = 4;             // uint64 words per tile row
constexpr int TILE_WORDS = TILE * TROW_WORDS;
constexpr uint32_t TILE_BITS = TILE * TILE;
constexpr int BLOCK = 256;                // threads per dense work-group: one per tile row
constexpr int WAVE = 64;                  // gfx950 wavefront
constexpr int QUARTERS = TILE / WAVE;     // 64-row quarters per tile
constexpr int OCC_ARRAYS = 2 + QUARTERS + 2;  // occR, occC, occQ[4], occD, occX

struct MatDev {
    uint64_t *bits;
    uint32_t *cnt;
    uint64_t *rmask;
    uint64_t *colq;
    uint64_t *colT;
    uint32_t *occR;
    uint32_t *occC;
    uint32_t *occQ;
    uint32_t *occD;
    uint32_t *occX;
    // per tile: the sweep step (counted from 1 within a fixpoint call) in
    // which the tile last gained bits as a rule head; written and read by
    // head and rule sweeps only (any value is safe for a tile that did not
    // change during the call)
    uint32_t *stamp;
    uint32_t T;
    uint32_t ow;
};

// A group of rules sharing one head: the head's full matrix (mask), the
// buffer new bits go to on even / odd steps, its semi-naive terms [t0, t1)
// (two per rule: dL*M_R then M_L*dR) and its full terms [f0, f1) (M_L*M_R,
// one per rule), both ranges into one term table.
struct HeadDev {
    uint32_t target;
    uint32_t out[2];
    uint32_t t0, t1;  // live delta terms: those whose delta operand can be non-empty
    uint32_t f0, f1;  // full terms, one per rule
    uint32_t p0, p1;  // every delta term, two per rule in rule order (scheduler features)
    uint32_t T;
    uint32_t colT;    // 1: the plan reads the head as a left operand, so the column-major
                      // form of the head and of its deltas is kept up to date
};

// One product term  left * right  contributing to a head; operands are
// indices into the matrix-descriptor array, chosen by step parity.  `rule`
// is the rule's index in its query (the scheduler's rule id).
struct TermDev {
    uint32_t left[2];
    uint32_t right[2];
    uint32_t rule;
};

// The operand arrays one term of one step parity reads, resolved on the host
// so that the sparse kernel reaches them in one dependent load instead of
// term -> matrix id -> descriptor -> array.  Table layout: [parity][term].
struct TermPtr {
    const uint32_t *occQ;   // left operand: per-quarter tile-row occupancy
    const uint32_t *occC;   // right operand: tile-column occupancy
    const uint64_t *lrmask; // left operand: row mask per (tile, quarter)
    const uint64_t *lcolq;  // left operand: column union per (tile, quarter)
    const uint64_t *lcolT;  // left operand: quarters in column-major form
    const uint64_t *lbits;  // left operand tiles
    const uint64_t *rrmask; // right operand: row mask per tile
    const uint64_t *rbits;  // right operand tiles
    const uint32_t *rcnt;   // right operand: bits per tile
};

// Statistics and the work cursor of the lane-parallel sparse kernel, one
// copy per shard of its work-groups.  Shard = blockIdx.x mod SHARDS: the
// dispatcher deals work-groups round-robin over the 8 XCDs, so a shard is the
// waves of one XCD, and the same-address atomics of thousands of waves are
// spread over 8 cache lines instead of queueing on one.  The host folds the
// shards into the step's totals.
constexpr int SHARDS = 8;
struct alignas(128) CounterShard {
    uint64_t tile_products, a_rows, b_rows, out_rows, cand_hits, new_bits, busy, busy_max;
    uint64_t a_cols;          // column words of A quarters read (column-major gathers)
    // work units and their clocks, and those of the units that added no bit
    // (what a scheduler that skipped exactly the unproductive units would save)
    uint64_t unit_clk, idle_clk, idle_units, idle_brows, units_done;
    uint64_t unit_max;        // longest single work unit (clocks)
    // the longest unit's shape, as (clocks << 24) | info so that one 64-bit
    // max keeps the info of the longest unit: a = sub-units of its quarter
    // (bits 0-4), candidates (5-13, capped), pairs (14-23, capped);
    // b = B-row gathers (bits 0-23, capped); c = clocks / 256 in the filter
    // (bits 12-23) and pair (0-11) stages; d = the same for the candidate scan
    // (12-23) and the epilogue (0-11)
    uint64_t unit_max_a, unit_max_b, unit_max_c, unit_max_d;
    uint32_t unit_next;       // claims taken from this shard's blocks of 64 units
    uint32_t prio_next;       // the same, over the first phase's list (scheduler ordering)
    uint32_t max_unit_cands;
    // entries of this shard's segment of the touched-tile list, on a cache
    // line of its own (apart from the cursor the same waves claim from)
    alignas(128) uint32_t n_touched;
};

// Per-step counters (a ring of them lives in device memory).
struct StepCounters {
    uint32_t n_units;      // sparse work units (head, tile, quarter) built
    uint32_t unit_next;    // sparse product cursor
    uint32_t n_dense;      // dense work items (head, tile) built
    uint32_t dense_next;   // dense product cursor
    uint32_t n_touched;    // output tiles that received new bits
    uint32_t n_mfma;       // MFMA work items (head, tile) built
    uint32_t mfma_next;    // MFMA product cursor
    uint32_t n_deferred;   // (head, tile) items the scheduler deferred
    uint32_t n_full;       // (head, tile) items evaluated over full operands
    uint32_t n_scored;     // (rule, tile) items scored by the scheduler
    uint32_t trace0;       // first trace row written by this step
    uint32_t dirty_new;    // tiles that became dirty (deferred) in this step
    uint32_t dirty_cleared; // dirty tiles evaluated in full (or found full) in this step
    uint32_t max_unit_cands; // most candidate k-tiles of one sparse unit
    uint32_t split_extra;    // heavy-list entries requested (budget counter)
    uint32_t n_split;        // heavy units split
    uint32_t n_heavy;        // heavy-list entries (sub-units of split units, heavy unsplit quarters), claimed first
    uint32_t n_prio;         // units the scheduler ranked first (evaluated by the step's first phase)
    uint32_t shards_done;    // lane-parallel sparse kernel: bit s = shard s's units all claimed
    uint32_t prio_shards_done;  // the same for the scheduler's first phase
    uint64_t new_bits;
    uint64_t tile_products;   // (term, C tile, k) products evaluated (sparse: per quarter)
    uint64_t tp_dense;        // of which on the LDS / Four Russians paths
    uint64_t tp_mfma;         // of which on the MFMA path
    uint64_t tp_unchanged;    // sweeps: tile-products skipped, both operand tiles unchanged since the
                              // entry's last evaluation (dense kernel)
    uint64_t a_rows;          // non-empty A rows visited
    uint64_t a_cols;          // A-quarter column words read (column-major gathers, 8 B each)
    uint64_t unit_clk, idle_clk, idle_units, idle_brows, units_done;  // (CounterShard)
    uint64_t b_rows;          // B rows gathered from HBM/L2
    uint64_t out_rows;        // C-tile rows that received new bits
    uint64_t b_rows_lds;      // B rows OR-ed from LDS
    uint64_t m4r_lookups;     // Four-Russians table lookups
    uint64_t mfma_instr;      // v_mfma_i32_32x32x32_i8 issued (per wave)
    uint64_t cand_hits;       // sparse candidates that passed the row / column filter
    uint64_t a_rows_dense;    // as a_rows / b_rows / out_rows, for the dense and MFMA kernels
    uint64_t b_rows_dense;    // (a_rows / b_rows / out_rows count the sparse kernels only)
    uint64_t out_rows_dense;
    uint64_t sparse_busy;     // sum over sparse-product waves of their busy clocks (s_memtime)
    uint64_t sparse_max;      // longest sparse-product wave (clocks)
    uint64_t sparse_unit_max; // longest single sparse work unit (clocks)
    uint64_t sparse_unit_max_a, sparse_unit_max_b, sparse_unit_max_c, sparse_unit_max_d;  // its shape
                                                                                     // (CounterShard)
    CounterShard shard[SHARDS];  // lane-parallel sparse kernel (folded into the fields above by the host)
};

// Host side: add the shards' partial sums into the step's totals.
inline void fold_shards(StepCounters &c) {
    for (int s = 0; s < SHARDS; ++s) {
        CounterShard &h = c.shard[s];
        c.tile_products += h.tile_products;
        c.a_rows += h.a_rows;
        c.a_cols += h.a_cols;
        c.unit_clk += h.unit_clk;
        c.idle_clk += h.idle_clk;
        c.idle_units += h.idle_units;
        c.idle_brows += h.idle_brows;
        c.units_done += h.units_done;
        c.b_rows += h.b_rows;
        c.out_rows += h.out_rows;
        c.cand_hits += h.cand_hits;
        c.new_bits += h.new_bits;
        c.sparse_busy += h.busy;
        if (h.busy_max > c.sparse_max) c.sparse_max = h.busy_max;
        if (h.unit_max > c.sparse_unit_max) c.sparse_unit_max = h.unit_max;
        if (h.unit_max_a > c.sparse_unit_max_a) c.sparse_unit_max_a = h.unit_max_a;
        if (h.unit_max_b > c.sparse_unit_max_b) c.sparse_unit_max_b = h.unit_max_b;
        if (h.unit_max_c > c.sparse_unit_max_c) c.sparse_unit_max_c = h.unit_max_c;
        if (h.unit_max_d > c.sparse_unit_max_d) c.sparse_unit_max_d = h.unit_max_d;
        if (h.max_unit_cands > c.max_unit_cands) c.max_unit_cands = h.max_unit_cands;
        c.n_touched += h.n_touched;
        h = CounterShard{};
    }
}

constexpr int RING = 64;  // step-counter ring length

// Touched-tile lists: segment 0 is appended to by the dense, MFMA and
// row-per-lane sparse kernels (count StepCounters::n_touched), segment 1 + s
// by shard s of the lane-parallel sparse kernel (count shard[s].n_touched).
// Each segment holds `seg_cap` entries, one per head tile of the plan.
struct TouchedLists {
    uint64_t *list;
    uint32_t seg_cap;
};
// Entry e of the concatenated segments (e below the sum of the counts).
// (the entries are written by kernels with plain stores: read through the
// vector path, never the scalar cache)
__device__ __forceinline__ uint64_t touched_entry(const StepCounters *c, TouchedLists tl, uint32_t e) {
    uint32_t n = c->n_touched;
    if (e < n) return __hip_atomic_load(&tl.list[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    e -= n;
    for (int s = 0; s < SHARDS; ++s) {
        n = c->shard[s].n_touched;
        if (e < n)
            return __hip_atomic_load(&tl.list[(size_t)(1 + s) * tl.seg_cap + e], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
        e -= n;
    }
    return 0;  // not reached for e below the total
}
__device__ __forceinline__ uint32_t touched_total(const StepCounters *c) {
    uint32_t n = c->n_touched;
    for (int s = 0; s < SHARDS; ++s) n += c->shard[s].n_touched;
    return n;
}

static inline uint32_t tiles_per_side(uint32_t n) { return (n + TILE - 1) / TILE; }
static inline uint32_t occ_words(uint32_t T) { return (T + 31) / 32; }

// Item encodings: dense item = (head << 40) | tile; sparse unit =
// (head << 40) | (tile << 2) | quarter.  A heavy sparse unit is split over
// nsub waves (sub-units sub = 0 .. nsub-1, sub-unit sub taking the quarter's
// rows r = sub (mod nsub) of every candidate): (sub << 55) | ((nsub - 1) << 59).
// Bit 63 set = evaluate the head's full terms (M_L * M_R) instead of its
// semi-naive terms.
constexpr uint64_t ITEM_TILE_MASK = 0xffffffffffull;
// A tile unit (lane-parallel sparse kernel, sparse_mode 2) covers every quarter
// of its tile with candidates: bit ITEM_QSET set, the quarter set in bits
// 34..37, the tile in bits 2..33 (tiles below 2^32).
constexpr uint64_t ITEM_QSET = 1ull << 38;
constexpr int ITEM_QSET_SHIFT = 34;
constexpr uint64_t ITEM_FULL = 1ull << 63;
// A sub-unit of a quarter split by rows: sub-unit s of n takes every
// candidate of the quarter but only its rows [64 s / n, 64 (s + 1) / n), and
// finishes them itself (no merge pass); without the flag a split quarter's
// sub-units take the candidates g = s (mod n) and leave partial rows.
constexpr uint64_t ITEM_ROWSPLIT = 1ull << 39;
constexpr uint32_t MAX_HEADS = 1u << 15;   // heads per plan (15-bit field)
constexpr uint32_t SPLIT_MAX = 16;         // sub-units per heavy sparse unit
// Unsplit quarters of more than this many 16ths of the split threshold go to
// the heavy list (claimed first).  On the 65,536-vertex forest the longest
// unit of the odd steps is such a quarter (117-119 candidates against a
// threshold of 128); claimed first it no longer ends its step late.  12/16
// measured best (13/16 the same within noise; 4, 8, 10, 14 and 15 slower:
// more units moved to the front crowd the first claims, fewer miss the hub).
constexpr uint32_t HEAVY_UNSPLIT_16THS = 12;
constexpr uint32_t SPLIT_CANDS = 64;       // default of cfpq_opts.split_cands: a unit with more
                                           // candidates than this is split, one sub-unit per
                                           // split_cands candidates
__host__ __device__ inline uint32_t item_head(uint64_t it) { return (uint32_t)((it >> 40) & 0x7fffu); }
__host__ __device__ inline uint32_t unit_sub(uint64_t it) { return (uint32_t)((it >> 55) & 15u); }
__host__ __device__ inline uint32_t unit_tile(uint64_t it) { return (uint32_t)((it >> 2) & 0xffffffffu); }
__host__ __device__ inline uint32_t unit_quarters(uint64_t it) {
    return (it & ITEM_QSET) ? (uint32_t)((it >> ITEM_QSET_SHIFT) & 15u) : 1u << (it & 3u);
}
__host__ __device__ inline uint32_t unit_nsub(uint64_t it) { return (uint32_t)((it >> 59) & 15u) + 1u; }

}  // namespace cfpq
