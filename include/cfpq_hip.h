/*
 * cfpq_hip.h -- C ABI of the MI355X (gfx950) CFPQ engine.
 *
 * Plain C: opaque handles, fixed-width integers, status codes.  No C++
 * types, no exceptions and no exit() cross this boundary (the reference's
 * GPU backends call exit() on any CUDA error, cuda/naive-gpu/Multiplication.cu:10-17).
 *
 * What each entry point replaces in the reference (JetBrains-Research/CFPQ-on-GPGPU):
 *   cfpq_mat_alloc / cfpq_mat_free   gpuMatrix::allocate_device_matrix / deallocate
 *                                    (cuda/naive-gpu/GpuMatrix.cpp:31-37,
 *                                    Multiplication.cu:103-112)
 *   cfpq_mat_upload                  gpuMatrix::transfer_to_gpu + the host mirror set_bit
 *                                    (GpuMatrix.cpp:15-18,51-53; called from
 *                                    gpuMatricesEnv::environment_preprocessing :56-64)
 *   cfpq_mat_set_bits                Grammar::intersection_with_graph's set_bit loops
 *                                    (cfpq-cpp/Grammar.h:41-51), batched
 *   cfpq_mat_download                gpuMatrix::transfer_from_gpu (GpuMatrix.cpp:55-57;
 *                                    environment_postprocessing :66-77)
 *   cfpq_add_mul                     Matrix::add_mul (cfpq-cpp/Matrix.h:17), as
 *                                    implemented by gpu_lib::matrix_product_add_wrapper
 *                                    (cuda/naive-gpu/Multiplication.cu:145-166) and
 *                                    the oracle M4riMatrix::add_mul (m4ri/m4ri_matrix.cpp:20-28)
 *   cfpq_fixpoint                    the worklist loop of Grammar::intersection_with_graph
 *                                    (cfpq-cpp/Grammar.h:58-69), moved onto the device
 *   cfpq_batch_fixpoint              no counterpart: the reference runs one query per
 *                                    process (m4ri/main.cpp:7-15, test_utils/runners.py:24)
 *   cfpq_mat_fingerprint             Grammar::print_results (cfpq-cpp/Grammar.cpp:81-101)
 *   cfpq_pairs_fixpoint              the sparse backends' add_mul, GraphBLAS GrB_mxm
 *                                    (graphBlas/graphblas_matrix.cpp:19-31) and scipy CSR
 *                                    (python/matmul.py:18-28), under a device fixpoint
 *   cfpq_mat_copy                    the D2D copy MethodOf4RusMatrix::add_mul makes when
 *                                    the head aliases an operand
 *                                    (cuda/methodOf4RusBooleanSemiringGpu/methodOf4RusBooleanSemiringMatrix.cu:62-67)
 *
 * Conventions
 *   - Every function returns 0 (CFPQ_OK) on success or a negative CFPQ_E* code;
 *     cfpq_last_error(ctx) then describes the failure.
 *   - The library owns device memory; the caller owns every host buffer.
 *   - One context per host thread; calls on a context are serialised on the
 *     context's HIP stream and return when their results are complete.
 *   - Host bit matrices are row-major: n rows of `words_per_row` uint64
 *     words, column j in word j/64 at bit j%64 (the m4ri convention used by
 *     the reference's oracle backend).
 */
#ifndef CFPQ_HIP_H
#define CFPQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFPQ_OK 0
#define CFPQ_E_INVALID (-1)   /* bad argument: null handle, size mismatch, index out of range */
#define CFPQ_E_DEVICE (-2)    /* a HIP runtime call failed */
#define CFPQ_E_NOMEM (-3)     /* device or host allocation failed */
#define CFPQ_E_LIMIT (-4)     /* iteration cap reached before the fixpoint */
#define CFPQ_E_CAPACITY (-5)  /* an output array is too small: the sizes needed are written back */

typedef struct cfpq_ctx cfpq_ctx;
typedef struct cfpq_mat cfpq_mat;

/* Product kernel choice for a tile-product (C tile |= A tile * B tile). */
#define CFPQ_KERNEL_AUTO 0    /* per tile, by the popcount of the A tile */
#define CFPQ_KERNEL_GATHER 1  /* bit-packed OR-of-rows, B rows gathered from HBM */
#define CFPQ_KERNEL_LDS 2     /* bit-packed OR-of-rows, B tile staged in LDS */
#define CFPQ_KERNEL_M4R 3     /* Four Russians, 4-bit tables in LDS */
#define CFPQ_KERNEL_MFMA 4    /* exact int8 0/1 matrix-core product (v_mfma_i32_32x32x32_i8),
                                 i32 accumulation, > 0 */

typedef struct cfpq_opts {
    int32_t kernel;             /* CFPQ_KERNEL_*                                         */
    uint32_t dense_bits;        /* AUTO: A-tile popcount at or above which the tile is dense
                                   (0 = library default)                                 */
    uint32_t mfma_bits;         /* AUTO: A-tile popcount at or above which MFMA is used
                                   (0 = library default, 0xffffffff = never)             */
    uint32_t m4r_bits;          /* AUTO: A-tile popcount at or above which the Four
                                   Russians tables are used (0 = library default)        */
    uint32_t max_iterations;    /* 0 = unlimited                                         */
    int32_t final_full_pass;    /* 1: after the delta fixpoint, verify with one exact
                                   full pass (every rule over full matrices)             */
    int32_t iters_per_sync;     /* fixpoint steps per window: the host reads a window's step
                                   counters while the next window runs (0 = default) */
    int32_t use_graph;          /* 1: replay one captured hipGraph per fixpoint step     */
    int32_t sparse_batch;       /* k-tiles per batch in the sparse product: 1, 2 or 4 (0 = default) */
    int32_t sparse_waves_per_cu; /* persistent sparse-product waves per CU (0 = default)   */
    int32_t sparse_chunk;       /* work units a sparse-product wave claims per atomic (0 = default) */
    int32_t sched;              /* 1: the learned tile scheduler may defer (head, tile) work
                                   (needs weights: cfpq_sched_set_weights / _train)     */
    uint32_t sched_threshold;   /* defer an item whose predicted activity probability is
                                   below sched_threshold / 65536 (0 = library default)   */
    int32_t sparse_mode;        /* sparse product and its work-tile shape: 0 = lane-parallel
                                   candidates, one work unit per 64-row quarter (default);
                                   3 = the same with one unit per 128-row half tile; 2 = one
                                   unit per 256-row tile (each covering its quarters with
                                   candidates); 1 = row-per-lane with sparse_batch
                                   candidates per round                                  */
    int32_t fuse_apply;         /* 1: fixpoint steps merge new rows into the head matrices
                                   inside the product kernels (no separate apply pass)  */
    int32_t schedule;           /* CFPQ_SCHEDULE_*: how cfpq_fixpoint orders the rules    */
    int32_t m4r_keys;           /* Four Russians key width: 4 = 16-entry tables over whole
                                   tile rows, shared by the work-group (default); 8 =
                                   256-entry tables per wavefront over a 64-column slice */
    int32_t sched_order;        /* 1: the learned tile scheduler ranks the work of each delta
                                   step: the (head, tile) items it predicts active are
                                   evaluated in a first phase, the rest in a second one, which
                                   with fuse_apply reads the heads the first phase updated
                                   (needs weights; independent of deferral)              */
    uint32_t order_threshold;   /* sched_order: an item goes to the first phase when its
                                   predicted activity probability is at least
                                   order_threshold / 65536 (0 = library default, 1/2)    */
    uint32_t split_cands;       /* lane-parallel sparse product: a 64-row quarter with more
                                   candidate k-tiles than this is split into sub-units of
                                   this many candidates each, merged afterwards
                                   (0 = library default, 64; 0xffffffff = never)         */
    int32_t split_rows;         /* how a split quarter is shared out: 0 = by candidates,
                                   the sub-units' partial rows OR-ed by a merge pass
                                   (default); 1 = by rows, each sub-unit evaluating every
                                   candidate for its own rows and writing them itself    */
    uint32_t split_piece;       /* candidates per sub-unit of a split quarter, at most 16
                                   sub-units (0 = split_cands)                           */
    int32_t no_kernel_timing;   /* 1: no timing events around the product kernels of each
                                   step (cfpq_stats sparse_ms, merge_ms, dense_ms, mfma_ms
                                   and product_ms stay 0; fixpoint_ms is still measured):
                                   each event costs a few microseconds of gap, which
                                   matters for small graphs of many short steps          */
} cfpq_opts;

/* cfpq_opts.schedule */
#define CFPQ_SCHEDULE_DELTA 0   /* semi-naive steps, every rule at once over the previous
                                   step's delta (default)                                */
#define CFPQ_SCHEDULE_SWEEP 1   /* head sweeps: one head at a time, its rules over the full
                                   matrices, each head seeing the bits of the heads before
                                   it; a sweep revisits only heads one of whose operands
                                   changed in the previous sweep, and the fixpoint is the
                                   first sweep that adds no bit.  The reference
                                   worklist's order of evaluation (cfpq-cpp/Grammar.h:58-69)
                                   without a host round trip per rule; it wins on
                                   saturating grammars whose tiles fill early         */
#define CFPQ_SCHEDULE_RULES 2   /* as CFPQ_SCHEDULE_SWEEP, one rule at a time: a sweep
                                   revisits only the rules one of whose operands
                                   changed                                               */

typedef struct cfpq_rule {
    uint32_t head, left, right; /* indices into the matrix array: head -> left right */
} cfpq_rule;

typedef struct cfpq_stats {
    uint32_t iterations;        /* delta steps run, including the final empty one       */
    uint32_t full_passes;       /* exact full passes run                                 */
    uint64_t new_bits;          /* bits added to all matrices                            */
    uint64_t items;             /* (head, C-tile) work items evaluated                   */
    uint64_t tile_products;     /* (rule, C tile, k) tile-products evaluated, all kernels */
    uint64_t tile_products_dense; /* of which on the LDS / M4R paths                     */
    uint64_t tile_products_mfma;  /* of which on the MFMA path                           */
    uint64_t a_rows;            /* non-empty A-tile rows visited (sparse kernels)        */
    uint64_t b_rows;            /* B rows gathered from HBM/L2 (sparse kernels)          */
    uint64_t b_rows_lds;        /* B rows OR-ed from LDS (dense kernel)                  */
    uint64_t m4r_lookups;       /* Four Russians table lookups (32 B each)               */
    uint64_t out_rows;          /* C-tile rows that received new bits (sparse kernels)   */
    double fixpoint_ms;         /* device time of the whole call (hipEvent)              */
    double product_ms;          /* device time inside the product kernel (hipEvent)      */
    uint32_t product_launches;  /* fixpoint steps whose product kernels were timed       */
    uint32_t host_syncs;        /* host waits on the device flag                         */
    uint64_t mfma_instr;        /* v_mfma_i32_32x32x32_i8 issued (per wave; 2*32^3 ops each) */
    double sparse_ms;           /* device time in the sparse product (k_product_sparse*),
                                   the merge of split quarters excluded (merge_ms)       */
    double dense_ms;            /* device time in the LDS / Four Russians product        */
    double mfma_ms;             /* device time in the MFMA product                       */
    uint64_t sched_scored;      /* (rule, tile) items scored by the learned scheduler    */
    uint64_t sched_deferred;    /* (head, tile) items it deferred                        */
    uint64_t sched_full;        /* deferred tiles later evaluated over full operands     */
    uint32_t flush_steps;       /* steps that evaluated every deferred tile (no deferral) */
    uint32_t max_unit_cands;    /* most candidate k-tiles one sparse work unit visited   */
    uint64_t cand_hits;         /* sparse candidates that passed the row / column filter */
    uint64_t a_rows_dense;      /* A rows visited by the dense kernel                    */
    uint64_t b_rows_dense;      /* B rows the dense kernel gathered from HBM/L2          */
    uint64_t out_rows_dense;    /* C-tile rows written by the dense and MFMA kernels     */
    uint64_t a_cols;            /* A-quarter column words read by the sparse kernel's
                                   column-major gathers (8 B each)                       */
    uint64_t units_done;        /* work units the lane-parallel sparse kernel evaluated  */
    uint64_t units_idle;        /* of which added no bit (split sub-units not counted)   */
    uint64_t unit_clocks;       /* summed clocks of those units (s_memtime)              */
    uint64_t unit_clocks_idle;  /* summed clocks of the units that added no bit: what a
                                   scheduler skipping exactly those would save         */
    uint64_t b_rows_idle;       /* B rows gathered by the units that added no bit        */
    double merge_ms;            /* device time merging split quarters' partial rows
                                   (k_merge_split; part of product_ms)                   */
    uint64_t tile_products_unchanged; /* head and rule sweeps: dense tile-products skipped
                                   because neither operand tile changed since the rule
                                   was last evaluated (not in tile_products)           */
} cfpq_stats;

/* ---- lifecycle ---------------------------------------------------------- */
void cfpq_default_opts(cfpq_opts *opts);
int cfpq_create(int device, const cfpq_opts *opts, cfpq_ctx **out);
int cfpq_destroy(cfpq_ctx *ctx);
const char *cfpq_last_error(const cfpq_ctx *ctx);
int cfpq_set_opts(cfpq_ctx *ctx, const cfpq_opts *opts);
int cfpq_sync(cfpq_ctx *ctx);

/* ---- matrices ------------------------------------------------------------ */
int cfpq_mat_alloc(cfpq_ctx *ctx, uint32_t n, cfpq_mat **out);   /* all-zero n x n */
int cfpq_mat_free(cfpq_ctx *ctx, cfpq_mat *mat);
uint32_t cfpq_mat_n(const cfpq_mat *mat);
int cfpq_mat_upload(cfpq_ctx *ctx, cfpq_mat *mat, const uint64_t *rows, size_t words_per_row);
int cfpq_mat_download(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *rows, size_t words_per_row);
/* OR-in `count` bits (rows[i], cols[i]); indices must be < n. */
int cfpq_mat_set_bits(cfpq_ctx *ctx, cfpq_mat *mat, const uint32_t *rows, const uint32_t *cols, size_t count);
int cfpq_mat_popcount(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *count);
/* dst := src (same n), on the device. */
int cfpq_mat_copy(cfpq_ctx *ctx, cfpq_mat *dst, const cfpq_mat *src);
/* (hash, count) of Grammar::print_results for this matrix. */
int cfpq_mat_fingerprint(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *hash, uint64_t *count);

/* ---- products and fixpoints --------------------------------------------- */
/* c |= a * b over (OR, AND).  a and/or b may alias c: the product is taken over
 * the operands as they were on entry.  *changed = 1 if c gained a bit. */
int cfpq_add_mul(cfpq_ctx *ctx, cfpq_mat *c, const cfpq_mat *a, const cfpq_mat *b, int *changed);

/* cfpq_add_mul without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops with a
 * fixed memory budget, harnesses that forbid hipMalloc during a run).  Same
 * result, aliasing rules and changed flag as cfpq_add_mul, for matrices of
 * n <= CFPQ_NOALLOC_MAX_N (larger n: CFPQ_E_INVALID).  Everything the step
 * needs beyond a, b and c (the delta matrix, work lists, descriptor tables,
 * counters) lives in a file-scope device array of the library, about
 * 2.4 MiB, which each GPU it runs on holds until the process ends; no call
 * to hipMalloc/hipFree is made.  Calls on one GPU share that array: the
 * library serialises them within a process (a per-device lock), so calls
 * from several threads on one GPU run one after another. */
#define CFPQ_NOALLOC_MAX_N 4096u
int cfpq_add_mul_noalloc(cfpq_ctx *ctx, cfpq_mat *c, const cfpq_mat *a, const cfpq_mat *b, int *changed);

/* Least fixpoint of  M[head] |= M[left] * M[right]  for all rules, in place on
 * mats[0..nmats) (all of one size n).  stats may be NULL. */
int cfpq_fixpoint(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, cfpq_mat *const *mats,
                  size_t nmats, cfpq_stats *stats);

/* One independent query of a batch: its own matrices (one size n per query)
 * and rules indexing them. */
typedef struct cfpq_query {
    cfpq_mat *const *mats;
    size_t nmats;
    const cfpq_rule *rules;
    size_t nrules;
} cfpq_query;

/* Least fixpoints of all queries, advanced together: one set of kernel
 * launches per step covers every query still changing. */
int cfpq_batch_fixpoint(cfpq_ctx *ctx, const cfpq_query *queries, size_t nqueries, cfpq_stats *stats);

/* ---- single-matrix formulation --------------------------------------------
 * Replaces the CUTLASS backend's single matrix of nonterminal sets
 * (cuda/cutlass/CutlassMatrix.cu:291-316 fixpoint loop, :353-437 entry
 * MultMatrSquare; grammar packed by cfpq-cpp/Grammar.cpp:63-75, matrix filled
 * by cfpq-cpp/Graph.cpp:22-48).  elems: n * n uint32 in row-major order, bit x
 * of element (i, j) set iff nonterminal x derives a path i -> j (terminal
 * edges, and the diagonal of epsilon nonterminals, set by the caller); closed
 * under `rules` on the device and written back in place.  Nonterminals < 32,
 * at most 256 rules.  stats: iterations (products run, the last adds nothing),
 * items (elements that gained nonterminals), fixpoint_ms.  Device memory:
 * 3 * n * n * 4 bytes for the call. */
int cfpq_nt_closure(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, uint32_t *elems, uint32_t n,
                    cfpq_stats *stats);

/* ---- sparse pair-list formulation ------------------------------------------
 * Replaces the reference's sparse matrix backends: GraphBLAS GrB_mxm over the
 * LOR semiring (graphBlas/graphblas_matrix.cpp:19-31, behind the same
 * Matrix::add_mul of cfpq-cpp/Matrix.h:17) and scipy CSR products
 * (python/matmul.py:18-28, python/matrix_utils.py:44).  For graphs whose
 * vertex count is beyond the bit tiles' reach (n up to 2^31) or whose
 * relations are very sparse: device memory grows with the pairs, not n^2.
 * One set per nonterminal: keys[0 .. count) hold pairs row << 32 | col, in
 * any order, duplicates allowed (terminal edges and the epsilon diagonal, set
 * by the caller); on success they hold the closed relation, sorted ascending
 * without duplicates, and count its size.  If a closed relation does not fit
 * its set's cap, CFPQ_E_CAPACITY is returned, no keys are written and every
 * count holds the size needed.  The device arrays of a call stay with the
 * context for the next call (no allocation once their sizes are reached);
 * cfpq_destroy frees them, and a call that runs out of device memory frees
 * them before it retries. */
typedef struct cfpq_pairs {
    uint64_t *keys;
    uint64_t count;
    uint64_t cap;
} cfpq_pairs;
/* stats: iterations (steps, the last adds nothing), new_bits (pairs added),
 * items (join outputs before deduplication), fixpoint_ms, product_launches
 * (join expansions), host_syncs. */
int cfpq_pairs_fixpoint(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, cfpq_pairs *sets, uint32_t nsets,
                        uint32_t n, cfpq_stats *stats);

/* ---- learned tile scheduler ----------------------------------------------
 * No counterpart in the reference.  A tile-activity network (16 inputs:
 * 8 tile features + 8-dim rule embedding; 64-64 ReLU hidden; 1 logit),
 * bf16 parameters, fp32 arithmetic.  CFPQ_SCHED_PARAMS parameters:
 * E[64][8], W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3. */
#define CFPQ_SCHED_PARAMS 5825
#define CFPQ_SCHED_FEATURES 8
int cfpq_sched_set_weights(cfpq_ctx *ctx, const uint16_t *bf16, size_t count);
int cfpq_sched_get_weights(cfpq_ctx *ctx, uint16_t *bf16, size_t count);
/* Deterministic initialisation (scaled uniform, from `seed`). */
int cfpq_sched_init(cfpq_ctx *ctx, uint64_t seed);
/* Forward (logits, may be NULL), mean binary cross-entropy (loss, may be NULL)
 * and its fp32 gradient (grad[CFPQ_SCHED_PARAMS], may be NULL) of the network
 * with parameters w (bf16) on n samples: x[n][8], rule[n], y[n] in {0, 1}. */
int cfpq_sched_eval(cfpq_ctx *ctx, const uint16_t *w, const float *x, const uint32_t *rule, const float *y,
                    size_t n, float *logits, float *loss, float *grad);
/* Adam on the context's network over the samples, on the device; loss[epoch]
 * receives each epoch's mean training loss (may be NULL). */
int cfpq_sched_train(cfpq_ctx *ctx, const float *x, const uint32_t *rule, const float *y, size_t n,
                     uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss);
/* Trace collection: while enabled, every fixpoint step records one row per
 * scored (rule, tile) item -- features, rule id and whether the item's head
 * tile gained a bit in that step -- up to `capacity` rows. */
int cfpq_sched_trace_begin(cfpq_ctx *ctx, size_t capacity);
int cfpq_sched_trace_read(cfpq_ctx *ctx, float *x, uint32_t *rule, float *y, size_t capacity, size_t *count);
int cfpq_sched_trace_end(cfpq_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* CFPQ_HIP_H */
