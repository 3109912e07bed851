// Host side of the C ABI (include/cfpq_hip.h): contexts, device matrices,
// the add_mul call and the device-resident delta fixpoint.
//
// The reference drives its fixpoint from the host, one add_mul per rule with
// a flag round-trip and a device-wide sync each time (cfpq-cpp/Grammar.h:58-69,
// cuda/naive-gpu/Multiplication.cu:145-166).  Here a fixpoint *step* evaluates
// every rule at once on the device:
//   new_H = (dL * M_R | M_L * dR) & ~M_H  for each rule H -> L R,
// with d the bits added by the previous step, and the host only reads a small
// counter block every `iters_per_sync` steps.  Steps after convergence are
// empty (no work items), so enqueueing a few ahead is harmless.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "cfpq_hip.h"
#include "ntmat.hpp"
#include "pairs.hpp"
#include "sched.hpp"

using namespace cfpq;

struct cfpq_mat {
    uint32_t n = 0;
    uint32_t T = 0;
    // the column-major form (MatDev::colT) is out of date: the matrix was the
    // head of a fixpoint that did not read it as a left operand; rebuilt
    // before any call reads it as one (colT_fresh)
    bool colT_stale = false;
    void *base = nullptr;  // single device allocation
    size_t bytes = 0;
    MatDev d = {};
};

struct cfpq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    cfpq_opts opts = {};
    std::string err;
    int num_cu = 256;

    // staging for uploads / downloads
    void *stage = nullptr;
    size_t stage_bytes = 0;

    // step counters (ring) and their pinned host mirror
    StepCounters *ring = nullptr;
    StepCounters *ring_host = nullptr;
    unsigned long long *scratch64 = nullptr;  // small reductions
    unsigned long long *scratch64_host = nullptr;
    // device arrays of the pair-list fixpoint kept between its calls
    // (pointer, capacity in uint64 elements); freed by cfpq_destroy
    std::vector<std::pair<uint64_t *, size_t>> pair_pool;

    // work lists
    uint64_t *items = nullptr;   // dense work items
    size_t items_cap = 0;
    uint64_t *units = nullptr;   // sparse work units
    size_t units_cap = 0;
    uint64_t *mitems = nullptr;  // MFMA work items
    size_t mitems_cap = 0;
    uint64_t *lists[2] = {nullptr, nullptr};
    size_t lists_cap = 0;
    // split quarters of the sparse product: partial rows (64 rows per slot),
    // their masks, and the list of split quarters (2 words each); at most
    // split_slots sub-units per step (2 KiB of partial rows each: 256 MiB)
    uint64_t *partials = nullptr;
    size_t partials_cap = 0;
    uint64_t *pmask = nullptr;
    size_t pmask_cap = 0;
    uint64_t *splits = nullptr;
    size_t splits_cap = 0;
    uint32_t split_slots = 131072;

    // descriptor tables
    MatDev *d_mats = nullptr;
    size_t d_mats_cap = 0;
    HeadDev *d_heads = nullptr;
    size_t d_heads_cap = 0;
    TermDev *d_terms = nullptr;  // [delta terms | full terms]
    size_t d_terms_cap = 0;
    TermPtr *d_tptr = nullptr;   // [parity][delta terms | full terms | live delta terms]
    size_t d_tptr_cap = 0;
    uint32_t *d_ids = nullptr;   // id lists for clear/copy kernels
    size_t d_ids_cap = 0;

    // pool of all-zero matrices keyed by n
    std::multimap<uint32_t, cfpq_mat *> pool;

    // events
    std::vector<hipEvent_t> ev;

    // captured fixpoint steps: graph_exec[variant * RING + slot], valid while
    // the launch configuration equals graph_sig
    std::vector<hipGraphExec_t> graph_exec;
    std::vector<uint64_t> graph_sig;

    // learned scheduler
    uint16_t *sched_w = nullptr;      // bf16 parameters on the device
    bool sched_have = false;
    uint32_t *dirty = nullptr;        // deferred-tile bitmaps, one per head
    size_t dirty_cap = 0;
    uint32_t *dirty_off = nullptr;    // word offset of each head's bitmap
    size_t dirty_off_cap = 0;
    // head and rule sweeps: per plan entry, the step (from 1) that last
    // evaluated it, 0 before its first evaluation in the call
    uint32_t *last_eval = nullptr;
    size_t last_eval_cap = 0;
    // trace rows
    float *tr_x = nullptr, *tr_y = nullptr;
    uint32_t *tr_rule = nullptr, *tr_item = nullptr, *tr_count = nullptr;
    size_t tr_cap = 0;
    // cfpq_add_mul_noalloc: this device's copy of the scratch array (its
    // device address), and the flag that turns every allocation into an error
    char *noalloc_base = nullptr;
    bool no_alloc = false;

    // training state
    float *t_master = nullptr, *t_m1 = nullptr, *t_m2 = nullptr, *t_grad = nullptr, *t_loss = nullptr;
    float *t_x = nullptr, *t_y = nullptr;
    uint32_t *t_rule = nullptr;
    size_t t_cap = 0, t_x_cap = 0, t_y_cap = 0, t_rule_cap = 0;
};

static constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Bytes of a matrix with T tiles per side (the layout of mat_layout).
static constexpr size_t mat_bytes(size_t T) {
    return 2 * align256(T * T * TILE_WORDS * 8) + 2 * align256(T * T * 4) + align256(T * T * 4 * 8) +
           align256(T * T * QUARTERS * TROW_WORDS * 8) + align256(OCC_ARRAYS * T * ((T + 31) / 32) * 4);
}

// ---------------------------------------------------------------------------
// scratch of cfpq_add_mul_noalloc
// ---------------------------------------------------------------------------
//
// Everything beyond its inputs and output that one add_mul step needs -- the
// delta matrix of the new bits, the descriptor tables, the work lists, the
// touched-tile lists, the counter ring and the scheduler's dirty bitmap --
// sized for n <= CFPQ_NOALLOC_MAX_N, lives in one file-scope device array.
// Each GPU holds its own copy for the life of the process; in host code the
// name `scratch` is not a device address, so the array is only ever reached
// through the address hipGetSymbolAddress returns for the current device
// (kept per device by noalloc_init).
constexpr uint32_t NOALLOC_T = (CFPQ_NOALLOC_MAX_N + TILE - 1) / TILE;
constexpr size_t NOALLOC_TILES = (size_t)NOALLOC_T * NOALLOC_T;
constexpr size_t NA_MATS_CAP = 4, NA_HEADS_CAP = 1, NA_TERMS_CAP = 3, NA_TPTR_CAP = 6, NA_IDS_CAP = 8;
constexpr size_t NA_UNITS_CAP = 3 * QUARTERS * NOALLOC_TILES, NA_LIST_CAP = (1 + SHARDS) * NOALLOC_TILES;
constexpr size_t NA_DIRTY_CAP = (NOALLOC_TILES + 31) / 32;
constexpr int EV_SLOT = 5;  // timing events per ring slot (sparse, merge, dense, MFMA bounds)
constexpr size_t NA_RING = align256(mat_bytes(NOALLOC_T));
constexpr size_t NA_MATS = NA_RING + align256(sizeof(StepCounters) * RING);
constexpr size_t NA_HEADS = NA_MATS + align256(sizeof(MatDev) * NA_MATS_CAP);
constexpr size_t NA_TERMS = NA_HEADS + align256(sizeof(HeadDev) * NA_HEADS_CAP);
constexpr size_t NA_TPTR = NA_TERMS + align256(sizeof(TermDev) * NA_TERMS_CAP);
constexpr size_t NA_IDS = NA_TPTR + align256(sizeof(TermPtr) * NA_TPTR_CAP);
constexpr size_t NA_ITEMS = NA_IDS + align256(4 * NA_IDS_CAP);
constexpr size_t NA_MITEMS = NA_ITEMS + align256(8 * NOALLOC_TILES);
constexpr size_t NA_UNITS = NA_MITEMS + align256(8 * NOALLOC_TILES);
constexpr size_t NA_LIST0 = NA_UNITS + align256(8 * NA_UNITS_CAP);
constexpr size_t NA_LIST1 = NA_LIST0 + align256(8 * NA_LIST_CAP);
constexpr size_t NA_DIRTY = NA_LIST1 + align256(8 * NA_LIST_CAP);
constexpr size_t NA_DIRTY_OFF = NA_DIRTY + align256(4 * NA_DIRTY_CAP);
constexpr uint32_t NA_SPLIT_SLOTS = 32;  // split sub-units per call
constexpr size_t NA_PARTIALS = NA_DIRTY_OFF + 256;
constexpr size_t NA_PMASK = NA_PARTIALS + align256(8 * NA_SPLIT_SLOTS * WAVE * TROW_WORDS);
constexpr size_t NA_SPLITS = NA_PMASK + align256(8 * NA_SPLIT_SLOTS);
constexpr size_t NOALLOC_BYTES = NA_SPLITS + align256(8 * 2 * (NA_SPLIT_SLOTS / 2 + 1));
static_assert(NOALLOC_BYTES <= (5u << 20), "the no-allocation scratch stays below 5 MiB");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_DEVICES = 64;
static std::mutex g_noalloc_init_mu;
static char *g_noalloc_ptr[NOALLOC_DEVICES];
static std::string g_noalloc_err[NOALLOC_DEVICES];
// calls on one GPU share its copy of the array: they are serialised per device
static std::mutex g_noalloc_call_mu[NOALLOC_DEVICES];

// Resolve this device's copy of the array once (the device must be current).
// A failure does not fail context creation: cfpq_add_mul_noalloc then
// reports it.
static void noalloc_init(cfpq_ctx *ctx, int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_noalloc_init_mu);
    if (!g_noalloc_ptr[device] && g_noalloc_err[device].empty()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p)
            g_noalloc_ptr[device] = (char *)p;
        else
            g_noalloc_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    }
    ctx->noalloc_base = g_noalloc_ptr[device];
}

// ---------------------------------------------------------------------------
// error helpers
// ---------------------------------------------------------------------------

static int fail(cfpq_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(ctx, e_ == hipErrorOutOfMemory ? CFPQ_E_NOMEM : CFPQ_E_DEVICE, "%s: %s (%s:%d)", \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

#define CHK(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != CFPQ_OK) return rc_; \
    } while (0)

template <class T>
static int grow(cfpq_ctx *ctx, T **ptr, size_t *cap, size_t need) {
    if (need <= *cap && *ptr) return CFPQ_OK;
    if (ctx->no_alloc) return fail(ctx, CFPQ_E_NOMEM, "a call that allocates nothing needed a larger device buffer");
    if (*ptr) HIPCHK(hipFree(*ptr));
    *ptr = nullptr;
    size_t n = std::max<size_t>(need, 16);
    n = std::max(n, *cap * 3 / 2);
    HIPCHK(hipMalloc((void **)ptr, n * sizeof(T)));
    *cap = n;
    return CFPQ_OK;
}

static int launch_check(cfpq_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, CFPQ_E_DEVICE, "launch %s: %s", what, hipGetErrorString(e));
    return CFPQ_OK;
}

// ---------------------------------------------------------------------------
// lifecycle
// ---------------------------------------------------------------------------

extern "C" void cfpq_default_opts(cfpq_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->kernel = CFPQ_KERNEL_AUTO;
    o->dense_bits = 0;
    o->mfma_bits = 0;
    o->max_iterations = 0;
    o->final_full_pass = 0;
    o->iters_per_sync = 0;
    o->use_graph = 0;
    o->schedule = CFPQ_SCHEDULE_DELTA;
}

static void normalise(cfpq_opts *o) {
    if (o->dense_bits == 0) o->dense_bits = 2048;   // >= 8 bits per tile row on average
    if (o->m4r_bits == 0) o->m4r_bits = 16384;      // >= 25% dense
    if (o->mfma_bits == 0) o->mfma_bits = 32768;    // >= 50% dense
    if (o->iters_per_sync <= 0) o->iters_per_sync = 4;
    if (o->sparse_batch != 1 && o->sparse_batch != 2 && o->sparse_batch != 4) o->sparse_batch = 2;
    if (o->sparse_waves_per_cu <= 0) o->sparse_waves_per_cu = 32;
    if (o->sparse_chunk <= 0) o->sparse_chunk = 1;
    // a divisor of 64 (claims stay inside one shard's block of 64 units)
    while (64 % o->sparse_chunk) --o->sparse_chunk;
    if (o->iters_per_sync > RING / 2) o->iters_per_sync = RING / 2;
    if (o->sched_threshold == 0) o->sched_threshold = 6554;  // P(active) < 0.1
    if (o->schedule != CFPQ_SCHEDULE_SWEEP && o->schedule != CFPQ_SCHEDULE_RULES) o->schedule = CFPQ_SCHEDULE_DELTA;
    if (o->m4r_keys != 8) o->m4r_keys = 4;
    if (o->order_threshold == 0) o->order_threshold = 32768;  // P(active) >= 1/2
    if (o->split_cands == 0) o->split_cands = SPLIT_CANDS;
    if (o->split_rows != 1) o->split_rows = 0;
    if (o->no_kernel_timing != 1) o->no_kernel_timing = 0;
}

extern "C" int cfpq_create(int device, const cfpq_opts *opts, cfpq_ctx **out) {
    if (!out) return CFPQ_E_INVALID;
    *out = nullptr;
    cfpq_ctx *ctx = new (std::nothrow) cfpq_ctx();
    if (!ctx) return CFPQ_E_NOMEM;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        delete ctx;
        return CFPQ_E_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        delete ctx;
        return CFPQ_E_INVALID;
    }
    ctx->device = device;
    if (opts) ctx->opts = *opts;
    else cfpq_default_opts(&ctx->opts);
    normalise(&ctx->opts);
    auto init = [&]() -> int {
        HIPCHK(hipSetDevice(device));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        ctx->num_cu = prop.multiProcessorCount;
        noalloc_init(ctx, device);
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void **)&ctx->ring, sizeof(StepCounters) * RING));
        HIPCHK(hipMemset(ctx->ring, 0, sizeof(StepCounters) * RING));
        HIPCHK(hipHostMalloc((void **)&ctx->ring_host, sizeof(StepCounters) * RING, hipHostMallocDefault));
        HIPCHK(hipMalloc((void **)&ctx->scratch64, sizeof(unsigned long long) * 4));
        HIPCHK(hipHostMalloc((void **)&ctx->scratch64_host, sizeof(unsigned long long) * 4, hipHostMallocDefault));
        return CFPQ_OK;
    };
    int rc = init();
    if (rc != CFPQ_OK) {
        cfpq_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return CFPQ_OK;
}

static void free_mat(cfpq_mat *m) {
    if (!m) return;
    if (m->base) (void)hipFree(m->base);
    delete m;
}

extern "C" int cfpq_destroy(cfpq_ctx *ctx) {
    if (!ctx) return CFPQ_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->pool) free_mat(kv.second);
    ctx->pool.clear();
    for (auto &b : ctx->pair_pool) (void)hipFree(b.first);
    ctx->pair_pool.clear();
    for (auto e : ctx->ev) (void)hipEventDestroy(e);
    for (auto g : ctx->graph_exec)
        if (g) (void)hipGraphExecDestroy(g);
    void *dev[] = {ctx->stage, ctx->ring, ctx->scratch64, ctx->items, ctx->units, ctx->mitems, ctx->lists[0], ctx->lists[1],
                   ctx->partials, ctx->pmask, ctx->splits,
                   ctx->d_mats, ctx->d_heads, ctx->d_terms, ctx->d_tptr, ctx->d_ids, ctx->sched_w, ctx->dirty, ctx->dirty_off,
                   ctx->last_eval,
                   ctx->tr_x, ctx->tr_y, ctx->tr_rule, ctx->tr_item, ctx->tr_count, ctx->t_master, ctx->t_m1,
                   ctx->t_m2, ctx->t_grad, ctx->t_loss, ctx->t_x, ctx->t_y, ctx->t_rule};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (ctx->ring_host) (void)hipHostFree(ctx->ring_host);
    if (ctx->scratch64_host) (void)hipHostFree(ctx->scratch64_host);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CFPQ_OK;
}

extern "C" const char *cfpq_last_error(const cfpq_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int cfpq_set_opts(cfpq_ctx *ctx, const cfpq_opts *opts) {
    if (!ctx || !opts) return CFPQ_E_INVALID;
    ctx->opts = *opts;
    normalise(&ctx->opts);
    return CFPQ_OK;
}

extern "C" int cfpq_sync(cfpq_ctx *ctx) {
    if (!ctx) return CFPQ_E_INVALID;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------

// Bytes of one matrix with T tiles per side in its single allocation, and
// (with base) the descriptor of a matrix laid out at base.
static size_t mat_layout(uint32_t Tn, char *base, MatDev *d) {
    const size_t T = Tn, ow = occ_words(Tn);
    const size_t b_bits = align256(T * T * TILE_WORDS * sizeof(uint64_t));
    const size_t b_cnt = align256(T * T * sizeof(uint32_t));
    const size_t b_rm = align256(T * T * 4 * sizeof(uint64_t));
    const size_t b_cq = align256(T * T * QUARTERS * TROW_WORDS * sizeof(uint64_t));
    const size_t b_occ = align256(OCC_ARRAYS * T * ow * sizeof(uint32_t));  // occR, occC, occQ, occD, occX
    const size_t b_ct = b_bits;  // the quarters in column-major form: as many words as the tiles
    const size_t b_st = b_cnt;   // sweep stamps, one word per tile
    if (base && d) {
        d->bits = (uint64_t *)base;
        d->cnt = (uint32_t *)(base + b_bits);
        d->rmask = (uint64_t *)(base + b_bits + b_cnt);
        d->colq = (uint64_t *)(base + b_bits + b_cnt + b_rm);
        d->colT = (uint64_t *)(base + b_bits + b_cnt + b_rm + b_cq + b_occ);
        d->occR = (uint32_t *)(base + b_bits + b_cnt + b_rm + b_cq);
        d->occC = d->occR + T * ow;
        d->occQ = d->occC + T * ow;
        d->occD = d->occQ + QUARTERS * T * ow;
        d->occX = d->occD + T * ow;
        d->stamp = (uint32_t *)(base + b_bits + b_cnt + b_rm + b_cq + b_occ + b_ct);
        d->T = Tn;
        d->ow = (uint32_t)ow;
    }
    return b_bits + b_cnt + b_rm + b_cq + b_occ + b_ct + b_st;
}

static int mat_create(cfpq_ctx *ctx, uint32_t n, cfpq_mat **out) {
    if (ctx->no_alloc) return fail(ctx, CFPQ_E_NOMEM, "a call that allocates nothing needed a device matrix");
    cfpq_mat *m = new (std::nothrow) cfpq_mat();
    if (!m) return fail(ctx, CFPQ_E_NOMEM, "host allocation failed");
    m->n = n;
    m->T = tiles_per_side(n);
    m->bytes = mat_layout(m->T, nullptr, nullptr);
    hipError_t e = hipMalloc(&m->base, m->bytes);
    if (e != hipSuccess) {
        delete m;
        return fail(ctx, e == hipErrorOutOfMemory ? CFPQ_E_NOMEM : CFPQ_E_DEVICE, "hipMalloc(%zu): %s", (size_t)0,
                    hipGetErrorString(e));
    }
    e = hipMemsetAsync(m->base, 0, m->bytes, ctx->stream);
    if (e != hipSuccess) {
        free_mat(m);
        return fail(ctx, CFPQ_E_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
    }
    mat_layout(m->T, (char *)m->base, &m->d);
    *out = m;
    return CFPQ_OK;
}

extern "C" int cfpq_mat_alloc(cfpq_ctx *ctx, uint32_t n, cfpq_mat **out) {
    if (!ctx || !out || n == 0) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_alloc: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    return mat_create(ctx, n, out);
}

extern "C" int cfpq_mat_free(cfpq_ctx *ctx, cfpq_mat *mat) {
    if (!ctx || !mat) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_free: bad argument");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_mat(mat);
    return CFPQ_OK;
}

extern "C" uint32_t cfpq_mat_n(const cfpq_mat *mat) { return mat ? mat->n : 0; }

static int ensure_stage(cfpq_ctx *ctx, size_t bytes) {
    if (bytes <= ctx->stage_bytes) return CFPQ_OK;
    if (ctx->no_alloc) return fail(ctx, CFPQ_E_NOMEM, "a call that allocates nothing needed a staging buffer");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->stage) HIPCHK(hipFree(ctx->stage));
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
    HIPCHK(hipMalloc(&ctx->stage, bytes));
    ctx->stage_bytes = bytes;
    return CFPQ_OK;
}

static uint32_t grid_for_tiles(cfpq_ctx *ctx, size_t tiles) {
    size_t g = std::min<size_t>(tiles, (size_t)ctx->num_cu * 16);
    return (uint32_t)std::max<size_t>(g, 1);
}

// Before a call reads m as a left operand: rebuild its column-major form if a
// fixpoint that did not keep it left it out of date.
static int colT_fresh(cfpq_ctx *ctx, const cfpq_mat *cm) {
    cfpq_mat *m = const_cast<cfpq_mat *>(cm);
    if (!m->colT_stale) return CFPQ_OK;
    hipLaunchKernelGGL(k_rebuild_colT, dim3(grid_for_tiles(ctx, (size_t)m->T * m->T)), dim3(BLOCK), 0, ctx->stream, m->d);
    CHK(launch_check(ctx, "k_rebuild_colT"));
    m->colT_stale = false;
    return CFPQ_OK;
}

extern "C" int cfpq_mat_upload(cfpq_ctx *ctx, cfpq_mat *mat, const uint64_t *rows, size_t words_per_row) {
    if (!ctx || !mat || !rows) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_upload: bad argument");
    const size_t wpr = (mat->n + 63) / 64;
    if (words_per_row < wpr) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_upload: words_per_row %zu < %zu", words_per_row, wpr);
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure_stage(ctx, (size_t)mat->n * wpr * 8));
    HIPCHK(hipMemcpy2DAsync(ctx->stage, wpr * 8, rows, words_per_row * 8, wpr * 8, mat->n, hipMemcpyHostToDevice,
                            ctx->stream));
    const size_t occ_bytes = (size_t)OCC_ARRAYS * mat->T * mat->d.ow * 4;
    HIPCHK(hipMemsetAsync(mat->d.occR, 0, occ_bytes, ctx->stream));
    hipLaunchKernelGGL(k_rows_to_tiles, dim3(grid_for_tiles(ctx, (size_t)mat->T * mat->T)), dim3(BLOCK), 0, ctx->stream,
                       mat->d, (const uint64_t *)ctx->stage, mat->n, (uint32_t)wpr);
    CHK(launch_check(ctx, "k_rows_to_tiles"));
    mat->colT_stale = false;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

extern "C" int cfpq_mat_download(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *rows, size_t words_per_row) {
    if (!ctx || !mat || !rows) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_download: bad argument");
    const size_t wpr = (mat->n + 63) / 64;
    if (words_per_row < wpr) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_download: words_per_row %zu < %zu", words_per_row, wpr);
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure_stage(ctx, (size_t)mat->n * wpr * 8));
    hipLaunchKernelGGL(k_tiles_to_rows, dim3(grid_for_tiles(ctx, (size_t)mat->T * mat->T)), dim3(BLOCK), 0, ctx->stream,
                       mat->d, (uint64_t *)ctx->stage, mat->n, (uint32_t)wpr);
    CHK(launch_check(ctx, "k_tiles_to_rows"));
    HIPCHK(hipMemcpy2DAsync(rows, words_per_row * 8, ctx->stage, wpr * 8, wpr * 8, mat->n, hipMemcpyDeviceToHost,
                            ctx->stream));
    if (words_per_row > wpr)
        for (uint32_t i = 0; i < mat->n; ++i) std::memset(rows + (size_t)i * words_per_row + wpr, 0, (words_per_row - wpr) * 8);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

extern "C" int cfpq_mat_set_bits(cfpq_ctx *ctx, cfpq_mat *mat, const uint32_t *rows, const uint32_t *cols, size_t count) {
    if (!ctx || !mat || (count && (!rows || !cols))) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_set_bits: bad argument");
    for (size_t q = 0; q < count; ++q)
        if (rows[q] >= mat->n || cols[q] >= mat->n)
            return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_set_bits: bit (%u, %u) outside %u x %u", rows[q], cols[q], mat->n, mat->n);
    if (!count) return CFPQ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure_stage(ctx, count * 8));
    uint32_t *dr = (uint32_t *)ctx->stage, *dc = dr + count;
    HIPCHK(hipMemcpyAsync(dr, rows, count * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dc, cols, count * 4, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t g = (uint32_t)std::min<size_t>((count + 255) / 256, (size_t)ctx->num_cu * 32);
    hipLaunchKernelGGL(k_set_bits, dim3(g), dim3(256), 0, ctx->stream, mat->d, (const uint32_t *)dr, (const uint32_t *)dc, count);
    CHK(launch_check(ctx, "k_set_bits"));
    const size_t occ_bytes = (size_t)OCC_ARRAYS * mat->T * mat->d.ow * 4;
    HIPCHK(hipMemsetAsync(mat->d.occR, 0, occ_bytes, ctx->stream));
    hipLaunchKernelGGL(k_recount, dim3(grid_for_tiles(ctx, (size_t)mat->T * mat->T)), dim3(BLOCK), 0, ctx->stream, mat->d);
    CHK(launch_check(ctx, "k_recount"));
    mat->colT_stale = false;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

extern "C" int cfpq_mat_popcount(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *count) {
    if (!ctx || !mat || !count) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_popcount: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemsetAsync(ctx->scratch64, 0, 8, ctx->stream));
    const size_t tiles = (size_t)mat->T * mat->T;
    const uint32_t g = (uint32_t)std::min<size_t>((tiles + 255) / 256, 1024);
    hipLaunchKernelGGL(k_popcount, dim3(g), dim3(256), 0, ctx->stream, mat->d, ctx->scratch64);
    CHK(launch_check(ctx, "k_popcount"));
    HIPCHK(hipMemcpyAsync(ctx->scratch64_host, ctx->scratch64, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *count = ctx->scratch64_host[0];
    return CFPQ_OK;
}

extern "C" int cfpq_mat_fingerprint(cfpq_ctx *ctx, const cfpq_mat *mat, uint64_t *hash, uint64_t *count) {
    if (!ctx || !mat || !hash || !count) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_fingerprint: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t n = mat->n;
    const uint64_t P = 1000000007ull;
    std::vector<uint64_t> pw(2 * (size_t)n);
    uint64_t a = 1, b = 1;
    for (uint32_t e = 0; e < n; ++e) {
        pw[e] = a;
        pw[n + e] = b;
        a = a * 11 % P;
        b = b * 13 % P;
    }
    CHK(ensure_stage(ctx, pw.size() * 8));
    HIPCHK(hipMemcpyAsync(ctx->stage, pw.data(), pw.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->scratch64, 0, 16, ctx->stream));
    const uint64_t *p11 = (const uint64_t *)ctx->stage;
    hipLaunchKernelGGL(k_fingerprint, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, ctx->stream, mat->d, n, p11, p11 + n,
                       ctx->scratch64);
    CHK(launch_check(ctx, "k_fingerprint"));
    HIPCHK(hipMemcpyAsync(ctx->scratch64_host, ctx->scratch64, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *hash = ctx->scratch64_host[0] % P;
    *count = ctx->scratch64_host[1];
    return CFPQ_OK;
}

extern "C" int cfpq_mat_copy(cfpq_ctx *ctx, cfpq_mat *dst, const cfpq_mat *src) {
    if (!ctx || !dst || !src || dst->n != src->n) return fail(ctx, CFPQ_E_INVALID, "cfpq_mat_copy: bad argument");
    if (dst == src) return CFPQ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    CHK(grow(ctx, &ctx->d_mats, &ctx->d_mats_cap, 2));
    CHK(grow(ctx, &ctx->d_ids, &ctx->d_ids_cap, 2));
    const MatDev md[2] = {dst->d, src->d};
    const uint32_t ids[2] = {0, 1};
    HIPCHK(hipMemcpyAsync(ctx->d_mats, md, sizeof md, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_ids, ids, sizeof ids, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t g = grid_for_tiles(ctx, (size_t)dst->T * dst->T);
    hipLaunchKernelGGL(k_clear_full, dim3(g, 1), dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats,
                       (const uint32_t *)ctx->d_ids, 1u);
    hipLaunchKernelGGL(k_copy, dim3(g, 1), dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats,
                       (const uint32_t *)ctx->d_ids, (const uint32_t *)ctx->d_ids + 1, 1u);
    CHK(launch_check(ctx, "cfpq_mat_copy"));
    dst->colT_stale = src->colT_stale;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

// ---------------------------------------------------------------------------
// fixpoint machinery shared by add_mul, fixpoint and batch fixpoint
// ---------------------------------------------------------------------------

static int pool_take(cfpq_ctx *ctx, uint32_t n, cfpq_mat **out) {
    auto it = ctx->pool.find(n);
    if (it != ctx->pool.end()) {
        *out = it->second;
        ctx->pool.erase(it);
        return CFPQ_OK;
    }
    return mat_create(ctx, n, out);
}

static void pool_give(cfpq_ctx *ctx, cfpq_mat *m) { ctx->pool.emplace(m->n, m); }

static int ensure_events(cfpq_ctx *ctx, size_t n) {
    if (ctx->no_alloc && ctx->ev.size() < n) return fail(ctx, CFPQ_E_NOMEM, "a call that allocates nothing needed events");
    while (ctx->ev.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ctx->ev.push_back(e);
    }
    return CFPQ_OK;
}

// A fixpoint problem laid out for the kernels.
struct Plan {
    std::vector<MatDev> mats;         // descriptor table
    std::vector<HeadDev> heads;
    std::vector<TermDev> terms_delta;  // semi-naive terms (two per rule)
    std::vector<TermDev> terms_full;   // full terms (one per rule), stored after the delta terms
    // the semi-naive terms whose delta operand can be non-empty (its
    // nonterminal heads some rule; a terminal-only one never changes after
    // the first step, which evaluates full terms), stored after the full terms
    std::vector<TermDev> terms_live;
    std::vector<uint32_t> all_delta[2];  // ids of every delta buffer of parity q (buffer Dq)
    size_t max_tiles_per_head = 0;     // max T*T over heads
    size_t total_head_tiles = 0;       // sum of T*T over heads
    size_t max_occ_words = 0;          // max T*ow over matrices
    size_t ids_all_off = 0;            // offset of the "every matrix" id list in d_ids
    std::vector<uint32_t> dirty_off;   // per head: word offset of its deferred-tile bitmap
    size_t dirty_words = 0;
    bool rule_heads = false;           // one head entry per rule (rule sweeps); entries share targets
};

// The plan with one head entry per rule, for rule sweeps: entry (h, r) has
// head h's target and delta buffers and rule r's full term only.  Entries of
// one head share a target, so they never run in the same step.
static Plan rule_plan(const Plan &plan) {
    Plan rp = plan;
    rp.heads.clear();
    rp.total_head_tiles = 0;
    for (const HeadDev &hd : plan.heads)
        for (uint32_t t = hd.f0; t < hd.f1; ++t) {
            HeadDev rh = hd;
            rh.f0 = t;
            rh.f1 = t + 1;
            rh.t1 = rh.t0;  // no delta terms: every step of a sweep is a full step
            rh.p1 = rh.p0;
            rp.heads.push_back(rh);
            rp.total_head_tiles += (size_t)hd.T * hd.T;
        }
    rp.rule_heads = true;
    return rp;
}

// Waves of the retire launch per matrix whose occupancy bitsets it zeroes:
// one per 128 words of the largest matrix (2 words per lane), at most 32.
static uint32_t occ_blocks_per_matrix(const Plan &plan) {
    return (uint32_t)std::min<size_t>(32, std::max<size_t>(1, (plan.max_occ_words + 127) / 128));
}

// terms in the device table (and per parity in the operand-pointer table)
static size_t plan_terms(const Plan &plan) {
    return plan.terms_delta.size() + plan.terms_full.size() + plan.terms_live.size();
}

struct StepStats {
    uint64_t new_bits = 0, items = 0, tp = 0, tpd = 0, tpm = 0, arows = 0, brows = 0, orows = 0, brows_lds = 0,
             lookups = 0, mfma_instr = 0, cand_hits = 0, arows_d = 0, brows_d = 0, orows_d = 0, acols = 0;
    uint64_t units_done = 0, units_idle = 0, unit_clk = 0, idle_clk = 0, idle_brows = 0;
    uint64_t tpu = 0;
    uint32_t max_unit_cands = 0;
};

// A-tile popcount thresholds of the product kernels.  A (head, C tile) item
// goes to the MFMA kernel if some term has an A tile with >= mfma_threshold
// bits, else to the dense kernel if some A tile has >= dense_threshold bits,
// else to the sparse kernel.  In the dense kernel, A tiles below
// dense_threshold gather B rows from HBM, from there LDS-staged OR-of-rows,
// from m4r_threshold the Four Russians tables.
static uint32_t dense_threshold(const cfpq_opts &o) {
    switch (o.kernel) {
    case CFPQ_KERNEL_GATHER: case CFPQ_KERNEL_MFMA: return 0xffffffffu;
    case CFPQ_KERNEL_LDS: case CFPQ_KERNEL_M4R: return 1u;
    default: return std::max<uint32_t>(o.dense_bits, 1u);
    }
}
static uint32_t mfma_threshold(const cfpq_opts &o) {
    switch (o.kernel) {
    case CFPQ_KERNEL_MFMA: return 1u;
    case CFPQ_KERNEL_AUTO: return std::max<uint32_t>(o.mfma_bits, 1u);
    default: return 0xffffffffu;
    }
}
static uint32_t m4r_threshold(const cfpq_opts &o) {
    switch (o.kernel) {
    case CFPQ_KERNEL_GATHER: case CFPQ_KERNEL_LDS: return 0xffffffffu;
    case CFPQ_KERNEL_M4R: return 0u;
    case CFPQ_KERNEL_MFMA: return 0xffffffffu;
    default: return o.m4r_bits;
    }
}

// Enqueue one step. parity q: deltas read from D_q, new bits to D_{1-q}.
// prev_slot < 0: the delta consumed by this step is the initial one
// (cleared in full); otherwise it was written by the step in prev_slot.
static bool sched_on(const cfpq_ctx *ctx) { return ctx->opts.sched != 0 && ctx->sched_have; }
static bool order_on(const cfpq_ctx *ctx) { return ctx->opts.sched_order != 0 && ctx->sched_have; }

// Sub-units (partial-row slots) the build may create per step.
static uint32_t split_slots_of(const cfpq_ctx *ctx, const Plan &plan) {
    return (uint32_t)std::min<size_t>(QUARTERS * plan.total_head_tiles, ctx->split_slots);
}

static SchedArgs sched_args(const cfpq_ctx *ctx, bool flush) {
    SchedArgs a = {};
    a.w = (sched_on(ctx) || order_on(ctx)) ? ctx->sched_w : nullptr;
    a.order = order_on(ctx) && !flush;
    const double po = std::min(0.999999, std::max(1e-6, ctx->opts.order_threshold / 65536.0));
    a.order_logit = (float)std::log(po / (1.0 - po));
    a.dirty = ctx->dirty;
    a.dirty_off = ctx->dirty_off;
    const double pth = std::min(0.999999, std::max(1e-6, ctx->opts.sched_threshold / 65536.0));
    a.logit_threshold = (float)std::log(pth / (1.0 - pth));
    a.defer = sched_on(ctx) && !flush;
    a.flush = flush;
    if (ctx->tr_cap) {
        a.trace_x = ctx->tr_x;
        a.trace_rule = ctx->tr_rule;
        a.trace_item = ctx->tr_item;
        a.trace_y = ctx->tr_y;
        a.trace_count = ctx->tr_count;
        a.trace_cap = (uint32_t)ctx->tr_cap;
    }
    return a;
}

// fuse: merge new rows into the head matrices inside the product kernels
// (fixpoint steps); add_mul keeps a separate apply pass so that aliased
// operands are read as they were on entry.
// zero_counters: the step zeroes its own counter block (a run's first step,
// or a slot the previous step's retire launch did not zero for it).
// head >= 0: only that head's items are built (a head-sweep step).
// Sweeps (head >= 0) pass stamp_step (the step number from 1): the retire
// stamps the head tiles that gained bits with it and records it as the
// entry's last evaluation; with `unchanged`, the dense kernel skips
// tile-products whose operand tiles both carry older stamps than that.
static int enqueue_step(cfpq_ctx *ctx, const Plan &plan, int q, uint32_t slot, int prev_slot, bool full_pass,
                        int ev_idx, bool flush, bool fuse, bool zero_counters, int head = -1,
                        uint32_t stamp_step = 0, bool unchanged = false) {
    StepCounters *c = ctx->ring + slot;
    const uint32_t nh = head >= 0 ? 1u : (uint32_t)plan.heads.size();
    const uint32_t h0 = head >= 0 ? (uint32_t)head : 0u;
    const TermDev *terms = ctx->d_terms;
    uint64_t *touched = ctx->lists[slot & 1];
    const uint32_t seg_cap = (uint32_t)plan.total_head_tiles;
    // a step's counters are zeroed by the previous step's retire launch,
    // except where the caller says otherwise
    if (zero_counters) hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, ctx->stream, ctx->ring, slot);
    // with the dense and MFMA paths both off, every item is sparse: the
    // heavy-tile bitsets are neither built nor read, and the dense and MFMA
    // kernels are not launched
    const int heavy = (dense_threshold(ctx->opts) <= TILE_BITS || mfma_threshold(ctx->opts) <= TILE_BITS) ? 1 : 0;
    // dense-tile bitsets of every matrix of the plan (operands of this step)
    const uint32_t nall = (uint32_t)plan.mats.size();
    const uint32_t *ids_all = ctx->d_ids + plan.ids_all_off;
    if (heavy)
        hipLaunchKernelGGL(k_dense_occ, dim3(std::max<uint32_t>(1, (uint32_t)((plan.max_occ_words + 255) / 256)),
                                             std::min<uint32_t>(nall, 65535)),
                           dim3(256), 0, ctx->stream, (const MatDev *)ctx->d_mats, ids_all, nall,
                           dense_threshold(ctx->opts), mfma_threshold(ctx->opts));
    const uint32_t bx = (uint32_t)((plan.max_tiles_per_head + BLOCK - 1) / BLOCK);
    // heavy sparse units are split over several waves (lane-parallel kernel,
    // head update not fused); the units buffer holds 4 entries per head tile
    // plus as many extra sub-units
    const int lane_parallel = ctx->opts.sparse_mode != 1;  // modes 0, 3, 2: quarter, half-tile, tile units
    // work-unit shape of the lane-parallel kernel: 2 = whole tiles, 3 = halves
    const int tile_units = ctx->opts.sparse_mode == 2 ? 1 : ctx->opts.sparse_mode == 3 ? 2 : 0;
    // candidate count above which a quarter is split (0: no splitting)
    // (0xffffffff: never split, so no merge launch either)
    const uint32_t split = (lane_parallel && !(fuse && ctx->opts.fuse_apply) && ctx->opts.split_cands != 0xffffffffu)
                               ? ctx->opts.split_cands : 0u;
    const uint32_t piece = ctx->opts.split_piece ? ctx->opts.split_piece : std::max<uint32_t>(split, 1u);
    const uint32_t split_budget = split_slots_of(ctx, plan);
    const uint32_t heavy_base = (uint32_t)(QUARTERS * plan.total_head_tiles);
    const bool use_sched =
        head < 0 && !full_pass && plan.dirty_words && (sched_on(ctx) || order_on(ctx) || ctx->tr_cap || flush);
    // the scheduler's first-phase list (its ranked units), after the build's lists
    uint64_t *prio_list = ctx->units + (size_t)2 * QUARTERS * plan.total_head_tiles;
    const bool ordered = use_sched && order_on(ctx) && !flush;
    const SchedArgs sa = sched_args(ctx, flush);
    if (use_sched) {
        if (ctx->tr_cap)
            hipLaunchKernelGGL(k_trace_begin, dim3(1), dim3(64), 0, ctx->stream, c, (const uint32_t *)ctx->tr_count,
                               (uint32_t)ctx->tr_cap);
        hipLaunchKernelGGL(k_build_sched, dim3(bx, nh), dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats,
                           (const HeadDev *)ctx->d_heads, terms, q, c, ctx->units, ctx->items, ctx->mitems, sa, split,
                           piece, split_budget, heavy, heavy_base, tile_units, ctx->splits, prio_list,
                           ctx->opts.split_rows);
    } else {
        hipLaunchKernelGGL(k_build, dim3(bx, nh), dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats,
                           (const HeadDev *)ctx->d_heads, terms, q, (int)full_pass, c, ctx->units, ctx->items,
                           ctx->mitems, split, piece, split_budget, heavy, heavy_base, h0,
                           (const StepCounters *)(prev_slot >= 0 && !full_pass ? ctx->ring + prev_slot : nullptr),
                           tile_units, ctx->splits, ctx->opts.split_rows);
    }
    const uint32_t pgrid = (uint32_t)ctx->num_cu * 4;
    // product kernels, timed by consecutive events of slot i: sparse
    // ev[5i]..ev[5i+1], merge of split quarters ..ev[5i+2], dense ..ev[5i+3],
    // MFMA ..ev[5i+4]
    hipEvent_t *ev = ev_idx >= 0 ? &ctx->ev[EV_SLOT * (size_t)ev_idx] : nullptr;
    if (ev) HIPCHK(hipEventRecord(ev[0], ctx->stream));
    // (no more waves than the step can have units: small problems launch
    // small grids, which dispatch faster)
    const uint64_t max_units = (uint64_t)QUARTERS * plan.total_head_tiles + split_budget;
    const dim3 sgrid((uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)ctx->num_cu * (uint64_t)ctx->opts.sparse_waves_per_cu, max_units)));
    const uint32_t chunk = (uint32_t)ctx->opts.sparse_chunk;
    const MatDev *dm = ctx->d_mats;
    const HeadDev *dh = ctx->d_heads;
    const uint64_t *un = ctx->units;
    const int fz = (fuse && ctx->opts.fuse_apply) ? 1 : 0;
    fuse = fz != 0;
    if (lane_parallel) {
        const TermPtr *tp = ctx->d_tptr + (size_t)q * plan_terms(plan);
        // with the scheduler's ranking, phase 1 evaluates its ranked units and
        // phase 0 the rest; the launch boundary makes phase 1's head updates
        // (fused apply) visible to every unit of phase 0
        for (int phase = ordered ? 1 : 0; phase >= 0; --phase) {
            const bool rs = ctx->opts.split_rows != 0;
            auto *kern = tile_units ? (rs ? k_product_sparse_lp<true, true> : k_product_sparse_lp<true, false>)
                                    : (rs ? k_product_sparse_lp<false, true> : k_product_sparse_lp<false, false>);
            hipLaunchKernelGGL(kern, sgrid, dim3(WAVE), 0, ctx->stream, dm, dh, tp, q, c, un, heavy_base,
                               TouchedLists{touched, seg_cap}, chunk, fz, ctx->partials, ctx->pmask, phase,
                               (const uint64_t *)prio_list);
        }
        if (ev) HIPCHK(hipEventRecord(ev[1], ctx->stream));
        // the split quarters, from their sub-units' partial rows (row-split
        // sub-units write their rows themselves)
        if (split && !ctx->opts.split_rows)
            // (a grid of 4 waves per CU measured best: larger grids cost more
            // in dispatch than the merge work they spread out)
            hipLaunchKernelGGL(k_merge_split, dim3((uint32_t)ctx->num_cu * 4), dim3(WAVE), 0, ctx->stream, dm, dh, q, c,
                               (const uint64_t *)ctx->splits, (const uint64_t *)ctx->partials,
                               (const uint64_t *)ctx->pmask, TouchedLists{touched, seg_cap});
    } else if (ctx->opts.sparse_batch == 1) {
        hipLaunchKernelGGL(k_product_sparse<1>, sgrid, dim3(WAVE), 0, ctx->stream, dm, dh, terms, q, c, un, touched, chunk,
                           fz);
    } else if (ctx->opts.sparse_batch == 2) {
        hipLaunchKernelGGL(k_product_sparse<2>, sgrid, dim3(WAVE), 0, ctx->stream, dm, dh, terms, q, c, un, touched, chunk,
                           fz);
    } else {
        hipLaunchKernelGGL(k_product_sparse<4>, sgrid, dim3(WAVE), 0, ctx->stream, dm, dh, terms, q, c, un, touched, chunk,
                           fz);
    }
    if (ev && !lane_parallel) HIPCHK(hipEventRecord(ev[1], ctx->stream));
    if (ev) HIPCHK(hipEventRecord(ev[2], ctx->stream));
    // (the dense and MFMA launches, and their events, only with a heavy-tile
    // path on; each kernel only when its own crossover can route a tile to it)
    if (heavy && dense_threshold(ctx->opts) <= TILE_BITS)
        hipLaunchKernelGGL(k_product_dense, dim3(pgrid), dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats,
                           (const HeadDev *)ctx->d_heads, terms, q, c, (const uint64_t *)ctx->items, touched,
                           dense_threshold(ctx->opts), m4r_threshold(ctx->opts), fz, ctx->opts.m4r_keys == 8 ? 1 : 0,
                           (const uint32_t *)(unchanged && head >= 0 && stamp_step ? ctx->last_eval : nullptr));
    if (ev && heavy) HIPCHK(hipEventRecord(ev[3], ctx->stream));
    if (heavy && mfma_threshold(ctx->opts) <= TILE_BITS)
        hipLaunchKernelGGL(k_product_mfma, dim3((uint32_t)ctx->num_cu), dim3(MFMA_BLOCK), 0, ctx->stream,
                           (const MatDev *)ctx->d_mats, (const HeadDev *)ctx->d_heads, terms, q, c,
                           (const uint64_t *)ctx->mitems, touched, fz);
    if (ev && heavy) HIPCHK(hipEventRecord(ev[4], ctx->stream));
    if (use_sched && ctx->tr_cap)
        hipLaunchKernelGGL(k_trace_label, dim3((uint32_t)ctx->num_cu), dim3(256), 0, ctx->stream,
                           (const MatDev *)ctx->d_mats, (const HeadDev *)ctx->d_heads, q, (const StepCounters *)c, sa);
    // the end of the step in one launch: apply (unless fused into the
    // products), retire the delta this step consumed (buffer D_q: rows,
    // metadata, occupancy) and zero the next slot's counters
    const uint32_t *ids_q = ctx->d_ids + (q ? plan.all_delta[0].size() : 0);
    const uint32_t nids_q = (uint32_t)plan.all_delta[q].size();
    const bool retire_delta = prev_slot >= 0;
    // the apply waves run in both modes: without fused apply they merge the
    // new rows into the heads, and either way they build the new quarters'
    // column-major form
    // (at most one wave per tile of the two lists)
    const uint32_t nw = (uint32_t)std::min<size_t>((size_t)ctx->num_cu * 32, std::max<size_t>(1, 2 * plan.total_head_tiles));
    const uint32_t nido = retire_delta ? nids_q : 0u, per = occ_blocks_per_matrix(plan);
    const StepCounters *prev = retire_delta ? ctx->ring + prev_slot : c;
    const TouchedLists ptl{ctx->lists[retire_delta ? (prev_slot & 1) : (slot & 1)], seg_cap};
    hipLaunchKernelGGL(k_retire, dim3(nw + nido * per + 1), dim3(WAVE), 0, ctx->stream,
                       (const MatDev *)ctx->d_mats, (const HeadDev *)ctx->d_heads, q, (const StepCounters *)c,
                       TouchedLists{touched, seg_cap}, prev, ptl, ids_q, nido, per, nw, 1, retire_delta ? 1 : 0,
                       ctx->ring + (slot + 1) % RING, fuse ? 0 : 1, head >= 0 ? stamp_step : 0u, head,
                       head >= 0 && stamp_step ? ctx->last_eval : nullptr);
    if (prev_slot < 0 && !full_pass) {
        // a first step over delta terms: its delta buffers are cleared in full
        hipLaunchKernelGGL(k_clear_full, dim3(std::max<uint32_t>(1, ctx->num_cu * 4 / std::max<uint32_t>(1, nids_q)),
                                              std::min<uint32_t>(nids_q, 65535)),
                           dim3(BLOCK), 0, ctx->stream, (const MatDev *)ctx->d_mats, ids_q, nids_q);
    }
    return launch_check(ctx, "fixpoint step");
}

static int upload_plan(cfpq_ctx *ctx, Plan &plan) {
    plan.ids_all_off = plan.all_delta[0].size() + plan.all_delta[1].size();
    plan.max_occ_words = 0;
    for (const MatDev &m : plan.mats) plan.max_occ_words = std::max<size_t>(plan.max_occ_words, (size_t)m.T * m.ow);
    CHK(grow(ctx, &ctx->d_mats, &ctx->d_mats_cap, plan.mats.size()));
    CHK(grow(ctx, &ctx->d_heads, &ctx->d_heads_cap, plan.heads.size()));
    std::vector<HeadDev> heads(plan.heads);
    // device table: delta terms, then full terms, then live delta terms
    const uint32_t nfull_end = (uint32_t)(plan.terms_delta.size() + plan.terms_full.size());
    for (HeadDev &hd : heads) {
        hd.f0 += (uint32_t)plan.terms_delta.size();
        hd.f1 += (uint32_t)plan.terms_delta.size();
        hd.t0 += nfull_end;
        hd.t1 += nfull_end;
    }
    CHK(grow(ctx, &ctx->d_terms, &ctx->d_terms_cap, plan_terms(plan)));
    const size_t nid = plan.all_delta[0].size() + plan.all_delta[1].size() + plan.mats.size();
    CHK(grow(ctx, &ctx->d_ids, &ctx->d_ids_cap, nid));
    CHK(grow(ctx, &ctx->items, &ctx->items_cap, plan.total_head_tiles));
    // units: up to QUARTERS per head tile, then the sub-units of split
    // quarters (at most the split budget plus one per quarter)
    CHK(grow(ctx, &ctx->units, &ctx->units_cap, 3 * QUARTERS * plan.total_head_tiles));
    CHK(grow(ctx, &ctx->mitems, &ctx->mitems_cap, plan.total_head_tiles));
    {
        const size_t slots = split_slots_of(ctx, plan);
        CHK(grow(ctx, &ctx->partials, &ctx->partials_cap, slots * WAVE * TROW_WORDS));
        CHK(grow(ctx, &ctx->pmask, &ctx->pmask_cap, slots));
        CHK(grow(ctx, &ctx->splits, &ctx->splits_cap, 2 * (slots / 2 + 1)));
    }
    // touched-tile lists: 1 + SHARDS segments of one entry per head tile
    const size_t list_need = (size_t)(1 + SHARDS) * plan.total_head_tiles;
    if (list_need > ctx->lists_cap || !ctx->lists[0]) {
        size_t cap0 = ctx->lists_cap;
        CHK(grow(ctx, &ctx->lists[0], &cap0, list_need));
        size_t cap1 = ctx->lists_cap;
        CHK(grow(ctx, &ctx->lists[1], &cap1, list_need));
        ctx->lists_cap = std::min(cap0, cap1);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_mats, plan.mats.data(), plan.mats.size() * sizeof(MatDev), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_heads, heads.data(), heads.size() * sizeof(HeadDev), hipMemcpyHostToDevice, ctx->stream));
    std::vector<TermDev> terms(plan.terms_delta);
    terms.insert(terms.end(), plan.terms_full.begin(), plan.terms_full.end());
    terms.insert(terms.end(), plan.terms_live.begin(), plan.terms_live.end());
    HIPCHK(hipMemcpyAsync(ctx->d_terms, terms.data(), terms.size() * sizeof(TermDev), hipMemcpyHostToDevice, ctx->stream));
    CHK(grow(ctx, &ctx->d_tptr, &ctx->d_tptr_cap, 2 * terms.size()));
    std::vector<TermPtr> tptr(2 * terms.size());
    for (int par = 0; par < 2; ++par)
        for (size_t t = 0; t < terms.size(); ++t) {
            const MatDev &L = plan.mats[terms[t].left[par]], &R = plan.mats[terms[t].right[par]];
            tptr[par * terms.size() + t] = TermPtr{L.occQ, R.occC, L.rmask, L.colq, L.colT, L.bits, R.rmask, R.bits, R.cnt};
        }
    HIPCHK(hipMemcpyAsync(ctx->d_tptr, tptr.data(), tptr.size() * sizeof(TermPtr), hipMemcpyHostToDevice, ctx->stream));
    std::vector<uint32_t> ids(plan.all_delta[0]);
    ids.insert(ids.end(), plan.all_delta[1].begin(), plan.all_delta[1].end());
    for (uint32_t x = 0; x < (uint32_t)plan.mats.size(); ++x) ids.push_back(x);  // at plan.ids_all_off
    HIPCHK(hipMemcpyAsync(ctx->d_ids, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->ring, 0, sizeof(StepCounters) * RING, ctx->stream));
    plan.dirty_off.clear();
    plan.dirty_words = 0;
    for (const HeadDev &hd : plan.heads) {
        plan.dirty_off.push_back((uint32_t)plan.dirty_words);
        plan.dirty_words += ((size_t)hd.T * hd.T + 31) / 32;
    }
    CHK(grow(ctx, &ctx->dirty, &ctx->dirty_cap, plan.dirty_words));
    CHK(grow(ctx, &ctx->dirty_off, &ctx->dirty_off_cap, plan.dirty_off.size()));
    HIPCHK(hipMemcpyAsync(ctx->dirty_off, plan.dirty_off.data(), plan.dirty_off.size() * 4, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->dirty, 0, plan.dirty_words * 4, ctx->stream));
    CHK(grow(ctx, &ctx->last_eval, &ctx->last_eval_cap, plan.heads.size()));
    HIPCHK(hipMemsetAsync(ctx->last_eval, 0, plan.heads.size() * 4, ctx->stream));
    // the tables above were copied from pageable host vectors that die here
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

static void accumulate(StepStats &s, const StepCounters &c) {
    s.new_bits += c.new_bits;
    s.items += c.n_units + c.n_heavy + c.n_dense + c.n_mfma;
    s.mfma_instr += c.mfma_instr;
    s.cand_hits += c.cand_hits;
    s.arows_d += c.a_rows_dense;
    s.brows_d += c.b_rows_dense;
    s.orows_d += c.out_rows_dense;
    s.max_unit_cands = std::max(s.max_unit_cands, c.max_unit_cands);
    s.tp += c.tile_products;
    s.tpd += c.tp_dense;
    s.tpm += c.tp_mfma;
    s.tpu += c.tp_unchanged;
    s.arows += c.a_rows;
    s.acols += c.a_cols;
    s.units_done += c.units_done;
    s.units_idle += c.idle_units;
    s.unit_clk += c.unit_clk;
    s.idle_clk += c.idle_clk;
    s.idle_brows += c.idle_brows;
    s.brows += c.b_rows;
    s.orows += c.out_rows;
    s.brows_lds += c.b_rows_lds;
    s.lookups += c.m4r_lookups;
}

// ---------------------------------------------------------------------------
// captured steps
// ---------------------------------------------------------------------------
//
// A step's launches depend only on the plan's launch configuration (table
// pointers, grid sizes, thresholds) and on the ring slot, which fixes the
// parity and the previous slot.  With use_graph, each (variant, slot) step
// is captured once into a hipGraph and replayed; the cache is dropped when
// the configuration changes.  Variants: 0 = first step of a fixpoint (clears
// the initial delta in full), 1 = delta step, 2 = exact full-pass step,
// 3 = scheduler flush step.
// Steps contain no memset nodes: counters are zeroed by a kernel (k_begin).

static std::vector<uint64_t> launch_signature(const cfpq_ctx *ctx, const Plan &plan) {
    const cfpq_opts &o = ctx->opts;
    return {(uint64_t)(uintptr_t)ctx->d_mats, (uint64_t)(uintptr_t)ctx->d_heads, (uint64_t)(uintptr_t)ctx->d_terms,
            (uint64_t)(uintptr_t)ctx->d_tptr,
            (uint64_t)(uintptr_t)ctx->d_ids, (uint64_t)(uintptr_t)ctx->units, (uint64_t)(uintptr_t)ctx->items,
            (uint64_t)(uintptr_t)ctx->mitems, (uint64_t)(uintptr_t)ctx->lists[0], (uint64_t)(uintptr_t)ctx->lists[1],
            (uint64_t)(uintptr_t)ctx->ring, (uint64_t)(uintptr_t)ctx->partials, (uint64_t)(uintptr_t)ctx->pmask,
            (uint64_t)(uintptr_t)ctx->splits, (uint64_t)ctx->split_slots, plan.heads.size(), plan.mats.size(), plan.terms_delta.size(), plan.terms_live.size(),
            plan.all_delta[0].size(), plan.all_delta[1].size(), plan.max_tiles_per_head, plan.max_occ_words,
            plan.ids_all_off, plan.total_head_tiles, (uint64_t)o.kernel, o.dense_bits, o.mfma_bits, o.m4r_bits, (uint64_t)o.sparse_batch,
            (uint64_t)o.sparse_waves_per_cu, (uint64_t)o.sparse_chunk, (uint64_t)o.sched, o.sched_threshold,
            (uint64_t)o.sparse_mode, (uint64_t)o.fuse_apply, (uint64_t)o.m4r_keys, (uint64_t)o.sched_order, (uint64_t)o.order_threshold,
            (uint64_t)o.split_cands, (uint64_t)o.split_rows, (uint64_t)o.split_piece, (uint64_t)o.no_kernel_timing,
            (uint64_t)ctx->sched_have, (uint64_t)(uintptr_t)ctx->sched_w, (uint64_t)(uintptr_t)ctx->dirty,
            (uint64_t)(uintptr_t)ctx->dirty_off, (uint64_t)ctx->tr_cap, (uint64_t)(uintptr_t)ctx->tr_x};
}

static int drop_graphs(cfpq_ctx *ctx) {
    for (auto &g : ctx->graph_exec)
        if (g) {
            HIPCHK(hipGraphExecDestroy(g));
            g = nullptr;
        }
    return CFPQ_OK;
}

// Enqueue one step through the graph cache.
constexpr int GRAPH_VARIANTS = 4;
static int launch_step_graph(cfpq_ctx *ctx, const Plan &plan, uint32_t slot, int prev_slot, bool full_pass,
                             bool flush) {
    const int variant = prev_slot < 0 ? 0 : full_pass ? 2 : flush ? 3 : 1;
    hipGraphExec_t &ex = ctx->graph_exec[(size_t)variant * RING + slot];
    if (!ex) {
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        // the first step, full passes and flush steps zero their own counters:
        // they may follow steps whose slots the host reuses
        int rc = enqueue_step(ctx, plan, (int)(slot & 1), slot, prev_slot, full_pass, -1, flush, true,
                              prev_slot < 0 || full_pass || flush);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != CFPQ_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(ctx, CFPQ_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            ex = nullptr;
            return fail(ctx, CFPQ_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
    }
    HIPCHK(hipGraphLaunch(ex, ctx->stream));
    return CFPQ_OK;
}

// Sweeps (CFPQ_SCHEDULE_SWEEP, CFPQ_SCHEDULE_RULES): each step evaluates one
// head entry of the plan -- a head's rules, or a single rule in a rule plan --
// over the full matrices; a sweep visits the entries in order, the first sweep
// all of them, later sweeps those that read a matrix some entry changed in the
// sweep before.  The first sweep that adds no bit ends the run.  Every entry
// is then stable: one last evaluated in sweep j saw no operand change after
// its step (else sweep j + 1 would have revisited it), and none since.  With
// `certify`, one more sweep over every entry must add no bit (counted as the
// exact full pass; were it to add bits, the sweeps would go on).
static int run_sweeps(cfpq_ctx *ctx, const Plan &plan, bool heavy_paths, bool certify, StepStats &tot,
                      double *part_ms, double &product_ms, uint32_t &steps, uint32_t &syncs, uint32_t &launches,
                      int &prev_slot, uint32_t &full_passes) {
    const size_t nh = plan.heads.size();
    std::vector<std::vector<uint32_t>> readers(plan.mats.size());  // entries whose rules read matrix m
    for (size_t h = 0; h < nh; ++h)
        for (uint32_t t = plan.heads[h].f0; t < plan.heads[h].f1; ++t)
            for (uint32_t m : {plan.terms_full[t].left[0], plan.terms_full[t].right[0]})
                if (readers[m].empty() || readers[m].back() != (uint32_t)h) readers[m].push_back((uint32_t)h);
    std::vector<char> visit(nh, 1), changed(nh, 0);
    std::vector<uint32_t> slots, hs;
    bool certifying = false;
    // stamps from earlier calls would only make tiles look changed
    const uint32_t nall = (uint32_t)plan.mats.size();
    hipLaunchKernelGGL(k_zero_stamps, dim3(std::max<uint32_t>(1, (uint32_t)((plan.max_tiles_per_head + 255) / 256)),
                                           std::min<uint32_t>(std::max<uint32_t>(nall, 1), 65535)),
                       dim3(256), 0, ctx->stream, (const MatDev *)ctx->d_mats, ctx->d_ids + plan.ids_all_off, nall);
    CHK(launch_check(ctx, "k_zero_stamps"));
    for (;;) {
        std::fill(changed.begin(), changed.end(), 0);
        bool any = false;
        for (size_t h = 0; h <= nh; ++h) {
            if (h < nh && visit[h]) {
                const uint32_t slot = steps % RING;
                // steps are numbered from 1 for the stamps; the certifying
                // sweep evaluates every tile-product
                CHK(enqueue_step(ctx, plan, (int)(steps & 1), slot, prev_slot, true,
                                 ctx->opts.no_kernel_timing ? -1 : (int)slots.size(), false, true,
                                 prev_slot < 0, (int)h, steps + 1, !certifying));
                ++launches;
                slots.push_back(slot);
                hs.push_back((uint32_t)h);
                prev_slot = (int)slot;
                ++steps;
            }
            // read the counters when the ring is about to wrap and at the end of the sweep
            if (slots.empty() || (h < nh && slots.size() < RING - 1)) continue;
            HIPCHK(hipMemcpyAsync(ctx->ring_host, ctx->ring, sizeof(StepCounters) * RING, hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ++syncs;
            for (size_t s = 0; s < slots.size(); ++s) {
                StepCounters &c = ctx->ring_host[slots[s]];
                fold_shards(c);
                float step_ms[4] = {0, 0, 0, 0};
                for (int k2 = 0; k2 < (heavy_paths ? 4 : 2) && !ctx->opts.no_kernel_timing; ++k2) {
                    HIPCHK(hipEventElapsedTime(&step_ms[k2], ctx->ev[EV_SLOT * s + k2], ctx->ev[EV_SLOT * s + k2 + 1]));
                    part_ms[k2] += step_ms[k2];
                    product_ms += step_ms[k2];
                }
                accumulate(tot, c);
                if (c.new_bits) {
                    changed[hs[s]] = 1;
                    any = true;
                }
            }
            slots.clear();
            hs.clear();
            if (ctx->opts.max_iterations && steps >= ctx->opts.max_iterations)
                return fail(ctx, CFPQ_E_LIMIT, "fixpoint: iteration cap %u reached", ctx->opts.max_iterations);
        }
        if (certifying) ++full_passes;
        if (!any) {
            if (!certify || certifying) return CFPQ_OK;
            certifying = true;
            std::fill(visit.begin(), visit.end(), 1);
            continue;
        }
        certifying = false;
        std::fill(visit.begin(), visit.end(), 0);
        for (size_t x = 0; x < nh; ++x)
            if (changed[x])
                for (uint32_t r : readers[plan.heads[x].target]) visit[r] = 1;
    }
}

// Run delta steps to the fixpoint (plus the optional exact full pass).
// The caller owns plan buffers; on return all delta buffers are zero again
// (on success) -- on failure they are released instead of pooled.
static int run_fixpoint(cfpq_ctx *ctx, Plan &plan, cfpq_stats *stats, bool single_step_full) {
    if (ctx->opts.schedule == CFPQ_SCHEDULE_RULES && !single_step_full && !plan.rule_heads) {
        Plan rp = rule_plan(plan);
        return run_fixpoint(ctx, rp, stats, single_step_full);
    }
    CHK(upload_plan(ctx, plan));
    const uint32_t nd = (uint32_t)plan.all_delta[0].size();
    // (the first step reads the full matrices, so the initial delta buffer
    // D0 stays empty: every initial bit is "new" through the full terms)
    const int per_sync = single_step_full ? 1 : ctx->opts.iters_per_sync;
    // graph replay needs the parity to follow the slot (RING is even)
    const bool use_graph = ctx->opts.use_graph != 0 && !single_step_full;
    // the dense and MFMA products (and their timing events) run only with a heavy-tile path on
    const bool heavy_paths = dense_threshold(ctx->opts) <= TILE_BITS || mfma_threshold(ctx->opts) <= TILE_BITS;
    if (use_graph) {
        std::vector<uint64_t> sig = launch_signature(ctx, plan);
        if (sig != ctx->graph_sig || ctx->graph_exec.size() != GRAPH_VARIANTS * (size_t)RING) {
            CHK(drop_graphs(ctx));
            ctx->graph_exec.assign(GRAPH_VARIANTS * (size_t)RING, nullptr);
            ctx->graph_sig = sig;
        }
    }
    CHK(ensure_events(ctx, EV_SLOT * (size_t)RING + 4));
    hipEvent_t ev_begin = ctx->ev[EV_SLOT * RING], ev_end = ctx->ev[EV_SLOT * RING + 1];
    HIPCHK(hipEventRecord(ev_begin, ctx->stream));
    StepStats tot;
    double product_ms = 0, part_ms[4] = {0, 0, 0, 0};  // sparse, merge, dense, MFMA
    uint32_t steps = 0, syncs = 0, launches = 0, full_passes = 0, flush_steps = 0;
    uint64_t sched_scored = 0, sched_deferred = 0, sched_full = 0;
    int64_t dirty_total = 0;  // tiles deferred and not yet evaluated in full
    static const bool step_log = std::getenv("CFPQ_STEP_LOG") != nullptr;  // per-step counters to stderr
    int prev_slot = -1;
    bool done = false;
    // next batch: 0 = delta steps, 1 = one exact full-pass step, 2 = one flush step
    int pending = single_step_full ? 1 : 0;
    const bool want_full = ctx->opts.final_full_pass != 0;
    if (ctx->opts.schedule != CFPQ_SCHEDULE_DELTA && !single_step_full) {
        CHK(run_sweeps(ctx, plan, heavy_paths, want_full, tot, part_ms, product_ms, steps, syncs, launches, prev_slot,
                       full_passes));
        done = true;
    }
    // A window: `per_sync` delta steps (or one full-pass or flush step), then
    // a copy of their counters to the host and an event.  Delta windows are
    // enqueued one ahead: while the host waits for a window's counters the
    // next one already runs, so the device does not idle for the host's wake-up
    // and enqueue at every window.  If the waited-for window converged, the
    // steps ahead found no delta (as the steps after the quiet one within a
    // window) and are not counted.  Not with deferred tiles, whose balance
    // every executed step changes, and not with graph replay (measured slower
    // with a window of graph launches ahead).
    struct Window {
        int kind = 0;
        uint32_t first = 0;  // step number of the first step
        std::vector<uint32_t> slots;
        hipEvent_t ev = nullptr;
    };
    Window win[2];
    win[0].ev = ctx->ev[EV_SLOT * RING + 2];
    win[1].ev = ctx->ev[EV_SLOT * RING + 3];
    const bool ahead_ok = !single_step_full && !use_graph && 2 * per_sync <= RING &&
                          !(plan.dirty_words && (sched_on(ctx) || order_on(ctx) || ctx->tr_cap));
    auto enqueue_window = [&](int kind, Window &w) -> int {
        const int k = kind ? 1 : per_sync;
        w.kind = kind;
        w.first = steps;
        w.slots.clear();
        for (int s = 0; s < k; ++s) {
            const uint32_t slot = steps % RING;
            // the first step evaluates every rule once over the full matrices
            // (with d = M the two semi-naive terms would repeat the same product)
            const bool full = kind == 1 || steps == 0;
            if (use_graph) CHK(launch_step_graph(ctx, plan, slot, prev_slot, full, kind == 2));
            else CHK(enqueue_step(ctx, plan, (int)(steps & 1), slot, prev_slot, full,
                                  ctx->opts.no_kernel_timing ? -1 : (int)slot, kind == 2, true,
                                  prev_slot < 0 || (uint32_t)(prev_slot + 1) % RING != slot));
            ++launches;
            w.slots.push_back(slot);
            prev_slot = (int)slot;
            ++steps;
        }
        // the counters of the window's steps (consecutive ring slots)
        const uint32_t first = w.slots[0], n1 = std::min<uint32_t>((uint32_t)k, RING - first);
        HIPCHK(hipMemcpyAsync(ctx->ring_host + first, ctx->ring + first, sizeof(StepCounters) * n1,
                              hipMemcpyDeviceToHost, ctx->stream));
        if (n1 < (uint32_t)k)
            HIPCHK(hipMemcpyAsync(ctx->ring_host, ctx->ring, sizeof(StepCounters) * (k - n1), hipMemcpyDeviceToHost,
                                  ctx->stream));
        HIPCHK(hipEventRecord(w.ev, ctx->stream));
        return CFPQ_OK;
    };
    int cw = 0;  // the window waited for next
    if (!done) CHK(enqueue_window(pending, win[cw]));
    while (!done) {
        Window &w = win[cw];
        Window &ahead = win[cw ^ 1];
        const bool have_ahead = ahead_ok && w.kind == 0;
        if (have_ahead) CHK(enqueue_window(0, ahead));
        HIPCHK(hipEventSynchronize(w.ev));
        ++syncs;
        const int kind = w.kind;
        const int k = (int)w.slots.size();
        const std::vector<uint32_t> &slots = w.slots;
        pending = 0;
        uint32_t used = 0;
        bool quiet = false;  // a step added no bit
        for (int s = 0; s < k; ++s) fold_shards(ctx->ring_host[slots[s]]);
        for (int s = 0; s < k; ++s) {
            const StepCounters &c = ctx->ring_host[slots[s]];
            // every executed step counts for the deferred-tile balance (a step
            // after a quiet one still retires dirty tiles whose head tile is full)
            dirty_total += (int64_t)c.dirty_new - (int64_t)c.dirty_cleared;
            if (quiet) continue;
            float step_ms[4] = {0, 0, 0, 0};
            const uint32_t e0 = EV_SLOT * slots[s];
            for (int k2 = 0; k2 < (heavy_paths ? 4 : 2) && !use_graph && !ctx->opts.no_kernel_timing; ++k2) {
                HIPCHK(hipEventElapsedTime(&step_ms[k2], ctx->ev[e0 + k2], ctx->ev[e0 + k2 + 1]));
                part_ms[k2] += step_ms[k2];
                product_ms += step_ms[k2];
            }
            if (step_log)
                std::fprintf(stderr,
                             "[cfpq step %u] units %u dense %u mfma %u touched %u new_bits %llu cand %llu hits %llu "
                             "max_unit %u split %u a_rows %llu b_rows %llu busy %.3f "
                             "wave_max %llu unit_max %llu clk (nsub %u cands %u pairs %u gathers %u; clk scan %u filter %u "
                             "pairs %u epilogue %u) | sparse %.3f "
                             "merge %.3f dense %.3f mfma %.3f ms\n",
                             w.first + s, c.n_units + c.n_heavy, c.n_dense, c.n_mfma, c.n_touched,
                             (unsigned long long)c.new_bits, (unsigned long long)c.tile_products,
                             (unsigned long long)c.cand_hits, c.max_unit_cands, c.n_split, (unsigned long long)c.a_rows,
                             (unsigned long long)c.b_rows,
                             c.sparse_max ? (double)c.sparse_busy / ((double)c.sparse_max * ctx->num_cu *
                                                                      ctx->opts.sparse_waves_per_cu) : 0.0,
                             (unsigned long long)c.sparse_max, (unsigned long long)c.sparse_unit_max,
                             (unsigned)(c.sparse_unit_max_a & 31u), (unsigned)((c.sparse_unit_max_a >> 5) & 511u),
                             (unsigned)((c.sparse_unit_max_a >> 14) & 1023u) * 16u,
                             (unsigned)(c.sparse_unit_max_b & 0xffffffu),
                             (unsigned)((c.sparse_unit_max_d >> 12) & 4095u) * 256u,
                             (unsigned)((c.sparse_unit_max_c >> 12) & 4095u) * 256u,
                             (unsigned)(c.sparse_unit_max_c & 4095u) * 256u, (unsigned)(c.sparse_unit_max_d & 4095u) * 256u,
                             step_ms[0], step_ms[1], step_ms[2], step_ms[3]);
            accumulate(tot, c);
            sched_scored += c.n_scored;
            sched_deferred += c.n_deferred;
            sched_full += c.n_full;
            ++used;
            if (c.new_bits == 0) quiet = true;
        }
        if (kind == 1) ++full_passes;
        if (kind == 2) ++flush_steps;
        // steps enqueued after a quiet one found no delta; count only up to it
        // (with the window ahead of this one)
        const uint32_t ahead_steps = have_ahead ? (uint32_t)ahead.slots.size() : 0u;
        steps -= (uint32_t)(k - used) + (quiet ? ahead_steps : 0u);
        if (quiet) {
            if (dirty_total > 0) {
                // deferred tiles remain: evaluate them all in full before
                // anything is declared converged
                pending = 2;
            } else if (single_step_full) {
                done = true;
            } else if (want_full && kind != 1) {
                pending = 1;  // the exact full pass certifies the fixpoint
            } else {
                done = true;
            }
        } else if (single_step_full && kind == 1) {
            done = true;
        }
        if (!done && ctx->opts.max_iterations && steps - (quiet ? 0u : ahead_steps) >= ctx->opts.max_iterations)
            return fail(ctx, CFPQ_E_LIMIT, "fixpoint: iteration cap %u reached", ctx->opts.max_iterations);
        if (done) break;
        if (have_ahead && !quiet) {
            cw ^= 1;  // the window ahead is waited for next
        } else {
            CHK(enqueue_window(pending, w));
        }
    }
    // the last step's delta (D of the step after it) is empty; the delta it
    // consumed was cleared by the step itself.  Retire the newest delta list
    // so that every buffer is zero for the pool.
    {
        const int q = (int)(steps & 1);  // parity the next step would have
        const uint32_t slot = steps % RING;
        const uint32_t *ids_q = ctx->d_ids + (q ? nd : 0);
        const uint32_t nids_q = (uint32_t)plan.all_delta[q].size();
        const uint32_t nw = (uint32_t)std::min<size_t>((size_t)ctx->num_cu * 16, std::max<size_t>(1, plan.total_head_tiles)),
                       per = occ_blocks_per_matrix(plan);
        const TouchedLists ptl{ctx->lists[prev_slot & 1], (uint32_t)plan.total_head_tiles};
        hipLaunchKernelGGL(k_retire, dim3(nw + nids_q * per + 1), dim3(WAVE), 0, ctx->stream,
                           (const MatDev *)ctx->d_mats, (const HeadDev *)ctx->d_heads, q,
                           (const StepCounters *)(ctx->ring + slot), ptl, (const StepCounters *)(ctx->ring + prev_slot),
                           ptl, ids_q, nids_q, per, nw, 0, 1, ctx->ring + slot, 1, 0u, -1, (uint32_t *)nullptr);
        CHK(launch_check(ctx, "final clear"));
    }
    HIPCHK(hipEventRecord(ev_end, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float total_ms = 0;
    HIPCHK(hipEventElapsedTime(&total_ms, ev_begin, ev_end));
    if (stats) {
        stats->iterations = steps;
        stats->full_passes = full_passes;
        stats->new_bits = tot.new_bits;
        stats->items = tot.items;
        stats->tile_products = tot.tp;
        stats->tile_products_dense = tot.tpd;
        stats->tile_products_mfma = tot.tpm;
        stats->tile_products_unchanged = tot.tpu;
        stats->a_rows = tot.arows;
        stats->b_rows = tot.brows;
        stats->out_rows = tot.orows;
        stats->b_rows_lds = tot.brows_lds;
        stats->m4r_lookups = tot.lookups;
        stats->fixpoint_ms = total_ms;
        stats->product_ms = product_ms;
        stats->product_launches = launches;
        stats->host_syncs = syncs;
        stats->mfma_instr = tot.mfma_instr;
        stats->sparse_ms = part_ms[0];
        stats->merge_ms = part_ms[1];
        stats->dense_ms = part_ms[2];
        stats->mfma_ms = part_ms[3];
        stats->sched_scored = sched_scored;
        stats->sched_deferred = sched_deferred;
        stats->sched_full = sched_full;
        stats->flush_steps = flush_steps;
        stats->cand_hits = tot.cand_hits;
        stats->a_rows_dense = tot.arows_d;
        stats->a_cols = tot.acols;
        stats->units_done = tot.units_done;
        stats->units_idle = tot.units_idle;
        stats->unit_clocks = tot.unit_clk;
        stats->unit_clocks_idle = tot.idle_clk;
        stats->b_rows_idle = tot.idle_brows;
        stats->b_rows_dense = tot.brows_d;
        stats->out_rows_dense = tot.orows_d;
        stats->max_unit_cands = tot.max_unit_cands;
    }
    return CFPQ_OK;
}

// Build the plan for one query: matrix ids [base, base+nm) = M, then D0, D1.
static void plan_query(Plan &plan, cfpq_mat *const *M, const std::vector<cfpq_mat *> &D0,
                       const std::vector<cfpq_mat *> &D1, size_t nm, const cfpq_rule *rules, size_t nrules) {
    const uint32_t base = (uint32_t)plan.mats.size();
    for (size_t x = 0; x < nm; ++x) plan.mats.push_back(M[x]->d);
    for (size_t x = 0; x < nm; ++x) plan.mats.push_back(D0[x]->d);
    for (size_t x = 0; x < nm; ++x) plan.mats.push_back(D1[x]->d);
    auto id_m = [&](uint32_t x) { return base + x; };
    auto id_d = [&](int q, uint32_t x) { return base + (uint32_t)nm * (1 + q) + x; };
    for (size_t x = 0; x < nm; ++x) {
        plan.all_delta[0].push_back(id_d(0, (uint32_t)x));
        plan.all_delta[1].push_back(id_d(1, (uint32_t)x));
    }
    const uint32_t T = M[0]->T;
    // group rules by head (stable)
    std::vector<std::vector<size_t>> by_head(nm);
    for (size_t r = 0; r < nrules; ++r) by_head[rules[r].head].push_back(r);
    for (size_t h = 0; h < nm; ++h) {
        if (by_head[h].empty()) continue;
        HeadDev hd;
        hd.target = id_m((uint32_t)h);
        hd.out[0] = id_d(1, (uint32_t)h);  // even step writes D1
        hd.out[1] = id_d(0, (uint32_t)h);  // odd step writes D0
        hd.p0 = (uint32_t)plan.terms_delta.size();
        hd.t0 = (uint32_t)plan.terms_live.size();  // rebased past the delta and full terms in upload_plan
        hd.f0 = (uint32_t)plan.terms_full.size();  // rebased past the delta terms in upload_plan
        hd.T = T;
        for (size_t r : by_head[h]) {
            const uint32_t a = rules[r].left, b = rules[r].right;
            TermDev t1, t2, f;
            t1.left[0] = id_d(0, a); t1.left[1] = id_d(1, a);
            t1.right[0] = id_m(b);   t1.right[1] = id_m(b);
            t2.left[0] = id_m(a);    t2.left[1] = id_m(a);
            t2.right[0] = id_d(0, b); t2.right[1] = id_d(1, b);
            t1.rule = t2.rule = (uint32_t)r;
            plan.terms_delta.push_back(t1);
            plan.terms_delta.push_back(t2);
            if (!by_head[a].empty()) plan.terms_live.push_back(t1);
            if (!by_head[b].empty()) plan.terms_live.push_back(t2);
            f.left[0] = f.left[1] = id_m(a);
            f.right[0] = f.right[1] = id_m(b);
            f.rule = (uint32_t)r;
            plan.terms_full.push_back(f);
        }
        hd.p1 = (uint32_t)plan.terms_delta.size();
        hd.t1 = (uint32_t)plan.terms_live.size();
        hd.f1 = (uint32_t)plan.terms_full.size();
        hd.colT = 0;
        for (size_t r = 0; r < nrules; ++r) hd.colT |= rules[r].left == h ? 1u : 0u;
        plan.heads.push_back(hd);
        plan.max_tiles_per_head = std::max<size_t>(plan.max_tiles_per_head, (size_t)T * T);
        plan.total_head_tiles += (size_t)T * T;
    }
}

static int check_query(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, cfpq_mat *const *mats, size_t nmats) {
    if (!mats || nmats == 0 || (nrules && !rules)) return fail(ctx, CFPQ_E_INVALID, "fixpoint: empty matrix list");
    for (size_t x = 0; x < nmats; ++x) {
        if (!mats[x]) return fail(ctx, CFPQ_E_INVALID, "fixpoint: matrix %zu is null", x);
        if (mats[x]->n != mats[0]->n) return fail(ctx, CFPQ_E_INVALID, "fixpoint: matrices of different sizes");
        for (size_t y = 0; y < x; ++y)
            if (mats[y] == mats[x]) return fail(ctx, CFPQ_E_INVALID, "fixpoint: matrix %zu repeated", x);
    }
    for (size_t r = 0; r < nrules; ++r)
        if (rules[r].head >= nmats || rules[r].left >= nmats || rules[r].right >= nmats)
            return fail(ctx, CFPQ_E_INVALID, "fixpoint: rule %zu names a matrix >= %zu", r, nmats);
    return CFPQ_OK;
}

extern "C" int cfpq_batch_fixpoint(cfpq_ctx *ctx, const cfpq_query *queries, size_t nqueries, cfpq_stats *stats) {
    if (!ctx || (nqueries && !queries)) return fail(ctx, CFPQ_E_INVALID, "batch_fixpoint: bad argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (nqueries == 0) return CFPQ_OK;
    for (size_t qi = 0; qi < nqueries; ++qi)
        CHK(check_query(ctx, queries[qi].rules, queries[qi].nrules, queries[qi].mats, queries[qi].nmats));
    if (nqueries > 1) {
        // a matrix that a query writes (a rule head) must belong to that query
        // alone: another query reading it would never see its new bits in its
        // own delta buffers.  Read-only operands may be shared.
        std::unordered_map<const cfpq_mat *, size_t> writer;
        for (size_t qi = 0; qi < nqueries; ++qi)
            for (size_t r = 0; r < queries[qi].nrules; ++r) writer.emplace(queries[qi].mats[queries[qi].rules[r].head], qi);
        for (size_t qi = 0; qi < nqueries; ++qi)
            for (size_t x = 0; x < queries[qi].nmats; ++x) {
                auto it = writer.find(queries[qi].mats[x]);
                if (it != writer.end() && it->second != qi)
                    return fail(ctx, CFPQ_E_INVALID,
                                "batch_fixpoint: matrix %zu of query %zu is a rule head of query %zu (a matrix a "
                                "query writes cannot be shared)", x, qi, it->second);
            }
    }
    HIPCHK(hipSetDevice(ctx->device));
    // column-major forms: exact for every matrix read as a left operand; the
    // heads no rule reads as one are not kept up to date by this call
    for (size_t qi = 0; qi < nqueries; ++qi)
        for (size_t r = 0; r < queries[qi].nrules; ++r) CHK(colT_fresh(ctx, queries[qi].mats[queries[qi].rules[r].left]));
    for (size_t qi = 0; qi < nqueries; ++qi)
        for (size_t r = 0; r < queries[qi].nrules; ++r) {
            const cfpq_query &q = queries[qi];
            bool left = false;
            for (size_t r2 = 0; r2 < q.nrules; ++r2) left |= q.rules[r2].left == q.rules[r].head;
            if (!left) q.mats[q.rules[r].head]->colT_stale = true;
        }
    Plan plan;
    std::vector<cfpq_mat *> taken;
    int rc = CFPQ_OK;
    for (size_t qi = 0; qi < nqueries && rc == CFPQ_OK; ++qi) {
        const cfpq_query &q = queries[qi];
        std::vector<cfpq_mat *> D0(q.nmats), D1(q.nmats);
        for (size_t x = 0; x < q.nmats && rc == CFPQ_OK; ++x) {
            rc = pool_take(ctx, q.mats[0]->n, &D0[x]);
            if (rc == CFPQ_OK) taken.push_back(D0[x]);
            if (rc == CFPQ_OK) rc = pool_take(ctx, q.mats[0]->n, &D1[x]);
            if (rc == CFPQ_OK) taken.push_back(D1[x]);
        }
        if (rc == CFPQ_OK) plan_query(plan, q.mats, D0, D1, q.nmats, q.rules, q.nrules);
    }
    if (rc == CFPQ_OK && plan.heads.size() >= MAX_HEADS)
        rc = fail(ctx, CFPQ_E_INVALID, "batch_fixpoint: %zu rule heads in one batch (limit %u); split the batch",
                  plan.heads.size(), MAX_HEADS - 1);
    if (rc == CFPQ_OK && plan.heads.empty()) {
        for (auto *m : taken) pool_give(ctx, m);
        if (stats) stats->iterations = 0;
        return CFPQ_OK;
    }
    if (rc == CFPQ_OK) rc = run_fixpoint(ctx, plan, stats, false);
    if (rc == CFPQ_OK) {
        for (auto *m : taken) pool_give(ctx, m);
    } else {
        (void)hipStreamSynchronize(ctx->stream);
        for (auto *m : taken) free_mat(m);
    }
    return rc;
}

extern "C" int cfpq_fixpoint(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, cfpq_mat *const *mats, size_t nmats,
                             cfpq_stats *stats) {
    if (!ctx) return CFPQ_E_INVALID;
    cfpq_query q;
    q.mats = mats;
    q.nmats = nmats;
    q.rules = rules;
    q.nrules = nrules;
    return cfpq_batch_fixpoint(ctx, &q, 1, stats);
}

// The plan of one add_mul: a single full step  S = (a*b) & ~c ;  c |= S ;
// S cleared again.  Descriptor ids: 0 = c, 1 = a, 2 = b, 3 = S.
static void add_mul_plan(Plan &plan, const cfpq_mat *c, const cfpq_mat *a, const cfpq_mat *b, const MatDev &S) {
    plan.mats = {c->d, a->d, b->d, S};
    HeadDev hd;
    hd.target = 0;
    hd.out[0] = hd.out[1] = 3;
    hd.t0 = 0;  // rebased to 2 by upload_plan
    hd.t1 = 1;
    hd.f0 = 0;  // rebased to 1 by upload_plan
    hd.f1 = 1;
    hd.p0 = 0;
    hd.p1 = 1;
    hd.T = c->T;
    hd.colT = 1;
    plan.heads.push_back(hd);
    TermDev t;
    t.left[0] = t.left[1] = 1;
    t.right[0] = t.right[1] = 2;
    t.rule = 0;
    plan.terms_delta.push_back(t);
    plan.terms_full.push_back(t);
    plan.terms_live.push_back(t);
    plan.all_delta[0] = {3};
    plan.all_delta[1] = {3};
    plan.max_tiles_per_head = plan.total_head_tiles = (size_t)c->T * c->T;
}

// Upload the plan, run its step and read the changed flag back.  On success
// S is all-zero again (the step's first-step clear).
static int add_mul_run(cfpq_ctx *ctx, Plan &plan, int *changed) {
    int rc = upload_plan(ctx, plan);
    if (rc == CFPQ_OK) rc = enqueue_step(ctx, plan, 0, 0, -1, false, -1, false, /*fuse=*/false, true);
    // enqueue_step cleared nothing useful for S yet (prev_slot < 0 clears all
    // non-empty tiles of "D_0" = S *after* apply), which is exactly the reset we need.
    if (rc == CFPQ_OK) {
        hipError_t e = hipMemcpyAsync(ctx->ring_host, ctx->ring, sizeof(StepCounters), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, CFPQ_E_DEVICE, "add_mul: %s", hipGetErrorString(e));
    }
    if (rc == CFPQ_OK) {
        fold_shards(ctx->ring_host[0]);
        if (changed) *changed = ctx->ring_host[0].new_bits != 0;
    }
    return rc;
}

extern "C" int cfpq_add_mul(cfpq_ctx *ctx, cfpq_mat *c, const cfpq_mat *a, const cfpq_mat *b, int *changed) {
    if (!ctx || !c || !a || !b) return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul: bad argument");
    if (a->n != c->n || b->n != c->n) return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul: size mismatch");
    HIPCHK(hipSetDevice(ctx->device));
    CHK(colT_fresh(ctx, a));
    CHK(colT_fresh(ctx, c));  // the head's form is kept up to date from here
    cfpq_mat *S = nullptr;
    CHK(pool_take(ctx, c->n, &S));
    Plan plan;
    add_mul_plan(plan, c, a, b, S->d);
    const int rc = add_mul_run(ctx, plan, changed);
    if (rc == CFPQ_OK) {
        pool_give(ctx, S);
    } else {
        (void)hipStreamSynchronize(ctx->stream);
        free_mat(S);
    }
    return rc;
}

// While alive, the context's work buffers are the regions of this device's
// scratch array (capacities set so that nothing grows) and every allocation
// path fails; the destructor puts the context's own buffers back.
struct NoallocBuffers {
    cfpq_ctx *ctx;
    MatDev *d_mats; HeadDev *d_heads; TermDev *d_terms; TermPtr *d_tptr; uint32_t *d_ids;
    uint64_t *items, *units, *mitems, *lists[2];
    StepCounters *ring;
    uint32_t *dirty, *dirty_off, *last_eval;
    uint64_t *partials, *pmask, *splits;
    uint32_t split_slots;
    size_t caps[15];
    explicit NoallocBuffers(cfpq_ctx *c) : ctx(c) {
        d_mats = c->d_mats; d_heads = c->d_heads; d_terms = c->d_terms; d_tptr = c->d_tptr; d_ids = c->d_ids;
        items = c->items; units = c->units; mitems = c->mitems; lists[0] = c->lists[0]; lists[1] = c->lists[1];
        ring = c->ring; dirty = c->dirty; dirty_off = c->dirty_off; last_eval = c->last_eval;
        partials = c->partials; pmask = c->pmask; splits = c->splits; split_slots = c->split_slots;
        const size_t cv[15] = {c->d_mats_cap, c->d_heads_cap, c->d_terms_cap, c->d_tptr_cap, c->d_ids_cap, c->items_cap,
                               c->units_cap, c->mitems_cap, c->lists_cap, c->dirty_cap, c->dirty_off_cap,
                               c->partials_cap, c->pmask_cap, c->splits_cap, c->last_eval_cap};
        std::copy(cv, cv + 15, caps);
        char *base = c->noalloc_base;
        c->d_mats = (MatDev *)(base + NA_MATS);       c->d_mats_cap = NA_MATS_CAP;
        c->d_heads = (HeadDev *)(base + NA_HEADS);    c->d_heads_cap = NA_HEADS_CAP;
        c->d_terms = (TermDev *)(base + NA_TERMS);    c->d_terms_cap = NA_TERMS_CAP;
        c->d_tptr = (TermPtr *)(base + NA_TPTR);      c->d_tptr_cap = NA_TPTR_CAP;
        c->d_ids = (uint32_t *)(base + NA_IDS);       c->d_ids_cap = NA_IDS_CAP;
        c->items = (uint64_t *)(base + NA_ITEMS);     c->items_cap = NOALLOC_TILES;
        c->mitems = (uint64_t *)(base + NA_MITEMS);   c->mitems_cap = NOALLOC_TILES;
        c->units = (uint64_t *)(base + NA_UNITS);     c->units_cap = NA_UNITS_CAP;
        c->lists[0] = (uint64_t *)(base + NA_LIST0);
        c->lists[1] = (uint64_t *)(base + NA_LIST1);  c->lists_cap = NA_LIST_CAP;
        c->ring = (StepCounters *)(base + NA_RING);
        c->dirty = (uint32_t *)(base + NA_DIRTY);     c->dirty_cap = NA_DIRTY_CAP;
        c->dirty_off = (uint32_t *)(base + NA_DIRTY_OFF); c->dirty_off_cap = 1;
        c->last_eval = (uint32_t *)(base + NA_DIRTY_OFF + 128); c->last_eval_cap = 1;  // (in the same 256 B)
        c->partials = (uint64_t *)(base + NA_PARTIALS); c->partials_cap = (size_t)NA_SPLIT_SLOTS * WAVE * TROW_WORDS;
        c->pmask = (uint64_t *)(base + NA_PMASK);     c->pmask_cap = NA_SPLIT_SLOTS;
        c->splits = (uint64_t *)(base + NA_SPLITS);   c->splits_cap = 2 * (NA_SPLIT_SLOTS / 2 + 1);
        c->split_slots = NA_SPLIT_SLOTS;
        c->no_alloc = true;
    }
    ~NoallocBuffers() {
        cfpq_ctx *c = ctx;
        c->d_mats = d_mats; c->d_heads = d_heads; c->d_terms = d_terms; c->d_tptr = d_tptr; c->d_ids = d_ids;
        c->items = items; c->units = units; c->mitems = mitems; c->lists[0] = lists[0]; c->lists[1] = lists[1];
        c->ring = ring; c->dirty = dirty; c->dirty_off = dirty_off; c->last_eval = last_eval;
        c->partials = partials; c->pmask = pmask; c->splits = splits; c->split_slots = split_slots;
        c->partials_cap = caps[11]; c->pmask_cap = caps[12]; c->splits_cap = caps[13]; c->last_eval_cap = caps[14];
        c->d_mats_cap = caps[0]; c->d_heads_cap = caps[1]; c->d_terms_cap = caps[2]; c->d_tptr_cap = caps[3];
        c->d_ids_cap = caps[4]; c->items_cap = caps[5]; c->units_cap = caps[6]; c->mitems_cap = caps[7];
        c->lists_cap = caps[8]; c->dirty_cap = caps[9]; c->dirty_off_cap = caps[10];
        c->no_alloc = false;
    }
};

extern "C" int cfpq_add_mul_noalloc(cfpq_ctx *ctx, cfpq_mat *c, const cfpq_mat *a, const cfpq_mat *b, int *changed) {
    if (!ctx || !c || !a || !b) return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul_noalloc: bad argument");
    if (a->n != c->n || b->n != c->n) return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul_noalloc: size mismatch");
    if (c->n > CFPQ_NOALLOC_MAX_N)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul_noalloc: n = %u is above the limit of %u of its scratch array "
                    "(use cfpq_add_mul)", c->n, (unsigned)CFPQ_NOALLOC_MAX_N);
    if (!ctx->noalloc_base) {
        const bool dev_ok = ctx->device >= 0 && ctx->device < NOALLOC_DEVICES;
        return fail(ctx, CFPQ_E_DEVICE, "cfpq_add_mul_noalloc: no scratch array on device %d: %s", ctx->device,
                    dev_ok ? g_noalloc_err[ctx->device].c_str() : "device index out of range");
    }
    std::lock_guard<std::mutex> lock(g_noalloc_call_mu[ctx->device]);
    HIPCHK(hipSetDevice(ctx->device));
    CHK(colT_fresh(ctx, a));
    CHK(colT_fresh(ctx, c));
    // the delta matrix S at the start of the array, cleared here: the array
    // keeps whatever the last call on this device left in it
    MatDev S = {};
    const size_t sbytes = mat_layout(c->T, ctx->noalloc_base, &S);
    if (sbytes > NA_RING) return fail(ctx, CFPQ_E_INVALID, "cfpq_add_mul_noalloc: delta matrix exceeds its region");
    HIPCHK(hipMemsetAsync(ctx->noalloc_base, 0, sbytes, ctx->stream));
    Plan plan;
    add_mul_plan(plan, c, a, b, S);
    int rc;
    {
        NoallocBuffers guard(ctx);
        rc = add_mul_run(ctx, plan, changed);
        if (rc != CFPQ_OK) (void)hipStreamSynchronize(ctx->stream);
    }
    return rc;
}

// ---------------------------------------------------------------------------
// learned tile scheduler: weights, evaluation, training, traces
// ---------------------------------------------------------------------------

static int ensure_sched_weights(cfpq_ctx *ctx) {
    if (!ctx->sched_w) HIPCHK(hipMalloc((void **)&ctx->sched_w, MLP_PARAMS * sizeof(uint16_t)));
    return CFPQ_OK;
}

extern "C" int cfpq_sched_set_weights(cfpq_ctx *ctx, const uint16_t *w, size_t count) {
    if (!ctx || !w || count != (size_t)MLP_PARAMS)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_set_weights: need %d bf16 values", MLP_PARAMS);
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure_sched_weights(ctx));
    HIPCHK(hipMemcpyAsync(ctx->sched_w, w, MLP_PARAMS * 2, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->sched_have = true;
    // the master copy used by training restarts from these weights
    if (ctx->t_master) {
        std::vector<float> f(MLP_PARAMS);
        for (int i = 0; i < MLP_PARAMS; ++i) f[i] = bf16_to_f32(w[i]);
        HIPCHK(hipMemcpy(ctx->t_master, f.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(ctx->t_m1, 0, MLP_PARAMS * 4));
        HIPCHK(hipMemset(ctx->t_m2, 0, MLP_PARAMS * 4));
    }
    return CFPQ_OK;
}

extern "C" int cfpq_sched_get_weights(cfpq_ctx *ctx, uint16_t *w, size_t count) {
    if (!ctx || !w || count != (size_t)MLP_PARAMS)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_get_weights: need room for %d bf16 values", MLP_PARAMS);
    if (!ctx->sched_have) return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_get_weights: no weights set");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(w, ctx->sched_w, MLP_PARAMS * 2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

extern "C" int cfpq_sched_init(cfpq_ctx *ctx, uint64_t seed) {
    if (!ctx) return CFPQ_E_INVALID;
    // uniform(-a, a) with a = sqrt(6 / fan_in) for the weight matrices (He
    // uniform), small uniform embeddings, zero biases; SplitMix64 stream
    uint64_t st = seed;
    auto next = [&st]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    auto unif = [&](float a) { return (float)(((next() >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a; };
    std::vector<uint16_t> w(MLP_PARAMS, 0);
    for (int i = 0; i < MLP_RULES * MLP_E; ++i) w[MLP_OFF_E + i] = f32_to_bf16(unif(0.5f));
    for (int i = 0; i < MLP_IN * MLP_H; ++i) w[MLP_OFF_W1 + i] = f32_to_bf16(unif(std::sqrt(6.f / MLP_IN)));
    for (int i = 0; i < MLP_H * MLP_H; ++i) w[MLP_OFF_W2 + i] = f32_to_bf16(unif(std::sqrt(6.f / MLP_H)));
    for (int i = 0; i < MLP_H; ++i) w[MLP_OFF_W3 + i] = f32_to_bf16(unif(std::sqrt(3.f / MLP_H)));
    return cfpq_sched_set_weights(ctx, w.data(), w.size());
}

static int upload_samples(cfpq_ctx *ctx, const float *x, const uint32_t *rule, const float *y, size_t n) {
    CHK(grow(ctx, &ctx->t_x, &ctx->t_x_cap, n * MLP_F));
    CHK(grow(ctx, &ctx->t_rule, &ctx->t_rule_cap, n));
    CHK(grow(ctx, &ctx->t_y, &ctx->t_y_cap, n));
    HIPCHK(hipMemcpyAsync(ctx->t_x, x, n * MLP_F * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->t_rule, rule, n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->t_y, y, n * 4, hipMemcpyHostToDevice, ctx->stream));
    return CFPQ_OK;
}

static int ensure_train_state(cfpq_ctx *ctx) {
    if (ctx->t_master) return CFPQ_OK;
    for (float **p : {&ctx->t_master, &ctx->t_m1, &ctx->t_m2, &ctx->t_grad})
        HIPCHK(hipMalloc((void **)p, MLP_PARAMS * sizeof(float)));
    HIPCHK(hipMalloc((void **)&ctx->t_loss, 64 * sizeof(float)));
    HIPCHK(hipMemset(ctx->t_m1, 0, MLP_PARAMS * 4));
    HIPCHK(hipMemset(ctx->t_m2, 0, MLP_PARAMS * 4));
    if (ctx->sched_have) {
        std::vector<uint16_t> w(MLP_PARAMS);
        HIPCHK(hipMemcpy(w.data(), ctx->sched_w, MLP_PARAMS * 2, hipMemcpyDeviceToHost));
        std::vector<float> f(MLP_PARAMS);
        for (int i = 0; i < MLP_PARAMS; ++i) f[i] = bf16_to_f32(w[i]);
        HIPCHK(hipMemcpy(ctx->t_master, f.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice));
    }
    return CFPQ_OK;
}

static uint32_t mlp_grid(cfpq_ctx *ctx, size_t n) {
    return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + MLP_BLOCK - 1) / MLP_BLOCK, (size_t)ctx->num_cu * 8));
}

extern "C" int cfpq_sched_eval(cfpq_ctx *ctx, const uint16_t *w, const float *x, const uint32_t *rule, const float *y,
                               size_t n, float *logits, float *loss, float *grad) {
    if (!ctx || !w || (n && (!x || !rule || !y)) || n > 0xffffffffu)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_eval: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ensure_train_state(ctx));
    uint16_t *dw = nullptr;
    float *dlog = nullptr;
    HIPCHK(hipMalloc((void **)&dw, MLP_PARAMS * 2));
    int rc = CFPQ_OK;
    auto run = [&]() -> int {
        HIPCHK(hipMemcpyAsync(dw, w, MLP_PARAMS * 2, hipMemcpyHostToDevice, ctx->stream));
        if (n) CHK(upload_samples(ctx, x, rule, y, n));
        if (logits && n) HIPCHK(hipMalloc((void **)&dlog, n * 4));
        hipLaunchKernelGGL(k_fill_f32, dim3(24), dim3(256), 0, ctx->stream, ctx->t_grad, (uint32_t)MLP_PARAMS, 0.f);
        hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, ctx->t_loss, 1u, 0.f);
        if (n)
            hipLaunchKernelGGL(k_mlp_eval, dim3(mlp_grid(ctx, n)), dim3(MLP_BLOCK), 0, ctx->stream,
                               (const uint16_t *)dw, (const float *)ctx->t_x, (const uint32_t *)ctx->t_rule,
                               (const float *)ctx->t_y, (uint32_t)n, 1.f / (float)n, 1.f / (float)n, dlog,
                               ctx->t_loss, grad ? ctx->t_grad : nullptr);
        CHK(launch_check(ctx, "k_mlp_eval"));
        if (logits && n) HIPCHK(hipMemcpyAsync(logits, dlog, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        if (loss) HIPCHK(hipMemcpyAsync(loss, ctx->t_loss, 4, hipMemcpyDeviceToHost, ctx->stream));
        if (grad) HIPCHK(hipMemcpyAsync(grad, ctx->t_grad, MLP_PARAMS * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return CFPQ_OK;
    };
    rc = run();
    (void)hipStreamSynchronize(ctx->stream);
    if (dlog) (void)hipFree(dlog);
    (void)hipFree(dw);
    return rc;
}

extern "C" int cfpq_sched_train(cfpq_ctx *ctx, const float *x, const uint32_t *rule, const float *y, size_t n,
                                uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss) {
    if (!ctx || !x || !rule || !y || n == 0 || n > 0xffffffffu || epochs == 0 || epochs > 64)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_train: bad argument (n > 0, 1 <= epochs <= 64)");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->sched_have) CHK(cfpq_sched_init(ctx, seed));
    CHK(ensure_train_state(ctx));
    if (batch == 0) batch = 4096;
    // one fixed shuffle of the samples (host, from the seed); batches are
    // contiguous slices of it
    std::vector<uint32_t> perm(n);
    for (size_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    uint64_t st = seed ^ 0x5851f42d4c957f2dull;
    for (size_t i = n - 1; i > 0; --i) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        std::swap(perm[i], perm[(st >> 33) % (i + 1)]);
    }
    std::vector<float> sx(n * MLP_F), sy(n);
    std::vector<uint32_t> sr(n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(&sx[i * MLP_F], x + (size_t)perm[i] * MLP_F, MLP_F * 4);
        sr[i] = rule[perm[i]];
        sy[i] = y[perm[i]];
    }
    CHK(upload_samples(ctx, sx.data(), sr.data(), sy.data(), n));
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, ctx->stream, ctx->t_loss, 64u, 0.f);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    uint64_t t = 0;
    const size_t nb = (n + batch - 1) / batch;
    for (uint32_t e = 0; e < epochs; ++e) {
        for (size_t bi = 0; bi < nb; ++bi) {
            const size_t b0 = bi * batch, bn = std::min<size_t>(batch, n - b0);
            ++t;
            hipLaunchKernelGGL(k_fill_f32, dim3(24), dim3(256), 0, ctx->stream, ctx->t_grad, (uint32_t)MLP_PARAMS, 0.f);
            // gradient of the batch mean; the epoch's mean loss accumulates in t_loss[e]
            hipLaunchKernelGGL(k_mlp_eval, dim3(mlp_grid(ctx, bn)), dim3(MLP_BLOCK), 0, ctx->stream,
                               (const uint16_t *)ctx->sched_w, (const float *)(ctx->t_x + b0 * MLP_F),
                               (const uint32_t *)(ctx->t_rule + b0), (const float *)(ctx->t_y + b0), (uint32_t)bn,
                               1.f / (float)bn, 1.f / (float)n, (float *)nullptr, ctx->t_loss + e, ctx->t_grad);
            const float c1 = 1.f - (float)std::pow((double)b1, (double)t);
            const float c2 = 1.f - (float)std::pow((double)b2, (double)t);
            hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, ctx->stream, ctx->t_master,
                               ctx->sched_w, ctx->t_m1, ctx->t_m2, (const float *)ctx->t_grad, lr, b1, b2, c1, c2,
                               eps);
            CHK(launch_check(ctx, "training step"));
        }
    }
    if (loss) HIPCHK(hipMemcpyAsync(loss, ctx->t_loss, epochs * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CFPQ_OK;
}

extern "C" int cfpq_sched_trace_begin(cfpq_ctx *ctx, size_t capacity) {
    if (!ctx || capacity == 0 || capacity > 0x7fffffffu) return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_trace_begin: bad capacity");
    HIPCHK(hipSetDevice(ctx->device));
    CHK(cfpq_sched_trace_end(ctx));
    HIPCHK(hipMalloc((void **)&ctx->tr_x, capacity * MLP_F * 4));
    HIPCHK(hipMalloc((void **)&ctx->tr_y, capacity * 4));
    HIPCHK(hipMalloc((void **)&ctx->tr_rule, capacity * 4));
    HIPCHK(hipMalloc((void **)&ctx->tr_item, 2 * capacity * 4));
    HIPCHK(hipMalloc((void **)&ctx->tr_count, 4));
    HIPCHK(hipMemsetAsync(ctx->tr_count, 0, 4, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->tr_cap = capacity;
    return CFPQ_OK;
}

extern "C" int cfpq_sched_trace_read(cfpq_ctx *ctx, float *x, uint32_t *rule, float *y, size_t capacity, size_t *count) {
    if (!ctx || !count || (capacity && (!x || !rule || !y))) return fail(ctx, CFPQ_E_INVALID, "cfpq_sched_trace_read: bad argument");
    *count = 0;
    if (!ctx->tr_cap) return CFPQ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    uint32_t cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, ctx->tr_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t n = std::min<size_t>(std::min<size_t>(cnt, ctx->tr_cap), capacity);
    if (n) {
        HIPCHK(hipMemcpyAsync(x, ctx->tr_x, n * MLP_F * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(rule, ctx->tr_rule, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(y, ctx->tr_y, n * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    *count = std::min<size_t>(cnt, ctx->tr_cap);
    return CFPQ_OK;
}

extern "C" int cfpq_sched_trace_end(cfpq_ctx *ctx) {
    if (!ctx) return CFPQ_E_INVALID;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (void *p : {(void *)ctx->tr_x, (void *)ctx->tr_y, (void *)ctx->tr_rule, (void *)ctx->tr_item, (void *)ctx->tr_count})
        if (p) HIPCHK(hipFree(p));
    ctx->tr_x = ctx->tr_y = nullptr;
    ctx->tr_rule = ctx->tr_item = ctx->tr_count = nullptr;
    ctx->tr_cap = 0;
    return CFPQ_OK;
}

// ---------------------------------------------------------------------------
// single-matrix formulation (one matrix of nonterminal sets)
// ---------------------------------------------------------------------------

extern "C" int cfpq_nt_closure(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, uint32_t *elems, uint32_t n,
                               cfpq_stats *stats) {
    if (!ctx || (n && !elems) || (nrules && !rules)) return fail(ctx, CFPQ_E_INVALID, "cfpq_nt_closure: bad argument");
    if (nrules > (size_t)NT_MAX_RULES)
        return fail(ctx, CFPQ_E_INVALID, "cfpq_nt_closure: %zu rules (at most %d)", nrules, NT_MAX_RULES);
    std::vector<uint32_t> packed(nrules);
    for (size_t r = 0; r < nrules; ++r) {
        if (rules[r].head >= 32 || rules[r].left >= 32 || rules[r].right >= 32)
            return fail(ctx, CFPQ_E_INVALID, "cfpq_nt_closure: rule %zu names a nonterminal >= 32", r);
        packed[r] = rules[r].head | rules[r].left << 8 | rules[r].right << 16;
    }
    if (stats) *stats = cfpq_stats{};
    if (n == 0) return CFPQ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t total = (size_t)n * n;
    uint32_t *buf = nullptr, *d_rules = nullptr;
    if (hipMalloc((void **)&buf, 3 * total * sizeof(uint32_t)) != hipSuccess)
        return fail(ctx, CFPQ_E_NOMEM, "cfpq_nt_closure: %zu bytes of device memory", 3 * total * sizeof(uint32_t));
    int rc = CFPQ_OK;
    uint32_t *T = buf, *D = buf + total, *N = buf + 2 * total;
    uint32_t steps = 0;
    uint64_t new_elems = 0;
    float ms = 0;
    auto run = [&]() -> int {
        HIPCHK(hipMalloc((void **)&d_rules, std::max<size_t>(1, nrules) * sizeof(uint32_t)));
        if (nrules)
            HIPCHK(hipMemcpyAsync(d_rules, packed.data(), nrules * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(T, elems, total * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        CHK(ensure_events(ctx, EV_SLOT * (size_t)RING + 2));
        HIPCHK(hipEventRecord(ctx->ev[EV_SLOT * RING], ctx->stream));
        const uint32_t nb = (n + NT_BLOCK - 1) / NT_BLOCK;
        const uint32_t ga = (uint32_t)std::min<size_t>((total + 255) / 256, (size_t)ctx->num_cu * 16);
        for (;;) {
            HIPCHK(hipMemsetAsync(ctx->scratch64, 0, 8, ctx->stream));
            hipLaunchKernelGGL(k_nt_product, dim3(nb, nb), dim3(NT_THREADS), 0, ctx->stream, (const uint32_t *)T,
                               (const uint32_t *)D, N, n, NtRules{d_rules, (uint32_t)nrules}, steps == 0 ? 1 : 0,
                               ctx->scratch64);
            CHK(launch_check(ctx, "nt product"));
            HIPCHK(hipMemcpyAsync(ctx->scratch64_host, ctx->scratch64, 8, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            ++steps;
            const uint64_t fresh = ctx->scratch64_host[0];
            if (!fresh) break;
            new_elems += fresh;
            hipLaunchKernelGGL(k_nt_apply, dim3(ga), dim3(256), 0, ctx->stream, T, (const uint32_t *)N, total);
            CHK(launch_check(ctx, "nt apply"));
            std::swap(D, N);  // this step's new elements are the next step's delta
            if (ctx->opts.max_iterations && steps >= ctx->opts.max_iterations)
                return fail(ctx, CFPQ_E_LIMIT, "cfpq_nt_closure: iteration cap %u reached", ctx->opts.max_iterations);
        }
        HIPCHK(hipEventRecord(ctx->ev[EV_SLOT * RING + 1], ctx->stream));
        HIPCHK(hipMemcpyAsync(elems, T, total * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[EV_SLOT * RING], ctx->ev[EV_SLOT * RING + 1]));
        return CFPQ_OK;
    };
    rc = run();
    (void)hipStreamSynchronize(ctx->stream);
    if (d_rules) (void)hipFree(d_rules);
    (void)hipFree(buf);
    if (rc == CFPQ_OK && stats) {
        stats->iterations = steps;
        stats->items = new_elems;
        stats->fixpoint_ms = ms;
        stats->product_launches = steps;
        stats->host_syncs = steps;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// sparse pair-list formulation (pairs.hpp)
// ---------------------------------------------------------------------------

namespace {
// A device array of uint64 that only grows; its contents are not kept across
// a growth (callers grow destinations only).
struct PairBuf {
    uint64_t *p = nullptr;
    size_t cap = 0;
};
}  // namespace

// Hands b back to the context's pool.  Work already enqueued on the stream
// that reads or writes it runs before any later use, so no wait is needed.
static void pairs_release(cfpq_ctx *ctx, PairBuf &b) {
    if (b.p) ctx->pair_pool.emplace_back(b.p, b.cap);
    b = PairBuf{};
}

static int pairs_grow(cfpq_ctx *ctx, PairBuf &b, size_t n) {
    if (n <= b.cap) return CFPQ_OK;
    pairs_release(ctx, b);
    // the smallest pooled array that fits
    size_t best = ctx->pair_pool.size();
    for (size_t i = 0; i < ctx->pair_pool.size(); ++i)
        if (ctx->pair_pool[i].second >= n && (best == ctx->pair_pool.size() || ctx->pair_pool[i].second < ctx->pair_pool[best].second))
            best = i;
    if (best < ctx->pair_pool.size()) {
        b.p = ctx->pair_pool[best].first;
        b.cap = ctx->pair_pool[best].second;
        ctx->pair_pool.erase(ctx->pair_pool.begin() + best);
        return CFPQ_OK;
    }
    const size_t c = std::max<size_t>(n + n / 2, 1024);
    if (hipMalloc((void **)&b.p, c * sizeof(uint64_t)) != hipSuccess) {
        (void)hipGetLastError();
        // give the pool back to the device and try once more
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (auto &e : ctx->pair_pool) (void)hipFree(e.first);
        ctx->pair_pool.clear();
        b.p = nullptr;
        if (hipMalloc((void **)&b.p, n * sizeof(uint64_t)) != hipSuccess) {
            (void)hipGetLastError();
            b.p = nullptr;
            return fail(ctx, CFPQ_E_NOMEM, "cfpq_pairs_fixpoint: %zu bytes of device memory", n * sizeof(uint64_t));
        }
        b.cap = n;
        return CFPQ_OK;
    }
    b.cap = c;
    return CFPQ_OK;
}

extern "C" int cfpq_pairs_fixpoint(cfpq_ctx *ctx, const cfpq_rule *rules, size_t nrules, cfpq_pairs *sets,
                                   uint32_t nsets, uint32_t n, cfpq_stats *stats) {
    if (!ctx || (nsets && !sets) || (nrules && !rules))
        return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: bad argument");
    if (n > (1u << 31)) return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: n = %u (at most 2^31)", n);
    for (size_t r = 0; r < nrules; ++r)
        if (rules[r].head >= nsets || rules[r].left >= nsets || rules[r].right >= nsets)
            return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: rule %zu names a set >= %u", r, nsets);
    for (uint32_t x = 0; x < nsets; ++x) {
        if (sets[x].count && !sets[x].keys) return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: set %u has no keys", x);
        if (sets[x].count > sets[x].cap)
            return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: set %u holds %llu keys over its cap %llu", x,
                        (unsigned long long)sets[x].count, (unsigned long long)sets[x].cap);
        for (uint64_t e = 0; e < sets[x].count; ++e) {
            const uint64_t k = sets[x].keys[e];
            if ((k >> 32) >= n || (k & 0xffffffffull) >= n)
                return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: set %u key %llu = (%llu, %llu) outside n = %u", x,
                            (unsigned long long)e, (unsigned long long)(k >> 32),
                            (unsigned long long)(k & 0xffffffffull), n);
        }
    }
    if (stats) *stats = cfpq_stats{};
    if (n == 0 || nsets == 0) return CFPQ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    uint32_t s = 1;
    while ((1ull << s) < n) ++s;

    struct Rel {
        PairBuf m[2], mt[2], d, dn;  // by row (current, spare), by column, delta, next delta
        uint64_t count = 0, nd = 0, ndn = 0;
    };
    std::vector<Rel> rel(nsets);
    PairBuf cand[2], cnt, start, incl, flags, tmp, words, sel, dparts;
    std::vector<uint32_t> heads;
    for (size_t r = 0; r < nrules; ++r)
        if (std::find(heads.begin(), heads.end(), rules[r].head) == heads.end()) heads.push_back(rules[r].head);
    uint32_t steps = 0, launches = 0, syncs = 0;
    uint64_t added = 0, joined = 0;
    float ms = 0;
    const hipStream_t st = ctx->stream;
    auto grid = [&](uint64_t m) {
        return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((m + PAIRS_THREADS - 1) / PAIRS_THREADS,
                                                                       (uint64_t)ctx->num_cu * 16)));
    };
    auto host_wait = [&]() -> int {
        HIPCHK(hipStreamSynchronize(st));
        ++syncs;
        return CFPQ_OK;
    };
    const uint32_t key_bits = 2 * s;
    // heads per shared sort: their tags must fit above the key bits
    const size_t group = std::min<size_t>(PAIRS_GROUP, (size_t)1 << std::min<uint32_t>(64 - key_bits, 5));
    auto tag_bits = [](size_t g) {
        uint32_t b = 0;
        while (((size_t)1 << b) < g) ++b;
        return b;
    };
    // Sorts cand[0][off, off + m) on bits [0, end_bit); cur := the buffer
    // index holding the result (at the same offset).
    auto sort_cand = [&](uint64_t m, uint32_t end_bit, int &cur, uint64_t off = 0) -> int {
        hipcub::DoubleBuffer<uint64_t> db(cand[0].p + off, cand[1].p + off);
        size_t bytes = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, db, m, 0, (int)end_bit, st));
        CHK(pairs_grow(ctx, tmp, (bytes + 7) / 8));
        HIPCHK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, db, m, 0, (int)end_bit, st));
        cur = db.Current() == cand[0].p + off ? 0 : 1;
        return CFPQ_OK;
    };
    // The keys of c[0, m) whose flags are set, into out (cap >= m); their
    // number into the device word *kept, which the host reads after its next wait.
    auto select_flagged = [&](const uint64_t *c, uint64_t m, uint64_t *out, uint64_t *kept) -> int {
        size_t bytes = 0;
        HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, c, (const uint8_t *)flags.p, out, kept, (int64_t)m, st));
        CHK(pairs_grow(ctx, tmp, (bytes + 7) / 8));
        HIPCHK(hipcub::DeviceSelect::Flagged(tmp.p, bytes, c, (const uint8_t *)flags.p, out, kept, (int64_t)m, st));
        return CFPQ_OK;
    };
    // host_words := k device words (one wait).
    std::vector<uint64_t> host_words;
    auto read_words = [&](const uint64_t *d, size_t k) -> int {
        host_words.assign(k, 0);
        if (k) HIPCHK(hipMemcpyAsync(host_words.data(), d, k * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        return host_wait();
    };
    // m[1] (and mt[1]) := merge of the current array with the sorted, disjoint b; swap.
    auto merge_into = [&](PairBuf (&two)[2], uint64_t na, const uint64_t *b, uint64_t nb) -> int {
        CHK(pairs_grow(ctx, two[1], na + nb));
        hipLaunchKernelGGL(k_pairs_merge, grid(na + nb), dim3(PAIRS_THREADS), 0, st, (const uint64_t *)two[0].p, na, b,
                           nb, two[1].p);
        CHK(launch_check(ctx, "pairs merge"));
        std::swap(two[0], two[1]);
        return CFPQ_OK;
    };
    // dst := the column-ordered copy of m row-ordered keys (through cand).
    auto build_transpose = [&](const uint64_t *keys, uint64_t m, PairBuf &dst) -> int {
        CHK(pairs_grow(ctx, cand[0], m));
        CHK(pairs_grow(ctx, cand[1], m));
        CHK(pairs_grow(ctx, dst, m));
        hipLaunchKernelGGL(k_pairs_transpose, grid(m), dim3(PAIRS_THREADS), 0, st, keys, cand[0].p, m, s, 0ull);
        CHK(launch_check(ctx, "pairs transpose"));
        int cur = 0;
        CHK(sort_cand(m, key_bits, cur));
        HIPCHK(hipMemcpyAsync(dst.p, cand[cur].p, m * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        return CFPQ_OK;
    };

    auto run = [&]() -> int {
        CHK(ensure_events(ctx, EV_SLOT * (size_t)RING + 2));
        HIPCHK(hipEventRecord(ctx->ev[EV_SLOT * RING], st));
        // load: pack, sort, drop duplicates (one wait for all sets), build the column order
        CHK(pairs_grow(ctx, words, std::max<size_t>(nsets, group + 2)));
        for (uint32_t x = 0; x < nsets; ++x) {
            Rel &r = rel[x];
            const uint64_t m = sets[x].count;
            if (!m) continue;
            CHK(pairs_grow(ctx, cand[0], m));
            CHK(pairs_grow(ctx, cand[1], m));
            HIPCHK(hipMemcpyAsync(cand[0].p, sets[x].keys, m * sizeof(uint64_t), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_pairs_pack, grid(m), dim3(PAIRS_THREADS), 0, st, cand[0].p, m, s);
            CHK(launch_check(ctx, "pairs pack"));
            int cur = 0;
            CHK(sort_cand(m, key_bits, cur));
            CHK(pairs_grow(ctx, r.m[0], m));
            CHK(pairs_grow(ctx, flags, (m + 7) / 8));
            hipLaunchKernelGGL(k_pairs_flag_new, grid(m), dim3(PAIRS_THREADS), 0, st, (const uint64_t *)cand[cur].p, m,
                               (const uint64_t *)nullptr, 0ull, (uint8_t *)flags.p);
            CHK(launch_check(ctx, "pairs flag"));
            CHK(select_flagged(cand[cur].p, m, r.m[0].p, words.p + x));
        }
        CHK(read_words(words.p, nsets));
        for (uint32_t x = 0; x < nsets; ++x) {
            rel[x].count = sets[x].count ? host_words[x] : 0;
            if (rel[x].count) CHK(build_transpose(rel[x].m[0].p, rel[x].count, rel[x].mt[0]));
        }
        std::vector<PairPart> parts;
        std::vector<size_t> first(heads.size() + 1);
        for (;;) {
            bool fresh = false;
            // every head's joins, in head order, each output tagged with its
            // head's index in its group
            parts.clear();
            uint64_t sum_m = 0, max_m = 0;
            for (size_t hi = 0; hi < heads.size(); ++hi) {
                first[hi] = parts.size();
                const uint64_t tag = (uint64_t)(hi % group) << key_bits;
                for (size_t i = 0; i < nrules; ++i) {
                    if (rules[i].head != heads[hi]) continue;
                    Rel &L = rel[rules[i].left], &R = rel[rules[i].right];
                    // step 0: every pair is delta, and dL |><| M_R = M_L |><| dR
                    const uint64_t *dl = steps ? L.d.p : L.m[0].p, *dr = steps ? R.d.p : R.m[0].p;
                    const uint64_t ndl = steps ? L.nd : L.count, ndr = steps ? R.nd : R.count;
                    if (ndl && R.count) {
                        parts.push_back({dl, R.m[0].p, ndl, R.count, sum_m, tag, 0, 0});
                        sum_m += ndl;
                        max_m = std::max(max_m, ndl);
                    }
                    if (steps && ndr && L.count) {
                        parts.push_back({dr, L.mt[0].p, ndr, L.count, sum_m, tag, 1, 0});
                        sum_m += ndr;
                        max_m = std::max(max_m, ndr);
                    }
                }
            }
            first[heads.size()] = parts.size();
            for (uint32_t h : heads) rel[h].ndn = 0;
            if (!parts.empty()) {
                const uint32_t np = (uint32_t)parts.size();
                if (np > 65535) return fail(ctx, CFPQ_E_INVALID, "cfpq_pairs_fixpoint: %u joins in one step", np);
                CHK(pairs_grow(ctx, cnt, sum_m));
                CHK(pairs_grow(ctx, start, sum_m));
                CHK(pairs_grow(ctx, incl, sum_m));
                CHK(pairs_grow(ctx, words, std::max<size_t>(np, group + 2)));
                CHK(pairs_grow(ctx, dparts, (np * sizeof(PairPart) + 7) / 8));
                // `parts` is left alone until the wait below, so the copy may read it late
                HIPCHK(hipMemcpyAsync(dparts.p, parts.data(), np * sizeof(PairPart), hipMemcpyHostToDevice, st));
                const PairPart *dp = (const PairPart *)dparts.p;
                hipLaunchKernelGGL(k_pairs_count_parts, dim3(grid(max_m).x, np), dim3(PAIRS_THREADS), 0, st, dp, s, cnt.p,
                                   start.p);
                CHK(launch_check(ctx, "pairs count"));
                size_t bytes = 0;
                HIPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt.p, incl.p, sum_m, st));
                CHK(pairs_grow(ctx, tmp, (bytes + 7) / 8));
                HIPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, cnt.p, incl.p, sum_m, st));
                hipLaunchKernelGGL(k_pairs_part_totals, grid(np), dim3(PAIRS_THREADS), 0, st, dp, np,
                                   (const uint64_t *)incl.p, words.p);
                CHK(launch_check(ctx, "pairs totals"));
                CHK(read_words(words.p, np));
                const std::vector<uint64_t> totals = host_words;
                uint64_t total_all = 0;
                for (uint64_t t : totals) total_all += t;
                if (total_all) {
                    joined += total_all;
                    CHK(pairs_grow(ctx, cand[0], total_all));
                    CHK(pairs_grow(ctx, cand[1], total_all));
                    hipLaunchKernelGGL(k_pairs_expand_parts, grid(total_all), dim3(PAIRS_THREADS), 0, st, dp, np,
                                       (const uint64_t *)incl.p, (const uint64_t *)start.p, sum_m, s, total_all,
                                       cand[0].p);
                    CHK(launch_check(ctx, "pairs expand"));
                    ++launches;
                }
                // per group of heads: its outputs are contiguous; one sort, what
                // is new kept, split by head (one wait)
                uint64_t off = 0;
                for (size_t g0 = 0; g0 < heads.size(); g0 += group) {
                    const size_t g1 = std::min(heads.size(), g0 + group);
                    uint64_t total = 0;
                    for (size_t i = first[g0]; i < first[g1]; ++i) total += totals[i];
                    if (!total) continue;
                    PairHeads tab{};
                    for (size_t hi = g0; hi < g1; ++hi) {
                        tab.m[hi - g0] = rel[heads[hi]].m[0].p;
                        tab.n[hi - g0] = rel[heads[hi]].count;
                    }
                    int cur = 0;
                    CHK(sort_cand(total, key_bits + tag_bits(g1 - g0), cur, off));
                    const uint64_t *sorted = cand[cur].p + off;
                    off += total;
                    CHK(pairs_grow(ctx, flags, (total + 7) / 8));
                    hipLaunchKernelGGL(k_pairs_flag_new_group, grid(total), dim3(PAIRS_THREADS), 0, st, sorted, total,
                                       tab, key_bits, (uint8_t *)flags.p);
                    CHK(launch_check(ctx, "pairs flag"));
                    CHK(pairs_grow(ctx, sel, total));
                    CHK(select_flagged(sorted, total, sel.p, words.p));
                    hipLaunchKernelGGL(k_pairs_group_bounds, dim3(1), dim3(64), 0, st, (const uint64_t *)sel.p,
                                       (const uint64_t *)words.p, key_bits, (uint32_t)(g1 - g0), words.p + 1);
                    CHK(launch_check(ctx, "pairs bounds"));
                    CHK(read_words(words.p, g1 - g0 + 2));
                    for (size_t hi = g0; hi < g1; ++hi) {
                        const uint64_t b = host_words[1 + hi - g0], e = host_words[2 + hi - g0];
                        Rel &H = rel[heads[hi]];
                        H.ndn = e - b;
                        if (!H.ndn) continue;
                        CHK(pairs_grow(ctx, H.dn, H.ndn));
                        hipLaunchKernelGGL(k_pairs_untag, grid(H.ndn), dim3(PAIRS_THREADS), 0, st,
                                           (const uint64_t *)sel.p + b, H.dn.p, H.ndn, key_bits);
                        CHK(launch_check(ctx, "pairs untag"));
                        fresh = true;
                    }
                }
            }
            ++steps;
            // every head's new pairs were found against the previous step's
            // relations; only now are they merged in, the column copies
            // through one tagged sort per group
            for (size_t g0 = 0; g0 < heads.size(); g0 += group) {
                const size_t g1 = std::min(heads.size(), g0 + group);
                uint64_t total = 0;
                for (size_t hi = g0; hi < g1; ++hi) total += rel[heads[hi]].ndn;
                if (!total) continue;
                CHK(pairs_grow(ctx, cand[0], total));
                CHK(pairs_grow(ctx, cand[1], total));
                uint64_t off = 0;
                for (size_t hi = g0; hi < g1; ++hi) {
                    Rel &H = rel[heads[hi]];
                    if (!H.ndn) continue;
                    added += H.ndn;
                    CHK(merge_into(H.m, H.count, H.dn.p, H.ndn));
                    hipLaunchKernelGGL(k_pairs_transpose, grid(H.ndn), dim3(PAIRS_THREADS), 0, st,
                                       (const uint64_t *)H.dn.p, cand[0].p + off, H.ndn, s,
                                       (uint64_t)(hi - g0) << key_bits);
                    CHK(launch_check(ctx, "pairs transpose"));
                    off += H.ndn;
                }
                int cur = 0;
                CHK(sort_cand(total, key_bits + tag_bits(g1 - g0), cur));
                off = 0;
                for (size_t hi = g0; hi < g1; ++hi) {
                    Rel &H = rel[heads[hi]];
                    if (!H.ndn) continue;
                    CHK(pairs_grow(ctx, H.d, H.ndn));  // H.d is free until the swap below
                    hipLaunchKernelGGL(k_pairs_untag, grid(H.ndn), dim3(PAIRS_THREADS), 0, st,
                                       (const uint64_t *)cand[cur].p + off, H.d.p, H.ndn, key_bits);
                    CHK(launch_check(ctx, "pairs untag"));
                    CHK(merge_into(H.mt, H.count, H.d.p, H.ndn));
                    H.count += H.ndn;
                    off += H.ndn;
                }
            }
            for (Rel &r : rel) {
                std::swap(r.d, r.dn);
                r.nd = r.ndn;
                r.ndn = 0;
            }
            if (!fresh) break;
            if (ctx->opts.max_iterations && steps >= ctx->opts.max_iterations)
                return fail(ctx, CFPQ_E_LIMIT, "cfpq_pairs_fixpoint: iteration cap %u reached", ctx->opts.max_iterations);
        }
        HIPCHK(hipEventRecord(ctx->ev[EV_SLOT * RING + 1], st));
        CHK(host_wait());
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[EV_SLOT * RING], ctx->ev[EV_SLOT * RING + 1]));
        bool fits = true;
        for (uint32_t x = 0; x < nsets; ++x) fits = fits && rel[x].count <= sets[x].cap;
        if (!fits) {
            for (uint32_t x = 0; x < nsets; ++x) sets[x].count = rel[x].count;
            return fail(ctx, CFPQ_E_CAPACITY, "cfpq_pairs_fixpoint: a closed relation exceeds its set's cap");
        }
        for (uint32_t x = 0; x < nsets; ++x) {
            const uint64_t m = rel[x].count;
            if (m) {
                CHK(pairs_grow(ctx, cand[0], m));
                hipLaunchKernelGGL(k_pairs_unpack, grid(m), dim3(PAIRS_THREADS), 0, st, (const uint64_t *)rel[x].m[0].p,
                                   cand[0].p, m, s);
                CHK(launch_check(ctx, "pairs unpack"));
                HIPCHK(hipMemcpyAsync(sets[x].keys, cand[0].p, m * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
            }
            sets[x].count = m;
        }
        return CFPQ_OK;
    };
    const int rc = run();
    (void)hipStreamSynchronize(ctx->stream);
    auto drop = [&](PairBuf &b) { pairs_release(ctx, b); };
    for (Rel &r : rel) {
        drop(r.m[0]), drop(r.m[1]), drop(r.mt[0]), drop(r.mt[1]), drop(r.d), drop(r.dn);
    }
    drop(cand[0]), drop(cand[1]), drop(cnt), drop(start), drop(incl), drop(flags), drop(tmp), drop(words), drop(sel), drop(dparts);
    if (rc == CFPQ_OK && stats) {
        stats->iterations = steps;
        stats->new_bits = added;
        stats->items = joined;
        stats->fixpoint_ms = ms;
        stats->product_launches = launches;
        stats->host_syncs = syncs;
    }
    return rc;
}
