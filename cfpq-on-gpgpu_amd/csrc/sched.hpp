// Learned tile scheduler: a small tile-activity network that decides, per
// fixpoint step, which (head, C tile) items to evaluate now and which to
// defer.  It never decides correctness:
//   * a deferred item's tile is marked dirty; a dirty tile is later
//     evaluated over the *full* operands (M_L * M_R for every rule of its
//     head), which covers every delta it skipped;
//   * the host declares convergence only after a flush step -- every dirty
//     tile evaluated in full, nothing deferred -- adds no bit (optionally
//     followed by one exact full pass over all tiles).
//
// Network: 16 inputs (8 tile features + an 8-dim embedding of the rule id)
// -> 64 ReLU -> 64 ReLU -> 1 logit of "this (rule, tile) item produces new
// bits in this step".  Weights are stored as bf16; all arithmetic and
// accumulation is fp32.  Forward, backward and Adam kernels are below; the
// features are computed on the device from the occupancy bitsets.
#pragma once
#include "kernels.hpp"

namespace cfpq {

constexpr int MLP_F = 8;       // tile features
constexpr int MLP_E = 8;       // rule embedding
constexpr int MLP_IN = MLP_F + MLP_E;
constexpr int MLP_H = 64;
constexpr int MLP_RULES = 64;  // embedding rows; rule ids are taken mod 64
// parameter layout (one flat array)
constexpr int MLP_OFF_E = 0;
constexpr int MLP_OFF_W1 = MLP_OFF_E + MLP_RULES * MLP_E;  // [MLP_IN][MLP_H]
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;             // [MLP_H][MLP_H]
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;             // [MLP_H]
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_PARAMS = MLP_OFF_B3 + 1;                 // 5825

// bf16 <-> fp32 (round to nearest even; inputs are finite)
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Load the bf16 parameters into an fp32 LDS copy (whole work-group).
__device__ __forceinline__ void mlp_load(const uint16_t *__restrict__ w, float *sw) {
    for (int i = threadIdx.x; i < MLP_PARAMS; i += blockDim.x) sw[i] = bf16_to_f32(w[i]);
    __syncthreads();
}

// Forward pass of one sample; keeps the activations for the backward pass.
__device__ __forceinline__ float mlp_forward(const float *sw, const float *x, uint32_t rule, float *in, float *h1,
                                             float *h2) {
    const float *E = sw + MLP_OFF_E + (rule % MLP_RULES) * MLP_E;
#pragma unroll
    for (int i = 0; i < MLP_F; ++i) in[i] = x[i];
#pragma unroll
    for (int i = 0; i < MLP_E; ++i) in[MLP_F + i] = E[i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float a = sw[MLP_OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a += in[i] * sw[MLP_OFF_W1 + i * MLP_H + j];
        h1[j] = a > 0.f ? a : 0.f;
    }
    float z = sw[MLP_OFF_B3];
    for (int k = 0; k < MLP_H; ++k) {
        float a = sw[MLP_OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) a += h1[j] * sw[MLP_OFF_W2 + j * MLP_H + k];
        h2[k] = a > 0.f ? a : 0.f;
        z += h2[k] * sw[MLP_OFF_W3 + k];
    }
    return z;
}

// Logit only (no activations kept): the scheduler's inference path.
__device__ __forceinline__ float mlp_logit(const float *sw, const float *x, uint32_t rule) {
    const float *E = sw + MLP_OFF_E + (rule % MLP_RULES) * MLP_E;
    float in[MLP_IN], h1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_F; ++i) in[i] = x[i];
#pragma unroll
    for (int i = 0; i < MLP_E; ++i) in[MLP_F + i] = E[i];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) {
        float a = sw[MLP_OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a += in[i] * sw[MLP_OFF_W1 + i * MLP_H + j];
        h1[j] = a > 0.f ? a : 0.f;
    }
    float z = sw[MLP_OFF_B3];
    for (int k = 0; k < MLP_H; ++k) {
        float a = sw[MLP_OFF_B2 + k];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) a += h1[j] * sw[MLP_OFF_W2 + j * MLP_H + k];
        z += (a > 0.f ? a : 0.f) * sw[MLP_OFF_W3 + k];
    }
    return z;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Gradient accumulation: sum over the wave, one atomic per parameter.
__device__ __forceinline__ void grad_add(float *g, int idx, float v) {
    v = wave_sum_f(v);
    if ((threadIdx.x & 63) == 0 && v != 0.f) atomicAdd(g + idx, v);
}

// Binary cross-entropy with logits over n samples: logits, loss (sum of the
// per-sample losses times loss_scale) and (if grad != nullptr) the gradient
// of the mean loss (per-sample gradients times inv_n) w.r.t. every
// parameter, all fp32.  One thread per sample; grad and loss accumulate.
constexpr int MLP_BLOCK = 128;
__global__ __launch_bounds__(MLP_BLOCK) void k_mlp_eval(const uint16_t *__restrict__ w, const float *__restrict__ X,
                                                        const uint32_t *__restrict__ rules,
                                                        const float *__restrict__ Y, uint32_t n, float inv_n,
                                                        float loss_scale, float *__restrict__ logits,
                                                        float *__restrict__ loss,
                                                        float *__restrict__ grad) {
    __shared__ float sw[MLP_PARAMS];
    mlp_load(w, sw);
    for (uint32_t base = blockIdx.x * MLP_BLOCK; base < n; base += gridDim.x * MLP_BLOCK) {
        const uint32_t s = base + threadIdx.x;
        const bool live = s < n;
        float x[MLP_F], in[MLP_IN], h1[MLP_H], h2[MLP_H];
#pragma unroll
        for (int i = 0; i < MLP_F; ++i) x[i] = live ? X[(size_t)s * MLP_F + i] : 0.f;
        const uint32_t r = live ? rules[s] : 0u;
        const float z = mlp_forward(sw, x, r, in, h1, h2);
        const float y = live ? Y[s] : 0.f;
        if (live && logits) logits[s] = z;
        // loss = softplus(z) - y z, computed stably
        const float l = live ? (fmaxf(z, 0.f) - y * z + log1pf(expf(-fabsf(z)))) * loss_scale : 0.f;
        const float lw = wave_sum_f(l);
        if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, lw);
        if (!grad) continue;
        const float dz = live ? (1.f / (1.f + expf(-z)) - y) * inv_n : 0.f;
        grad_add(grad, MLP_OFF_B3, dz);
        float dh2[MLP_H];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) {
            grad_add(grad, MLP_OFF_W3 + k, dz * h2[k]);
            dh2[k] = h2[k] > 0.f ? dz * sw[MLP_OFF_W3 + k] : 0.f;
            grad_add(grad, MLP_OFF_B2 + k, dh2[k]);
        }
        float dh1[MLP_H];
        for (int j = 0; j < MLP_H; ++j) {
            float a = 0.f;
#pragma unroll
            for (int k = 0; k < MLP_H; ++k) {
                grad_add(grad, MLP_OFF_W2 + j * MLP_H + k, h1[j] * dh2[k]);
                a += sw[MLP_OFF_W2 + j * MLP_H + k] * dh2[k];
            }
            dh1[j] = h1[j] > 0.f ? a : 0.f;
            grad_add(grad, MLP_OFF_B1 + j, dh1[j]);
        }
        for (int i = 0; i < MLP_IN; ++i) {
            float a = 0.f;
#pragma unroll
            for (int j = 0; j < MLP_H; ++j) {
                grad_add(grad, MLP_OFF_W1 + i * MLP_H + j, in[i] * dh1[j]);
                a += sw[MLP_OFF_W1 + i * MLP_H + j] * dh1[j];
            }
            if (i >= MLP_F) {
                // embedding rows differ per lane: no wave reduction here
                if (live && a != 0.f) atomicAdd(grad + MLP_OFF_E + (r % MLP_RULES) * MLP_E + (i - MLP_F), a);
            }
        }
    }
}

// Adam step on the fp32 master weights; the bf16 copy is what the network uses.
__global__ void k_mlp_adam(float *__restrict__ master, uint16_t *__restrict__ w, float *__restrict__ m1,
                           float *__restrict__ m2, const float *__restrict__ grad, float lr, float b1, float b2,
                           float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_PARAMS) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float v = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float p = master[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    master[i] = p;
    w[i] = f32_to_bf16(p);
}

__global__ void k_fill_f32(float *__restrict__ p, uint32_t n, float v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// scheduled work-list construction
// ---------------------------------------------------------------------------

__device__ __forceinline__ float log2p1(uint32_t v) { return __log2f(1.f + (float)v); }

__device__ __forceinline__ uint32_t occ_popc(const uint32_t *a, uint32_t ow) {
    uint32_t s = 0;
    for (uint32_t w = 0; w < ow; ++w) s += __popc(a[w]);
    return s;
}

// Features of rule item (dL * M_R, M_L * dR) for C tile (I, J) of a head.
__device__ __forceinline__ void rule_features(const MatDev &dL, const MatDev &MR, const MatDev &ML,
                                              const MatDev &dR, uint32_t I, uint32_t J, float head_fill,
                                              float head_delta, float *x) {
    uint32_t cl = 0, cr = 0;
    for (uint32_t w = 0; w < dL.ow; ++w) {
        cl += __popc(dL.occR[I * dL.ow + w] & MR.occC[J * MR.ow + w]);
        cr += __popc(ML.occR[I * ML.ow + w] & dR.occC[J * dR.ow + w]);
    }
    const float s = 0.125f;
    x[0] = s * log2p1(cl);
    x[1] = s * log2p1(cr);
    x[2] = s * log2p1(occ_popc(dL.occR + I * dL.ow, dL.ow));
    x[3] = s * log2p1(occ_popc(dR.occC + J * dR.ow, dR.ow));
    x[4] = s * log2p1(occ_popc(ML.occR + I * ML.ow, ML.ow));
    x[5] = s * log2p1(occ_popc(MR.occC + J * MR.ow, MR.ow));
    x[6] = head_fill;
    x[7] = head_delta;
}

struct SchedArgs {
    const uint16_t *w;     // bf16 parameters
    uint32_t *dirty;       // per head: T*T bits (offset dirty_off[h] words)
    const uint32_t *dirty_off;
    float logit_threshold; // defer an item whose best logit is below this
    int defer;             // 1: may defer this step
    int flush;             // 1: evaluate every dirty tile in full, defer nothing
    int order;             // 1: rank items: predicted active (logit >= order_logit) go to the
                           //    first phase of the step
    float order_logit;
    float *trace_x;        // trace rows (nullptr: no trace)
    uint32_t *trace_rule;
    uint32_t *trace_item;  // (head << 40 | tile) per row, for the labels
    float *trace_y;
    uint32_t *trace_count;
    uint32_t trace_cap;
};

// One thread per (head, C tile): as k_build, plus the network's decision.
__global__ __launch_bounds__(BLOCK) void k_build_sched(const MatDev *__restrict__ mats,
                                                       const HeadDev *__restrict__ heads,
                                                       const TermDev *__restrict__ terms, int parity,
                                                       StepCounters *__restrict__ c, uint64_t *__restrict__ units,
                                                       uint64_t *__restrict__ dense_items,
                                                       uint64_t *__restrict__ mfma_items, SchedArgs sa, uint32_t split, uint32_t piece,
                                                       uint32_t split_budget, int heavy, uint32_t heavy_base,
                                                       int tile_units, uint64_t *__restrict__ splits,
                                                       uint64_t *__restrict__ prio_list, int split_rows) {
    __shared__ float sw[MLP_PARAMS];
    if (sa.w) mlp_load(sa.w, sw);
    const uint32_t h = blockIdx.y;
    const uint32_t T = uniform(heads[h].T), target = uniform(heads[h].target);
    const uint32_t t0 = uniform(heads[h].t0), t1 = uniform(heads[h].t1);
    const uint32_t f0 = uniform(heads[h].f0), f1 = uniform(heads[h].f1);
    const uint32_t p0 = uniform(heads[h].p0), p1 = uniform(heads[h].p1);
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    bool dense = false, mfma = false, full = false;
    uint32_t qmask = 0;
    uint32_t qcnt[QUARTERS] = {0, 0, 0, 0};
    uint32_t deferred = 0, fulls = 0, scored = 0, newly = 0, cleared = 0;
    bool prio = false;  // ranked into the step's first phase
    if (p < T * T) {
        const uint32_t I = p / T, J = p - I * T;
        uint32_t *dw = sa.dirty + sa.dirty_off[h] + (p >> 5);
        const uint32_t dbit = 1u << (p & 31);
        const bool dirty = (*dw & dbit) != 0;
        const MatDev M = mats[target];
        if (M.cnt[p] < TILE_BITS) {
            scan_terms(mats, terms, t0, t1, parity, T, I, J, qmask, dense, mfma, qcnt, heavy);
            bool evaluate = qmask != 0;
            if (evaluate && !sa.flush && (sa.w || sa.trace_x)) {
                // score every rule of the head; the item's score is the best one
                const MatDev Dh = mats[heads[h].out[1 - parity]];  // this head's delta consumed now
                const float fill = (float)M.cnt[p] * (1.f / (float)TILE_BITS);
                const float hd = log2p1(Dh.cnt[p]) * (1.f / 16.f);
                float best = -1e30f;
                for (uint32_t t = p0; t + 1 < p1; t += 2) {
                    const TermDev ta = terms[t], tb = terms[t + 1];
                    float x[MLP_F];
                    rule_features(mats[ta.left[parity]], mats[ta.right[parity]], mats[tb.left[parity]],
                                  mats[tb.right[parity]], I, J, fill, hd, x);
                    if (sa.w) {
                        best = fmaxf(best, mlp_logit(sw, x, ta.rule));
                        ++scored;
                    }
                    if (sa.trace_x) {
                        const uint32_t row = atomicAdd(sa.trace_count, 1u);
                        if (row < sa.trace_cap) {
                            for (int i = 0; i < MLP_F; ++i) sa.trace_x[(size_t)row * MLP_F + i] = x[i];
                            sa.trace_rule[row] = ta.rule;
                            sa.trace_item[row] = p;
                            sa.trace_item[sa.trace_cap + row] = h;
                            sa.trace_y[row] = -1.f;
                        }
                    }
                }
                prio = sa.w && sa.order && best >= sa.order_logit;
                if (sa.w && sa.defer && best < sa.logit_threshold) {
                    evaluate = false;
                    if (!dirty) {
                        atomicOr(dw, dbit);
                        ++newly;
                    }
                    ++deferred;
                }
            }
            if (dirty && (sa.flush || evaluate)) {
                // a dirty tile is evaluated over the full operands, which covers
                // every delta it skipped and this step's delta as well
                atomicAnd(dw, ~dbit);
                ++cleared;
                full = true;
                prio = false;
                qmask = 0;
                dense = mfma = false;
                qcnt[0] = qcnt[1] = qcnt[2] = qcnt[3] = 0;
                scan_terms(mats, terms, f0, f1, parity, T, I, J, qmask, dense, mfma, qcnt, heavy);
                evaluate = qmask != 0;
                ++fulls;
            }
            if (!evaluate) {
                qmask = 0;
                dense = mfma = false;
            }
        } else if (dirty) {
            atomicAnd(dw, ~dbit);  // a full head tile needs no more work
            ++cleared;
        }
    }
    const uint64_t flag = full ? ITEM_FULL : 0;
    wave_append(mfma, flag | ((uint64_t)h << 40) | p, &c->n_mfma, mfma_items);
    wave_append(dense && !mfma, flag | ((uint64_t)h << 40) | p, &c->n_dense, dense_items);
    dense = dense || mfma;
    append_units(!dense, qmask, qcnt, flag | ((uint64_t)h << 40) | ((uint64_t)p << 2), split, piece, split_budget, c, units,
                 heavy_base, tile_units, splits, prio, sa.order ? prio_list : nullptr, split_rows);
    deferred = wave_sum(deferred);
    fulls = wave_sum(fulls);
    scored = wave_sum(scored);
    newly = wave_sum(newly);
    cleared = wave_sum(cleared);
    if ((threadIdx.x & 63) == 0) {
        if (deferred) atomicAdd(&c->n_deferred, deferred);
        if (fulls) atomicAdd(&c->n_full, fulls);
        if (scored) atomicAdd(&c->n_scored, scored);
        if (newly) atomicAdd(&c->dirty_new, newly);
        if (cleared) atomicAdd(&c->dirty_cleared, cleared);
    }
}

// Labels of this step's trace rows: did the row's head tile gain a bit?
__global__ void k_trace_label(const MatDev *__restrict__ mats, const HeadDev *__restrict__ heads, int parity,
                              const StepCounters *__restrict__ c, SchedArgs sa) {
    const uint32_t end = *sa.trace_count < sa.trace_cap ? *sa.trace_count : sa.trace_cap;
    for (uint32_t row = c->trace0 + blockIdx.x * blockDim.x + threadIdx.x; row < end; row += gridDim.x * blockDim.x) {
        const uint32_t p = sa.trace_item[row], h = sa.trace_item[sa.trace_cap + row];
        sa.trace_y[row] = mats[heads[h].out[parity]].cnt[p] ? 1.f : 0.f;
    }
}

__global__ void k_trace_begin(StepCounters *__restrict__ c, const uint32_t *__restrict__ trace_count, uint32_t cap) {
    if (threadIdx.x == 0) c->trace0 = *trace_count < cap ? *trace_count : cap;
}

}  // namespace cfpq
