// HipMatrix / HipMatricesEnv: the driver's Matrix plugin over the C ABI.
#include "HipMatrix.h"

#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

namespace {

cfpq_ctx *g_ctx = nullptr;
std::once_flag g_once;

void check(int rc, const char *what) {
    if (rc != CFPQ_OK) {
        std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
        if (g_ctx) msg += ": " + std::string(cfpq_last_error(g_ctx));
        throw std::runtime_error(msg);
    }
}

HipMatrix *as_hip(Matrix *m) {
    auto *h = dynamic_cast<HipMatrix *>(m);
    if (!h) throw std::runtime_error("HipMatrix: operand is not a HipMatrix");
    return h;
}

}  // namespace

cfpq_ctx *cfpq_host_context() {
    std::call_once(g_once, [] {
        int dev = 0;
        if (const char *e = std::getenv("CFPQ_DEVICE")) dev = std::atoi(e);
        cfpq_opts o;
        cfpq_default_opts(&o);
        if (const char *k = std::getenv("CFPQ_KERNEL")) o.kernel = std::atoi(k);
        const int rc = cfpq_create(dev, &o, &g_ctx);
        if (rc != CFPQ_OK) {
            g_ctx = nullptr;
            throw std::runtime_error("cfpq_create failed (" + std::to_string(rc) + "): no usable GPU");
        }
        std::atexit([] {
            if (g_ctx) cfpq_destroy(g_ctx);
            g_ctx = nullptr;
        });
    });
    if (!g_ctx) throw std::runtime_error("cfpq engine context unavailable");
    return g_ctx;
}

HipMatrix::HipMatrix(unsigned int n) : Matrix(n), n_(n), wpr_((n + 63) / 64), host_((size_t)n * ((n + 63) / 64), 0) {}

HipMatrix::~HipMatrix() {
    if (dev_ && g_ctx) cfpq_mat_free(g_ctx, dev_);
}

void HipMatrix::set_bit(unsigned int row, unsigned col) {
    if (row >= n_ || col >= n_) throw std::out_of_range("HipMatrix::set_bit");
    host_[(size_t)row * wpr_ + (col >> 6)] |= 1ull << (col & 63);
    have_fp_ = false;
}

unsigned int HipMatrix::get_bit(unsigned int row, unsigned col) {
    if (row >= n_ || col >= n_) return 0;
    return (unsigned int)((host_[(size_t)row * wpr_ + (col >> 6)] >> (col & 63)) & 1ull);
}

bool HipMatrix::add_mul(Matrix *left, Matrix *right) {
    HipMatrix *a = as_hip(left), *b = as_hip(right);
    if (n_ == 0) return false;
    if (!dev_ || !a->dev_ || !b->dev_) throw std::runtime_error("HipMatrix::add_mul outside preprocessing/postprocessing");
    int changed = 0;
    check(cfpq_add_mul(cfpq_host_context(), dev_, a->dev_, b->dev_, &changed), "cfpq_add_mul");
    return changed != 0;
}

bool HipMatrix::fingerprint(uint64_t *hash, uint64_t *count) {
    if (dev_) {
        check(cfpq_mat_fingerprint(cfpq_host_context(), dev_, hash, count), "cfpq_mat_fingerprint");
        return true;
    }
    if (have_fp_) {
        *hash = fp_hash_;
        *count = fp_count_;
        return true;
    }
    return false;
}

void HipMatrix::to_device() {
    if (n_ == 0) return;  // an empty graph: nothing to hold
    cfpq_ctx *ctx = cfpq_host_context();
    if (!dev_) check(cfpq_mat_alloc(ctx, n_, &dev_), "cfpq_mat_alloc");
    check(cfpq_mat_upload(ctx, dev_, host_.data(), wpr_), "cfpq_mat_upload");
}

void HipMatrix::from_device() {
    if (!dev_) return;
    cfpq_ctx *ctx = cfpq_host_context();
    check(cfpq_mat_download(ctx, dev_, host_.data(), wpr_), "cfpq_mat_download");
    // keep the fingerprint computed on the device so print_results stays O(nnz)
    check(cfpq_mat_fingerprint(ctx, dev_, &fp_hash_, &fp_count_), "cfpq_mat_fingerprint");
    have_fp_ = true;
    check(cfpq_mat_free(ctx, dev_), "cfpq_mat_free");
    dev_ = nullptr;
}

void HipMatricesEnv::environment_preprocessing(const std::vector<Matrix *> &matrices) {
    for (Matrix *m : matrices) as_hip(m)->to_device();
}

void HipMatricesEnv::environment_postprocessing(const std::vector<Matrix *> &matrices) {
    for (Matrix *m : matrices) as_hip(m)->from_device();
}

bool HipMatricesEnv::run_fixpoint(const std::vector<hip_rule> &rules, const std::vector<Matrix *> &matrices) {
    if (matrices.empty() || as_hip(matrices[0])->size() == 0) return true;  // nothing to compute
    std::vector<cfpq_rule> r(rules.size());
    for (size_t i = 0; i < rules.size(); ++i) r[i] = {rules[i].first, rules[i].second.first, rules[i].second.second};
    std::vector<cfpq_mat *> mats(matrices.size());
    for (size_t i = 0; i < matrices.size(); ++i) mats[i] = as_hip(matrices[i])->device();
    check(cfpq_fixpoint(cfpq_host_context(), r.data(), r.size(), mats.data(), mats.size(), &stats_), "cfpq_fixpoint");
    return true;
}
