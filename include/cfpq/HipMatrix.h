// gfx950 backend for the CFPQ driver: drops in beside the reference's
// M4riMatrix (m4ri/m4ri_matrix.h:9-25) and gpuMatrix / gpuMatricesEnv
// (cuda/naive-gpu/GpuMatrix.h:9-58):
//
//   grammar.intersection_with_graph<HipMatrix, HipMatricesEnv>(graph);   // device fixpoint
//   grammar.intersection_with_graph<HipMatrix, HipWorklistEnv>(graph);   // reference worklist,
//                                                                         // one cfpq_add_mul per rule
//
// Like gpuMatrix, set_bit writes a host mirror before preprocessing and
// get_bit reads it after postprocessing; in between the matrix lives in HBM
// (bit-packed 256 x 256 tiles) and is touched only through the C ABI of
// include/cfpq_hip.h.  All matrices share one process-wide engine context on
// the device named by $CFPQ_DEVICE (default 0).  Errors throw
// std::runtime_error carrying cfpq_last_error().
//
// Two driver interfaces are accepted.  <Matrix.h> is taken from the include
// path, so whichever driver the build names first is the one this backend
// derives from:
//   * the reference's own cfpq-cpp/Matrix.h (-I cfpq-cpp first): the plain
//     Matrix / MatricesEnv contract of cfpq-cpp/Matrix.h:7-37.  The driver's
//     worklist (cfpq-cpp/Grammar.h:58-69) runs and every add_mul is one exact
//     cfpq_add_mul; fingerprint() and run_fixpoint() are extra members the
//     reference driver never calls;
//   * this repository's include/cfpq/Matrix.h, which adds two optional hooks
//     (CFPQ_DRIVER_HOOKS): the same members then override them, so
//     HipMatricesEnv hands the whole fixpoint to cfpq_fixpoint and
//     print_results reads device fingerprints.
#ifndef CFPQ_HIP_MATRIX_H
#define CFPQ_HIP_MATRIX_H

#include <cstdint>
#include <utility>
#include <vector>

#include <Matrix.h>
#include "cfpq_hip.h"

#ifdef CFPQ_DRIVER_HOOKS
#define CFPQ_HIP_HOOK override
#else
#define CFPQ_HIP_HOOK
#endif

cfpq_ctx *cfpq_host_context();  // process-wide context (created on first use)

// A binary rule head -> (left, right) as the driver stores it: the element
// type of the reference's Grammar::rules (cfpq-cpp/Grammar.h:98).
using hip_rule = std::pair<unsigned int, std::pair<unsigned int, unsigned int>>;

class HipMatrix : public Matrix {
public:
    explicit HipMatrix(unsigned int n);
    ~HipMatrix() override;

    void set_bit(unsigned int row, unsigned col) override;
    unsigned int get_bit(unsigned int row, unsigned col) override;
    bool add_mul(Matrix *left, Matrix *right) override;
    // (hash, count) of Grammar::print_results, computed on the device
    bool fingerprint(uint64_t *hash, uint64_t *count) CFPQ_HIP_HOOK;

    void to_device();    // allocate + upload the host mirror
    void from_device();  // download into the host mirror, free device memory
    cfpq_mat *device() const { return dev_; }
    unsigned int size() const { return n_; }

private:
    unsigned int n_;
    size_t wpr_;
    std::vector<uint64_t> host_;
    cfpq_mat *dev_ = nullptr;
    bool have_fp_ = false;
    uint64_t fp_hash_ = 0, fp_count_ = 0;
};

// Uploads before and downloads after the algorithm; run_fixpoint hands the
// whole fixpoint to the device engine (cfpq_fixpoint) where the driver calls it.
class HipMatricesEnv : public MatricesEnv {
public:
    void environment_preprocessing(const std::vector<Matrix *> &matrices) override;
    void environment_postprocessing(const std::vector<Matrix *> &matrices) override;
    bool run_fixpoint(const std::vector<hip_rule> &rules, const std::vector<Matrix *> &matrices) CFPQ_HIP_HOOK;
    const cfpq_stats &stats() const { return stats_; }

private:
    cfpq_stats stats_{};
};

// Same residency, but the driver's worklist runs and calls add_mul per rule.
class HipWorklistEnv : public HipMatricesEnv {
public:
    bool run_fixpoint(const std::vector<hip_rule> &, const std::vector<Matrix *> &) CFPQ_HIP_HOOK { return false; }
};

#endif  // CFPQ_HIP_MATRIX_H
