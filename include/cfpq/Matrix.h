// Matrix plugin interface of the CFPQ driver.
//
// Same contract as the reference's cfpq-cpp/Matrix.h:7-37: a backend derives
// from Matrix and implements set_bit / get_bit / add_mul; an optional
// MatricesEnv brackets the algorithm with pre- and post-processing.  Two
// optional hooks are added, both with defaults that keep the reference
// behaviour for any plain backend:
//   * MatricesEnv::run_fixpoint -- lets an environment that can run the whole
//     fixpoint itself (the device-resident engine) take over the worklist loop;
//     returning false (the default) leaves the worklist to the driver;
//   * Matrix::fingerprint -- lets a backend compute the result fingerprint
//     without N^2 get_bit calls; returning false falls back to get_bit.
//
// The include guard differs from the reference header's, so that a build
// which sees both interfaces fails loudly instead of dropping the hooks.
// CFPQ_DRIVER_HOOKS tells backends (HipMatrix.h) that the hooks exist.
#ifndef CFPQ_HIP_DRIVER_MATRIX_H
#define CFPQ_HIP_DRIVER_MATRIX_H
#define CFPQ_DRIVER_HOOKS 1

#include <cstdint>
#include <utility>
#include <vector>

class Matrix {
public:
    explicit Matrix(unsigned int n) { (void)n; }
    virtual ~Matrix() = default;

    virtual void set_bit(unsigned int row, unsigned col) = 0;
    virtual unsigned int get_bit(unsigned int row, unsigned col) = 0;
    // this |= left * right over (OR, AND); true if a bit was added.
    // left and/or right may be this matrix.
    virtual bool add_mul(Matrix *left, Matrix *right) = 0;

    // (hash, count) as Grammar::print_results defines them; false = not provided.
    virtual bool fingerprint(uint64_t *hash, uint64_t *count) {
        (void)hash;
        (void)count;
        return false;
    }

    // One-hot nonterminal masks (cfpq-cpp/Matrix.h:19-25).
    static unsigned int toBoolVector(unsigned int number) { return 1u << number; }
    static unsigned char toBoolVector8(unsigned int number) { return (unsigned char)(1u << number); }
};

// A binary rule head -> (left, right), as indices into the matrix vector.
using cfpq_rule_t = std::pair<unsigned int, std::pair<unsigned int, unsigned int>>;

class MatricesEnv {
public:
    MatricesEnv() = default;
    virtual ~MatricesEnv() = default;

    virtual void environment_preprocessing(const std::vector<Matrix *> &matrices) { (void)matrices; }
    virtual void environment_postprocessing(const std::vector<Matrix *> &matrices) { (void)matrices; }

    // Compute the least fixpoint of matrices[h] |= matrices[l] * matrices[r]
    // over all rules in place.  Return false to let the driver run its worklist.
    virtual bool run_fixpoint(const std::vector<cfpq_rule_t> &rules, const std::vector<Matrix *> &matrices) {
        (void)rules;
        (void)matrices;
        return false;
    }
};

#endif  // CFPQ_HIP_DRIVER_MATRIX_H
