// Reference-driver harness (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile against the reference's own, unmodified
// cfpq-cpp/Grammar.cpp and cfpq-cpp/Graph.cpp (read where they lie, never
// copied).  It plugs the oracle's M4RM product into the reference's Matrix
// plugin interface (cfpq-cpp/Matrix.h:7-26) exactly as m4ri/main.cpp:7-15
// plugs in M4riMatrix, so the reference's grammar parser, graph parser,
// worklist fixpoint (Grammar.h:23-81) and fingerprint writer
// (Grammar.cpp:81-101) run as shipped.  libm4ri itself is absent from this
// image, so the matrix arithmetic is the oracle's restatement.
//
// usage: cfpq_ref <grammar> <graph> <out>
//   stdout: "<total_ms> <algorithm_ms>"  (as m4ri/main.cpp:11)
//   stderr: "add_mul_calls <count>"
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <vector>
#include "Grammar.h"
#include "Graph.h"
#include "Matrix.h"

extern "C" int oracle_addmul_m4rm(uint64_t *C, const uint64_t *A, const uint64_t *B, uint32_t n);

static unsigned long long g_calls = 0;

class OracleMatrix : public Matrix {
public:
    explicit OracleMatrix(unsigned int n) : Matrix(n), n_(n), w_((n + 63) / 64), bits_((size_t)n * ((n + 63) / 64), 0) {}
    void set_bit(unsigned int row, unsigned col) override {
        bits_[(size_t)row * w_ + (col >> 6)] |= 1ull << (col & 63);
    }
    unsigned int get_bit(unsigned int row, unsigned col) override {
        return (unsigned int)((bits_[(size_t)row * w_ + (col >> 6)] >> (col & 63)) & 1ull);
    }
    bool add_mul(Matrix *left, Matrix *right) override {
        auto *a = dynamic_cast<OracleMatrix *>(left);
        auto *b = dynamic_cast<OracleMatrix *>(right);
        ++g_calls;
        return oracle_addmul_m4rm(bits_.data(), a->bits_.data(), b->bits_.data(), n_) != 0;
    }
private:
    unsigned int n_;
    size_t w_;
    std::vector<uint64_t> bits_;
};

int main(int argc, char *argv[]) {
    if (argc < 4) {
        std::cerr << "usage: cfpq_ref grammar graph out\n";
        return 2;
    }
    Grammar grammar = Grammar(argv[1]);
    Graph graph = Graph(argv[2]);
    auto times = grammar.intersection_with_graph<OracleMatrix>(graph);
    std::cout << times.first << ' ' << times.second << std::endl;
    grammar.print_results(argv[3]);
    std::cerr << "add_mul_calls " << g_calls << std::endl;
    return 0;
}
