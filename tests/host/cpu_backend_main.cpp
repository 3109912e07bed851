// Test harness (CPU only): the C++ drop-in host (include/cfpq/Grammar.h,
// Graph.h; cfpq-on-gpgpu_amd/host/Grammar.cpp, Graph.cpp) driven by a
// plain bitset Matrix, so the parser, worklist and fingerprint of the host
// can be checked against the reference driver without a GPU.
//   usage: cfpq_host_cpu <grammar> <graph> <out>   -> stdout "calls <n>"
#include <cstdint>
#include <iostream>
#include <vector>
#include "Grammar.h"
#include "Graph.h"

class BitsetMatrix : public Matrix {
public:
    explicit BitsetMatrix(unsigned int n) : Matrix(n), n_(n), w_((n + 63) / 64), b_((size_t)n * ((n + 63) / 64)) {}
    void set_bit(unsigned int r, unsigned c) override { b_[(size_t)r * w_ + (c >> 6)] |= 1ull << (c & 63); }
    unsigned int get_bit(unsigned int r, unsigned c) override { return (b_[(size_t)r * w_ + (c >> 6)] >> (c & 63)) & 1; }
    bool add_mul(Matrix *l, Matrix *r) override {
        auto *A = dynamic_cast<BitsetMatrix *>(l);
        auto *B = dynamic_cast<BitsetMatrix *>(r);
        std::vector<uint64_t> out(b_);
        for (unsigned int i = 0; i < n_; ++i)
            for (unsigned int k = 0; k < n_; ++k)
                if (A->get_bit(i, k))
                    for (size_t j = 0; j < w_; ++j) out[(size_t)i * w_ + j] |= B->b_[(size_t)k * w_ + j];
        const bool changed = out != b_;
        b_.swap(out);
        return changed;
    }
private:
    unsigned int n_;
    size_t w_;
    std::vector<uint64_t> b_;
};

int main(int argc, char *argv[]) {
    if (argc < 4) return 2;
    Grammar g(argv[1]);
    Graph graph(argv[2]);
    g.intersection_with_graph<BitsetMatrix>(graph);
    g.print_results(argv[3]);
    std::cout << "calls " << g.get_add_mul_calls() << std::endl;
    return 0;
}
