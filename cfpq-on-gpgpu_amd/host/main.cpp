// Command-line runner with the reference backends' interface
// (m4ri/main.cpp:7-15, cuda/naive-gpu/main.cpp:7-16):
//
//   cfpq_hip <grammar> <graph> <out> [--worklist | --pairs | --bits]
//
// prints "<total_ms> <algorithm_ms>" and writes "name hash count" lines.
// --worklist runs the reference's per-rule worklist over cfpq_add_mul instead
// of the device-resident fixpoint.  --pairs runs the sparse pair-list
// fixpoint (cfpq_pairs_fixpoint, the counterpart of the GraphBLAS backend's
// runner, graphBlas/main.cpp:11-20): relations as sorted pair keys, no n x n
// matrices, so graphs far beyond the bit tiles' size run too.  Without a
// flag the bit matrices are used when they fit (kBitsBudget) and pair lists
// otherwise; --bits forces the matrices.
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Grammar.h"
#include "Graph.h"
#include "HipMatrix.h"

namespace {

constexpr uint64_t kHashMod = 1000000007ull;  // cfpq-cpp/Grammar.cpp:13
// Device bytes the bit-matrix engine may plan for: a matrix and its two delta
// buffers per nonterminal take about 6 n^2 / 8 bytes, the bits and their
// column-major copy (of 288 GB of HBM).
constexpr double kBitsBudget = 200.0 * (1ull << 30);

uint64_t pow_mod(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= kHashMod;
    for (; e; e >>= 1, b = b * b % kHashMod)
        if (e & 1) r = r * b % kHashMod;
    return r;
}

// The fingerprint of cfpq-cpp/Grammar.cpp:81-101 over a set of distinct
// pairs: bit (i, j) adds 11^(n-1-j) * 13^(n-1-i).
std::pair<uint64_t, uint64_t> fingerprint_pairs(const std::vector<uint64_t> &keys, uint64_t n) {
    uint64_t h = 0;
    for (uint64_t k : keys) {
        const uint64_t i = k >> 32, j = k & 0xffffffffull;
        h = (h + pow_mod(11, n - 1 - j) * pow_mod(13, n - 1 - i)) % kHashMod;
    }
    return {h, keys.size()};
}

std::pair<unsigned int, unsigned int> run_pairs(Grammar &grammar, Graph &graph, const std::string &out_path) {
    using clock = std::chrono::steady_clock;
    const unsigned int nnt = grammar.get_nonterminals_count();
    const uint64_t n = graph.vertices_count;
    std::vector<std::vector<uint64_t>> keys(nnt);
    const auto &by_terminal = grammar.get_nonterminal_from_terminal();
    for (const auto &e : graph.edges) {
        auto it = by_terminal.find(e.first);
        if (it == by_terminal.end()) continue;
        for (int nt : it->second) keys[nt].push_back((uint64_t)e.second.first << 32 | e.second.second);
    }
    for (unsigned int nt : grammar.get_epsilon_nonterminals())
        for (uint64_t v = 0; v < n; ++v) keys[nt].push_back(v << 32 | v);
    cfpq_ctx *ctx = cfpq_host_context();
    const auto &rules = grammar.get_rules();
    std::vector<cfpq_rule> rr;
    for (const auto &r : rules) rr.push_back(cfpq_rule{r.first, r.second.first, r.second.second});
    std::vector<uint64_t> initial(nnt);
    for (unsigned int x = 0; x < nnt; ++x) initial[x] = keys[x].size();
    const auto t_begin = clock::now();
    std::vector<cfpq_pairs> sets(nnt);
    cfpq_stats st{};
    int rc = CFPQ_E_CAPACITY;
    // a first capacity of 4x the input; when a closed relation outgrows it the
    // call reports the sizes needed (inputs untouched) and is made once more
    for (unsigned int x = 0; x < nnt; ++x) keys[x].resize(4 * initial[x] + 4096);
    for (int attempt = 0; attempt < 2 && rc == CFPQ_E_CAPACITY; ++attempt) {
        for (unsigned int x = 0; x < nnt; ++x) sets[x] = cfpq_pairs{keys[x].data(), initial[x], keys[x].size()};
        rc = cfpq_pairs_fixpoint(ctx, rr.data(), rr.size(), sets.data(), nnt, (uint32_t)n, &st);
        if (rc == CFPQ_E_CAPACITY)
            for (unsigned int x = 0; x < nnt; ++x)
                keys[x].resize(std::max<uint64_t>(sets[x].count, keys[x].size()));
    }
    if (rc != CFPQ_OK) throw std::runtime_error(std::string("cfpq_pairs_fixpoint: ") + cfpq_last_error(ctx));
    const auto t_end = clock::now();
    std::ofstream out(out_path);
    if (!out) throw std::runtime_error("cannot write " + out_path);
    for (const auto &kv : grammar.get_nonterminal_to_index()) {
        std::vector<uint64_t> &k = keys[kv.second];
        k.resize(sets[kv.second].count);
        const auto f = fingerprint_pairs(k, n);
        out << kv.first << ' ' << f.first << ' ' << f.second << '\n';
    }
    const unsigned int total =
        (unsigned int)std::chrono::duration_cast<std::chrono::milliseconds>(t_end - t_begin).count();
    return {total, (unsigned int)st.fixpoint_ms};
}

}  // namespace

int main(int argc, char *argv[]) {
    if (argc < 4) {
        std::cerr << "usage: cfpq_hip <grammar> <graph> <out> [--worklist | --pairs | --bits]\n";
        return 2;
    }
    const char *mode = argc > 4 ? argv[4] : "";
    if (*mode && std::strcmp(mode, "--worklist") && std::strcmp(mode, "--pairs") && std::strcmp(mode, "--bits")) {
        std::cerr << "cfpq_hip: unknown option " << mode << "\n";
        return 2;
    }
    const bool worklist = std::strcmp(mode, "--worklist") == 0;
    try {
        Grammar grammar(argv[1]);
        Graph graph(argv[2]);
        const double n = graph.vertices_count;
        const bool pairs = std::strcmp(mode, "--pairs") == 0 ||
                           (!*mode && 6.0 * grammar.get_nonterminals_count() * n * n / 8 > kBitsBudget);
        if (pairs) {
            auto times = run_pairs(grammar, graph, argv[3]);
            std::cout << times.first << ' ' << times.second << std::endl;
            return 0;
        }
        auto times = worklist ? grammar.intersection_with_graph<HipMatrix, HipWorklistEnv>(graph)
                              : grammar.intersection_with_graph<HipMatrix, HipMatricesEnv>(graph);
        std::cout << times.first << ' ' << times.second << std::endl;
        grammar.print_results(argv[3]);
    } catch (const std::exception &e) {
        std::cerr << "cfpq_hip: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
