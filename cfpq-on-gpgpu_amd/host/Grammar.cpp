// CNF grammar parser, worklist driver and result fingerprint.
// Behaviour follows the reference's cfpq-cpp/Grammar.cpp:17-101 and
// Grammar.h:58-69; see include/cfpq/Grammar.h.
#include "Grammar.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace {
constexpr uint64_t kHashMod = 1000000007ull;  // Grammar.cpp:13
constexpr uint64_t kBaseCol = 11;             // Grammar.cpp:11, per column
constexpr uint64_t kBaseRow = 13;             // Grammar.cpp:12, per row
}  // namespace

unsigned int Grammar::index_of(const std::string &name) {
    auto it = nonterminal_to_index.find(name);
    if (it != nonterminal_to_index.end()) return it->second;
    nonterminal_to_index.emplace(name, nonterminals_count);
    return nonterminals_count++;
}

Grammar::Grammar(const std::string &grammar_filename) {
    std::ifstream in(grammar_filename);
    if (!in) throw std::runtime_error("cannot open grammar file " + grammar_filename);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::vector<std::string> tok;
        for (std::string t; ss >> t;) tok.push_back(t);
        if (tok.empty()) continue;
        const unsigned int head = index_of(tok[0]);
        switch (tok.size()) {
        case 1:  // NT -> epsilon
            epsilon_nonterminals.push_back(head);
            break;
        case 2:  // NT -> terminal
            terminal_to_nonterminals[tok[1]].push_back((int)head);
            break;
        case 3: {  // NT -> B1 B2
            const unsigned int b1 = index_of(tok[1]);
            const unsigned int b2 = index_of(tok[2]);
            rules.push_back({head, {b1, b2}});
            break;
        }
        default:  // longer lines carry no rule (as in the reference)
            break;
        }
    }
}

Grammar::Grammar(Grammar &&o) noexcept
    : nonterminals_count(o.nonterminals_count), vertices_count(o.vertices_count), add_mul_calls(o.add_mul_calls),
      nonterminal_to_index(std::move(o.nonterminal_to_index)), epsilon_nonterminals(std::move(o.epsilon_nonterminals)),
      terminal_to_nonterminals(std::move(o.terminal_to_nonterminals)), rules(std::move(o.rules)),
      matrices(std::move(o.matrices)) {
    o.matrices.clear();
}

Grammar::~Grammar() { release_matrices(); }

void Grammar::release_matrices() {
    for (Matrix *m : matrices) delete m;
    matrices.clear();
}

unsigned long long Grammar::worklist() {
    // rules reading each nonterminal
    std::vector<std::vector<unsigned int>> readers(nonterminals_count);
    for (unsigned int r = 0; r < rules.size(); ++r) {
        readers[rules[r].second.first].push_back(r);
        if (rules[r].second.second != rules[r].second.first) readers[rules[r].second.second].push_back(r);
    }
    std::deque<unsigned int> queue;
    std::vector<char> queued(rules.size(), 1);
    for (unsigned int r = 0; r < rules.size(); ++r) queue.push_back(r);
    unsigned long long calls = 0;
    while (!queue.empty()) {
        const unsigned int r = queue.front();
        queue.pop_front();
        queued[r] = 0;
        const unsigned int head = rules[r].first;
        ++calls;
        if (matrices[head]->add_mul(matrices[rules[r].second.first], matrices[rules[r].second.second])) {
            for (unsigned int q : readers[head])
                if (!queued[q]) {
                    queued[q] = 1;
                    queue.push_back(q);
                }
        }
    }
    return calls;
}

std::map<std::string, std::pair<uint64_t, uint64_t>> Grammar::results() {
    std::map<std::string, std::pair<uint64_t, uint64_t>> out;
    const unsigned int n = vertices_count;
    std::vector<uint64_t> pcol, prow;
    for (const auto &kv : nonterminal_to_index) {
        Matrix *m = kv.second < matrices.size() ? matrices[kv.second] : nullptr;
        uint64_t hash = 0, count = 0;
        if (m && !m->fingerprint(&hash, &count)) {
            // hash = sum_i line_i * 13^(n-1-i),  line_i = sum_{j: bit} 11^(n-1-j)  (mod P):
            // the closed form of the per-cell Horner loop of Grammar.cpp:84-97
            if (pcol.empty()) {
                pcol.resize(n);
                prow.resize(n);
                uint64_t a = 1, b = 1;
                for (unsigned int e = 0; e < n; ++e) {
                    pcol[e] = a;
                    prow[e] = b;
                    a = a * kBaseCol % kHashMod;
                    b = b * kBaseRow % kHashMod;
                }
            }
            for (unsigned int i = 0; i < n; ++i) {
                uint64_t line = 0;
                for (unsigned int j = 0; j < n; ++j)
                    if (m->get_bit(i, j)) {
                        ++count;
                        line = (line + pcol[n - 1 - j]) % kHashMod;
                    }
                hash = (hash + line * prow[n - 1 - i]) % kHashMod;
            }
        }
        out[kv.first] = {hash, count};
    }
    return out;
}

void Grammar::print_results(const std::string &output_filename) {
    std::ofstream out(output_filename);
    if (!out) throw std::runtime_error("cannot write " + output_filename);
    for (const auto &kv : results()) out << kv.first << ' ' << kv.second.first << ' ' << kv.second.second << '\n';
}

unsigned int Grammar::get_rules_size() { return (unsigned int)rules.size(); }

const std::unordered_map<std::string, std::vector<int>> &Grammar::get_nonterminal_from_terminal() {
    return terminal_to_nonterminals;
}

Matrix *Grammar::matrix(const std::string &nonterminal) {
    auto it = nonterminal_to_index.find(nonterminal);
    if (it == nonterminal_to_index.end() || it->second >= matrices.size()) return nullptr;
    return matrices[it->second];
}
