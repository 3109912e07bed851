// CNF grammar and the CFPQ driver (reference interface: cfpq-cpp/Grammar.h:17-100).
//
//   Grammar(file)                       CNF lines: "H B1 B2" | "NT term" | "NT" (epsilon),
//                                       nonterminals numbered by first appearance
//   intersection_with_graph<M, E>(g)    one M per nonterminal, terminal and epsilon
//                                       bits, then the fixpoint; returns
//                                       (total_ms incl. env pre/post, algorithm_ms)
//   print_results(file)                 "name hash count" per nonterminal, name order
//
// The fixpoint is the environment's run_fixpoint when it provides one (the
// device-resident engine), otherwise the worklist of Grammar.h:58-69: apply a
// rule, and when its head grew, schedule every rule that reads the head.  The
// least fixpoint is unique, so the scheduling order does not affect results.
#ifndef CFPQ_GRAMMAR_H
#define CFPQ_GRAMMAR_H

#include <chrono>
#include <deque>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "Graph.h"
#include "Matrix.h"

using nonterminals_pair = std::pair<unsigned int, unsigned int>;

class Grammar {
public:
    explicit Grammar(const std::string &grammar_filename);
    virtual ~Grammar();

    Grammar(const Grammar &) = delete;
    Grammar &operator=(const Grammar &) = delete;
    Grammar(Grammar &&other) noexcept;

    template <class T1 = Matrix, class T2 = MatricesEnv>
    std::pair<unsigned int, unsigned int> intersection_with_graph(Graph &graph) {
        using clock = std::chrono::steady_clock;
        T2 env;
        release_matrices();
        vertices_count = graph.vertices_count;
        for (unsigned int x = 0; x < nonterminals_count; ++x) matrices.push_back(new T1(vertices_count));
        for (const auto &e : graph.edges) {
            auto it = terminal_to_nonterminals.find(e.first);
            if (it == terminal_to_nonterminals.end()) continue;
            for (int nt : it->second) matrices[nt]->set_bit(e.second.first, e.second.second);
        }
        for (unsigned int nt : epsilon_nonterminals)
            for (unsigned int v = 0; v < vertices_count; ++v) matrices[nt]->set_bit(v, v);

        const auto t_begin = clock::now();
        env.environment_preprocessing(matrices);
        const auto t_algo = clock::now();
        if (!env.run_fixpoint(rules, matrices)) add_mul_calls = worklist();
        const auto t_algo_end = clock::now();
        env.environment_postprocessing(matrices);
        const auto t_end = clock::now();

        auto ms = [](clock::duration d) {
            return (unsigned int)std::chrono::duration_cast<std::chrono::milliseconds>(d).count();
        };
        return {ms(t_end - t_begin), ms(t_algo_end - t_algo)};
    }

    void print_results(const std::string &output_filename);
    // name -> (hash, count), in name order
    std::map<std::string, std::pair<uint64_t, uint64_t>> results();

    unsigned int get_rules_size();
    const std::unordered_map<std::string, std::vector<int>> &get_nonterminal_from_terminal();
    const std::vector<cfpq_rule_t> &get_rules() const { return rules; }
    unsigned int get_nonterminals_count() const { return nonterminals_count; }
    const std::vector<unsigned int> &get_epsilon_nonterminals() const { return epsilon_nonterminals; }
    const std::map<std::string, unsigned int> &get_nonterminal_to_index() const { return nonterminal_to_index; }
    unsigned long long get_add_mul_calls() const { return add_mul_calls; }
    Matrix *matrix(const std::string &nonterminal);

private:
    unsigned long long worklist();
    void release_matrices();
    unsigned int index_of(const std::string &name);

    unsigned int nonterminals_count = 0;
    unsigned int vertices_count = 0;
    unsigned long long add_mul_calls = 0;

    std::map<std::string, unsigned int> nonterminal_to_index;
    std::vector<unsigned int> epsilon_nonterminals;
    std::unordered_map<std::string, std::vector<int>> terminal_to_nonterminals;
    std::vector<cfpq_rule_t> rules;
    std::vector<Matrix *> matrices;
};

#endif  // CFPQ_GRAMMAR_H
