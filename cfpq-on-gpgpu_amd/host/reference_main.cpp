// The backend main a maintainer adds to the reference as hip/main.cpp: the
// shape of the reference's own backend mains (m4ri/main.cpp:7-15,
// cuda/naive-gpu/main.cpp:7-16) with HipMatrix plugged in.  It builds
// against the reference's unmodified cfpq-cpp sources (-I cfpq-cpp first),
// whose worklist then runs with one exact cfpq_add_mul per rule; against this
// repository's include/cfpq driver the same main hands the fixpoint to the
// device (HipMatricesEnv::run_fixpoint).
//
//   usage: <exe> <grammar> <graph> <out>    stdout: "<total_ms> <algorithm_ms>"
#include <iostream>

#include "Grammar.h"
#include "Graph.h"
#include "HipMatrix.h"

int main(int argc, char *argv[]) {
    if (argc < 4) {
        std::cerr << "usage: " << argv[0] << " <grammar> <graph> <out>" << std::endl;
        return 2;
    }
    Grammar grammar(argv[1]);
    Graph graph(argv[2]);
    auto times = grammar.intersection_with_graph<HipMatrix, HipMatricesEnv>(graph);
    std::cout << times.first << " " << times.second << std::endl;
    grammar.print_results(argv[3]);
    return 0;
}
