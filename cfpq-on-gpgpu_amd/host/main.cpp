// Command-line runner with the reference backends' interface
// (m4ri/main.cpp:7-15, cuda/naive-gpu/main.cpp:7-16):
//
//   cfpq_hip <grammar> <graph> <out> [--worklist]
//
// prints "<total_ms> <algorithm_ms>" and writes "name hash count" lines.
// --worklist runs the reference's per-rule worklist over cfpq_add_mul instead
// of the device-resident fixpoint.
#include <cstring>
#include <iostream>
#include <stdexcept>

#include "Grammar.h"
#include "Graph.h"
#include "HipMatrix.h"

int main(int argc, char *argv[]) {
    if (argc < 4) {
        std::cerr << "usage: cfpq_hip <grammar> <graph> <out> [--worklist]\n";
        return 2;
    }
    const bool worklist = argc > 4 && std::strcmp(argv[4], "--worklist") == 0;
    try {
        Grammar grammar(argv[1]);
        Graph graph(argv[2]);
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
