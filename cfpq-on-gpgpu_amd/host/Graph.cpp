// Edge-list reader (reference: cfpq-cpp/Graph.cpp:11-20).
#include "Graph.h"

#include <algorithm>
#include <fstream>
#include <stdexcept>

Graph::Graph(const std::string &graph_filename) {
    std::ifstream in(graph_filename);
    if (!in) throw std::runtime_error("cannot open graph file " + graph_filename);
    unsigned int from, to;
    std::string label;
    while (in >> from >> label >> to) add_edge(from, label, to);
}

void Graph::add_edge(unsigned int from, const std::string &label, unsigned int to) {
    edges.push_back({label, {from, to}});
    vertices_count = std::max(vertices_count, std::max(from, to) + 1);
}
