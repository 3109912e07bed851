// Labelled edge list (reference interface: cfpq-cpp/Graph.h, Graph.cpp:11-20).
// File format: one edge per line, "from label to"; vertices_count is the
// largest vertex id + 1.
#ifndef CFPQ_GRAPH_H
#define CFPQ_GRAPH_H

#include <string>
#include <utility>
#include <vector>

using edge = std::pair<std::string, std::pair<unsigned int, unsigned int>>;

class Graph {
public:
    explicit Graph(const std::string &graph_filename);
    Graph() = default;
    virtual ~Graph() = default;

    void add_edge(unsigned int from, const std::string &label, unsigned int to);

    std::vector<edge> edges;
    unsigned int vertices_count = 0;
};

#endif  // CFPQ_GRAPH_H
