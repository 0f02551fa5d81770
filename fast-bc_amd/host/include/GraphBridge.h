/*
 * Bridge from the reference-style graph interface to the engine's CSR.
 * A DirectedWeightedGraph hands over its own CSR; any other IGraph (a
 * SubGraph, a user implementation) is read through vertices()/forwardStar(),
 * with vertex ids 0..vertices().size()-1 as the reference requires of the
 * graphs its computers accept (brandes/IBrandesBC.h:17-23).
 */
#ifndef FASTBC_AMD_GRAPHBRIDGE_H
#define FASTBC_AMD_GRAPHBRIDGE_H

#include "DirectedWeightedGraph.h"

#include <memory>

namespace fastbc {
namespace bridge {

template <typename V, typename W>
std::shared_ptr<const fbchost::Graph> host_graph_ptr(const IGraph<V, W> &g)
{
    if (auto *dwg = dynamic_cast<const DirectedWeightedGraph<V, W> *>(&g))
        return std::shared_ptr<const fbchost::Graph>(std::shared_ptr<const fbchost::Graph>(), &dwg->hostGraph());
    std::vector<int> s, d;
    std::vector<double> w;
    for (const V &v : g.vertices())
        for (const auto &e : g.forwardStar(v)) {
            s.push_back((int)v);
            d.push_back((int)e.first);
            w.push_back((double)e.second);
        }
    return std::make_shared<const fbchost::Graph>(
        fbchost::from_arcs_n((int)g.vertices().size(), s.size(), s.data(), d.data(), w.data()));
}

/* Reference to the CSR of a DirectedWeightedGraph (the only IDegreeGraph
 * the reference's partitioner is given, main.cpp:208-209). */
template <typename V, typename W>
const fbchost::Graph &host_graph(const IGraph<V, W> &g)
{
    if (auto *dwg = dynamic_cast<const DirectedWeightedGraph<V, W> *>(&g)) return dwg->hostGraph();
    thread_local std::shared_ptr<const fbchost::Graph> keep;
    keep = host_graph_ptr<V, W>(g);
    return *keep;
}

} // namespace bridge
} // namespace fastbc

#endif
