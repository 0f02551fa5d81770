/*
 * SubGraph (SubGraph.h:59-148): a vertex set of a reference graph.  Border =
 * vertex with an arc leaving or entering the set; forwardStar/backwardStar are
 * restricted to the set; edges() counts the internal arcs.
 */
#ifndef FASTBC_AMD_SUBGRAPH_H
#define FASTBC_AMD_SUBGRAPH_H

#include "ISubGraph.h"

#include <map>
#include <memory>
#include <set>
#include <vector>

namespace fastbc {

template <typename V, typename W>
class SubGraph : public ISubGraph<V, W> {
public:
    SubGraph(const std::vector<V> &subGraphVertices, std::shared_ptr<const IGraph<V, W>> referenceGraph)
        : _ref(referenceGraph), _vertices(subGraphVertices), _edges(0)
    {
        std::set<V> in(_vertices.begin(), _vertices.end());
        for (const V &v : _vertices) {
            auto &fs = _fs[v];
            auto &bs = _bs[v];
            for (const auto &e : _ref->forwardStar(v)) {
                if (in.count(e.first)) {
                    fs.emplace_hint(fs.end(), e.first, e.second);
                    ++_edges;
                } else {
                    _borders.insert(v);
                }
            }
            for (const auto &e : _ref->backwardStar(v)) {
                if (in.count(e.first)) bs.emplace_hint(bs.end(), e.first, e.second);
                else _borders.insert(v);
            }
        }
    }

    W edge(V src, V dest) const override
    {
        const auto &f = forwardStar(src);
        auto it = f.find(dest);
        return it == f.end() ? W(0) : it->second;
    }
    const std::map<V, W> &forwardStar(V src) const override
    {
        auto it = _fs.find(src);
        return it != _fs.end() ? it->second : _ref->forwardStar(src);
    }
    const std::map<V, W> &backwardStar(V dest) const override
    {
        auto it = _bs.find(dest);
        return it != _bs.end() ? it->second : _ref->backwardStar(dest);
    }
    const std::vector<V> &vertices() const override { return _vertices; }
    V edges() const override { return _edges; }
    const std::set<V> &borders() const override { return _borders; }
    bool isBorder(V vertex) const override { return _borders.count(vertex) > 0; }
    std::shared_ptr<const IGraph<V, W>> referenceGraph() const override { return _ref; }

private:
    const std::shared_ptr<const IGraph<V, W>> _ref;
    const std::vector<V> _vertices;
    V _edges;
    std::map<V, std::map<V, W>> _fs, _bs;
    std::set<V> _borders;
};

} // namespace fastbc

#endif
