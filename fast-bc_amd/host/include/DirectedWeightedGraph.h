/*
 * DirectedWeightedGraph (DirectedWeightedGraph.h:13-182) with the same
 * constructors and accessors.  The loaded graph is held as a CSR (what the
 * GPU path uploads); the std::map stars of the reference's interface are built
 * on first use.  Loader semantics: fbchost::load_edge_list.
 */
#ifndef FASTBC_AMD_DIRECTEDWEIGHTEDGRAPH_H
#define FASTBC_AMD_DIRECTEDWEIGHTEDGRAPH_H

#include "IDegreeGraph.h"
#include "../fbc_host.h"

#include <istream>
#include <map>
#include <mutex>
#include <vector>

namespace fastbc {

template <typename V, typename W>
class DirectedWeightedGraph : public IDegreeGraph<V, W> {
public:
    explicit DirectedWeightedGraph(std::istream &inputTextGraph)
        : _g(fbchost::load_edge_list(inputTextGraph)), _csr(true)
    {
        _edges = (V)_g.edges;
        _totalWeight = (W)_g.total_weight;
        _in.assign(_g.in_wdeg.begin(), _g.in_wdeg.end());
        _out.assign(_g.out_wdeg.begin(), _g.out_wdeg.end());
        initVertices();
    }

    DirectedWeightedGraph() : _csr(false), _maps(true) {}

    W edge(V src, V dest) const override
    {
        const auto &f = forwardStar(src);
        auto it = f.find(dest);
        return it == f.end() ? W(0) : it->second;
    }
    const std::map<V, W> &forwardStar(V src) const override
    {
        buildMaps();
        return _fs[src];
    }
    const std::map<V, W> &backwardStar(V dest) const override
    {
        buildMaps();
        return _bs[dest];
    }
    const std::vector<V> &vertices() const override { return _vertices; }
    V edges() const override { return _edges; }

    /* DirectedWeightedGraph.h:135-170 */
    void addEdge(V from, V to, W weight) override
    {
        buildMaps();
        std::lock_guard<std::mutex> lk(_mu);
        if ((V)_fs.size() <= from) _fs.resize((size_t)from + 1);
        if ((V)_bs.size() <= to) _bs.resize((size_t)to + 1);
        const V s = to > from ? to : from;
        if ((V)_in.size() <= s) _in.resize((size_t)s + 1, 0);
        if ((V)_out.size() <= s) _out.resize((size_t)s + 1, 0);
        auto f = _fs[from].find(to);
        if (f != _fs[from].end()) f->second += weight;
        else {
            _fs[from][to] = weight;
            ++_edges;
        }
        auto b = _bs[to].find(from);
        if (b != _bs[to].end()) b->second += weight;
        else _bs[to][from] = weight;
        _totalWeight += weight;
        _in[to] += weight;
        _out[from] += weight;
        _csr = false;
    }
    void initVertices() override
    {
        const size_t n = _maps ? _fs.size() : (size_t)_g.n;
        if (_maps) _bs.resize(n);
        _vertices.resize(n);
        for (size_t v = 0; v < n; ++v) _vertices[v] = (V)v;
    }
    W totalWeight() const override { return _totalWeight; }
    W inWeightedDegree(V v) const override { return _in[v]; }
    W outWeightedDegree(V v) const override { return _out[v]; }

    /* The CSR the GPU path and the Louvain engine read. */
    const fbchost::Graph &hostGraph() const
    {
        std::lock_guard<std::mutex> lk(_mu);
        if (!_csr) {
            std::vector<int> s, d;
            std::vector<double> w;
            for (size_t u = 0; u < _fs.size(); ++u)
                for (const auto &e : _fs[u]) {
                    s.push_back((int)u);
                    d.push_back((int)e.first);
                    w.push_back((double)e.second);
                }
            _g = fbchost::from_arcs_n((int)_vertices.size(), s.size(), s.data(), d.data(), w.data());
            _csr = true;
        }
        return _g;
    }

private:
    void buildMaps() const
    {
        std::lock_guard<std::mutex> lk(_mu);
        if (_maps) return;
        _fs.assign((size_t)_g.n, {});
        _bs.assign((size_t)_g.n, {});
        for (int u = 0; u < _g.n; ++u) {
            for (uint64_t e = _g.rp[u]; e < _g.rp[u + 1]; ++e) _fs[u].emplace_hint(_fs[u].end(), (V)_g.col[e], (W)_g.w[e]);
            for (uint64_t e = _g.rrp[u]; e < _g.rrp[u + 1]; ++e) _bs[u].emplace_hint(_bs[u].end(), (V)_g.rcol[e], (W)_g.rw[e]);
        }
        _maps = true;
    }

    mutable fbchost::Graph _g;
    mutable bool _csr = false;
    mutable bool _maps = false;
    mutable std::vector<std::map<V, W>> _fs, _bs;
    mutable std::mutex _mu;
    std::vector<V> _vertices;
    V _edges = 0;
    W _totalWeight = 0;
    std::vector<W> _in, _out;
};

} // namespace fastbc

#endif
