/*
 * Multi-instance directed Louvain, bit-identical to
 * louvain/LouvainGraphPartition.h:84-130 + louvain/Partition.h + louvain/LouvainGraph.h.
 *
 * What must stay alike, and why (the partition feeds every later step, so a
 * one-ulp difference in a modularity gain can change the result):
 *   - one std::mt19937 per seed, ascending seeds, libstdc++ std::shuffle of
 *     0..n-1 per level and instance (Partition.h:288-295);
 *   - gain = wic/m - (wout/m)*(winctot/m) - (win/m)*(woutctot/m) with a strict
 *     '>' over neighbour communities in discovery order (Partition.h:122-140,
 *     :330-345), in-links first then out-links (:154-197);
 *   - insert() adds the node's OUT-degree to winctot (Partition.h:118-119);
 *   - a pass repeats while moves > 0 and the modularity rose by more than the
 *     precision (:364); levels repeat while the best instance moved (:125);
 *   - the best instance is the first with the largest modularity (:108-111);
 *   - aggregated link weights are summed in node order, then link order, and
 *     emitted by ascending community (Partition.h:228-282).
 * The data layout is flat arrays; the per-community std::map of
 * partition2graph is replaced by a stable sort with the same summation order;
 * weighted degrees, which the reference re-sums on every call with the same
 * order, are summed once per level.
 */
#include "fbc_host.h"

#include <algorithm>
#include <numeric>
#include <random>
#include <set>

namespace fbchost {

namespace {

struct LGraph {
    unsigned nodes = 0;
    double total_weight = 0;
    std::vector<uint64_t> in_end, out_end;     /* cumulative link counts */
    std::vector<int> in_nb, out_nb;
    std::vector<double> in_w, out_w;
    std::vector<double> win, wout, self;       /* per node, reference sums */

    uint64_t in_begin(int v) const { return v == 0 ? 0 : in_end[v - 1]; }
    uint64_t out_begin(int v) const { return v == 0 ? 0 : out_end[v - 1]; }

    void derive()
    {
        win.assign(nodes, 0.0);
        wout.assign(nodes, 0.0);
        self.assign(nodes, 0.0);
        for (unsigned v = 0; v < nodes; ++v) {
            double r = 0;
            for (uint64_t i = in_begin((int)v); i < in_end[v]; ++i) r += in_w.empty() ? 1.0 : in_w[i];
            win[v] = in_w.empty() ? (double)(in_end[v] - in_begin((int)v)) : r;
            r = 0;
            for (uint64_t i = out_begin((int)v); i < out_end[v]; ++i) r += out_w.empty() ? 1.0 : out_w[i];
            wout[v] = out_w.empty() ? (double)(out_end[v] - out_begin((int)v)) : r;
            self[v] = 0.0;
            for (uint64_t i = out_begin((int)v); i < out_end[v]; ++i)
                if (out_nb[i] == (int)v) {
                    self[v] = out_w.empty() ? 1.0 : out_w[i];
                    break;
                }
        }
    }
};

LGraph from_graph(const Graph &g)
{
    LGraph L;
    L.nodes = (unsigned)g.n;
    L.in_end.resize(g.n);
    L.out_end.resize(g.n);
    L.in_nb.assign(g.rcol.begin(), g.rcol.end());
    L.in_w.assign(g.rw.begin(), g.rw.end());
    L.out_nb.assign(g.col.begin(), g.col.end());
    L.out_w.assign(g.w.begin(), g.w.end());
    L.total_weight = 0;
    for (int i = 0; i < g.n; ++i) {
        for (uint64_t e = g.rrp[i]; e < g.rrp[i + 1]; ++e) L.total_weight += g.rw[e];
        L.in_end[i] = g.rrp[i + 1];
        L.out_end[i] = g.rp[i + 1];
    }
    L.derive();
    return L;
}

struct Part {
    const LGraph *g;
    int size;
    double min_mod;
    std::vector<double> nw;
    std::vector<unsigned> npos;
    unsigned nlast = 0;
    std::vector<int> n2c;
    std::vector<double> woutc, winc, woutctot, winctot;

    Part(const LGraph &gr, double minm) : g(&gr), size((int)gr.nodes), min_mod(minm)
    {
        nw.assign(size, -1);
        npos.assign(size, 0);
        n2c.resize(size);
        woutc.resize(size);
        winc.resize(size);
        woutctot.resize(size);
        winctot.resize(size);
        for (int i = 0; i < size; ++i) {
            n2c[i] = i;
            woutc[i] = winc[i] = gr.self[i];
            woutctot[i] = gr.wout[i];
            winctot[i] = gr.win[i];
        }
    }

    void remove(int node)
    {
        woutc[node] = 0;
        winc[node] = 0;
        woutctot[n2c[node]] -= g->wout[node];
        winctot[n2c[node]] -= g->win[node];
        n2c[node] = -1;
    }

    void insert(int node, int comm)
    {
        n2c[node] = comm;
        for (uint64_t i = g->out_begin(node); i < g->out_end[node]; ++i)
            if (n2c[g->out_nb[i]] == comm) woutc[node] += g->out_w[i];
        for (uint64_t i = g->in_begin(node); i < g->in_end[node]; ++i)
            if (n2c[g->in_nb[i]] == comm) winc[node] += g->in_w[i];
        woutctot[comm] += g->wout[node];
        winctot[comm] += g->wout[node];
    }

    double gain(int node, int comm, double wic) const
    {
        const double woutn = g->wout[node], winn = g->win[node];
        const double wc_in = winctot[comm], wc_out = woutctot[comm];
        const double m = g->total_weight;
        return (wic / m - (woutn / m) * (wc_in / m) - (winn / m) * (wc_out / m));
    }

    double modularity() const
    {
        double q = 0.;
        const double m = g->total_weight;
        for (int i = 0; i < size; ++i)
            if (g->wout[i] > 0) q += woutc[i] / m - (g->wout[i] / m) * (winctot[n2c[i]] / m);
        return q;
    }

    void neigh_comm(int node)
    {
        for (unsigned i = 0; i < nlast; ++i) nw[npos[i]] = -1;
        nlast = 0;
        npos[0] = (unsigned)n2c[node];
        nw[npos[0]] = 0;
        nlast = 1;
        const bool win_unit = g->in_w.empty(), wout_unit = g->out_w.empty();
        for (uint64_t i = g->in_begin(node); i < g->in_end[node]; ++i) {
            const int nb = g->in_nb[i];
            if (nb == node) continue;
            const unsigned c = (unsigned)n2c[nb];
            if (nw[c] == -1) {
                nw[c] = 0.;
                npos[nlast++] = c;
            }
            nw[c] += win_unit ? 1. : g->in_w[i];
        }
        for (uint64_t i = g->out_begin(node); i < g->out_end[node]; ++i) {
            const int nb = g->out_nb[i];
            if (nb == node) continue;
            const unsigned c = (unsigned)n2c[nb];
            if (nw[c] == -1) {
                nw[c] = 0.;
                npos[nlast++] = c;
            }
            nw[c] += wout_unit ? 1. : g->out_w[i];
        }
    }

    bool one_level(std::mt19937 &eng)
    {
        std::vector<int> order((size_t)size);
        std::iota(order.begin(), order.end(), 0);
        std::shuffle(order.begin(), order.end(), eng);
        bool improvement = false;
        int moves;
        double new_mod = modularity(), cur_mod;
        do {
            cur_mod = new_mod;
            moves = 0;
            for (int t = 0; t < size; ++t) {
                const int node = order[t];
                const int node_comm = n2c[node];
                neigh_comm(node);
                remove(node);
                int best = node_comm;
                double best_inc = 0.;
                for (unsigned i = 0; i < nlast; ++i) {
                    const double inc = gain(node, (int)npos[i], nw[npos[i]]);
                    if (inc > best_inc) {
                        best = (int)npos[i];
                        best_inc = inc;
                    }
                }
                insert(node, best);
                if (best != node_comm) ++moves;
            }
            new_mod = modularity();
            if (moves > 0) improvement = true;
        } while (moves > 0 && new_mod - cur_mod > min_mod);
        return improvement;
    }

    /* partition2graph: link weights per (community, neighbour community)
     * summed in node order then link order, emitted by ascending key */
    LGraph aggregate(std::vector<int> &renumber) const
    {
        renumber.assign((size_t)size, -1);
        for (int node = 0; node < size; ++node) renumber[n2c[node]]++;
        int fin = 0;
        for (int i = 0; i < size; ++i)
            if (renumber[i] != -1) renumber[i] = fin++;
        std::vector<int> cptr((size_t)fin + 1, 0), members((size_t)size);
        for (int node = 0; node < size; ++node) cptr[renumber[n2c[node]] + 1]++;
        for (int c = 0; c < fin; ++c) cptr[c + 1] += cptr[c];
        {
            std::vector<int> f(cptr.begin(), cptr.end() - 1);
            for (int node = 0; node < size; ++node) members[f[renumber[n2c[node]]]++] = node;
        }
        LGraph g2;
        g2.nodes = (unsigned)fin;
        g2.total_weight = 0;
        g2.in_end.resize(fin);
        g2.out_end.resize(fin);
        std::vector<std::pair<int, double>> inl, outl;
        std::vector<size_t> idx;
        auto emit = [&](std::vector<std::pair<int, double>> &lst, std::vector<int> &nb,
                        std::vector<double> &wt, bool count_total) {
            idx.resize(lst.size());
            std::iota(idx.begin(), idx.end(), (size_t)0);
            std::stable_sort(idx.begin(), idx.end(),
                             [&](size_t a, size_t b) { return lst[a].first < lst[b].first; });
            size_t k = 0;
            while (k < idx.size()) {
                const int key = lst[idx[k]].first;
                double s = lst[idx[k]].second;
                size_t j = k + 1;
                while (j < idx.size() && lst[idx[j]].first == key) s += lst[idx[j++]].second;
                if (count_total) g2.total_weight += s;
                nb.push_back(key);
                wt.push_back(s);
                k = j;
            }
        };
        const bool win_unit = g->in_w.empty(), wout_unit = g->out_w.empty();
        for (int c = 0; c < fin; ++c) {
            inl.clear();
            outl.clear();
            for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                const int node = members[k];
                for (uint64_t i = g->in_begin(node); i < g->in_end[node]; ++i)
                    inl.push_back({renumber[n2c[g->in_nb[i]]], win_unit ? 1. : g->in_w[i]});
                for (uint64_t i = g->out_begin(node); i < g->out_end[node]; ++i)
                    outl.push_back({renumber[n2c[g->out_nb[i]]], wout_unit ? 1. : g->out_w[i]});
            }
            emit(inl, g2.in_nb, g2.in_w, true);
            emit(outl, g2.out_nb, g2.out_w, false);
            g2.in_end[c] = g2.in_nb.size();
            g2.out_end[c] = g2.out_nb.size();
        }
        g2.derive();
        return g2;
    }
};

} // namespace

std::vector<std::vector<int>> louvain(const Graph &g, const std::vector<uint64_t> &seed_list,
                                      double precision)
{
    std::set<uint64_t> seeds(seed_list.begin(), seed_list.end());
    std::vector<std::mt19937> eng;
    for (uint64_t s : seeds) eng.push_back(std::mt19937((std::mt19937::result_type)s));
    const int P = (int)eng.size();
    if (P == 0 || g.n == 0) return {};
    LGraph lg = from_graph(g);
    std::vector<int> n2c(lg.nodes);
    std::iota(n2c.begin(), n2c.end(), 0);
    bool improvement;
    do {
        std::vector<Part> parts;
        parts.reserve(P);
        for (int i = 0; i < P; ++i) parts.emplace_back(lg, precision);
        std::vector<char> imp((size_t)P);
        std::vector<double> mods((size_t)P);
#pragma omp parallel for schedule(static, 1)
        for (int i = 0; i < P; ++i) {
            imp[i] = parts[i].one_level(eng[i]);
            mods[i] = parts[i].modularity();
        }
        int best = 0;
        for (int i = 1; i < P; ++i)
            if (mods[i] > mods[best]) best = i;
        improvement = imp[best] != 0;
        std::vector<int> renumber;
        LGraph next = parts[best].aggregate(renumber);
        /* renumber_communities (LouvainGraphPartition.h:27-46) */
        const std::vector<int> &lvl = parts[best].n2c;
        for (size_t i = 0; i < n2c.size(); ++i) n2c[i] = renumber[lvl[n2c[i]]];
        lg = std::move(next);
    } while (improvement);
    int mx = 0;
    for (int c : n2c) mx = std::max(mx, c);
    std::vector<std::vector<int>> r((size_t)mx + 1);
    for (size_t i = 0; i < n2c.size(); ++i) r[n2c[i]].push_back((int)i);
    return r;
}

} // namespace fbchost
