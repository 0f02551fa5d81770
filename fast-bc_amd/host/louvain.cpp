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
 *
 * A pass is a chain of dependent random reads (the shuffled node, its links,
 * its neighbours' communities, their sums), so each level packs a node's
 * in-links then out-links into one run of (neighbour, weight) records shared
 * by the instances, keeps a community's two sums side by side, and prefetches
 * that chain a few nodes ahead along the shuffled order.  Aggregation reduces
 * the communities in parallel (each one's sums keep their order) and adds the
 * level's total in the reference's order afterwards.
 */
#include "fbc_host.h"

#include <sys/mman.h>

#include <algorithm>
#include <cstdlib>
#include <new>
#include <numeric>
#include <random>
#include <set>

namespace fbchost {

namespace {

/* Arrays of a pass (link runs, node records, community sums, the node order)
 * are read at random over hundreds of MB, so with 4 KiB pages nearly every
 * step of the chain also misses the TLB.  Arrays of 2 MiB or more are placed
 * on 2 MiB boundaries and offered to transparent huge pages
 * (MADV_HUGEPAGE; where the kernel declines, they stay ordinary pages). */
template <class T>
struct HugeAlloc {
    using value_type = T;
    static constexpr size_t HUGE = (size_t)2 << 20;
    HugeAlloc() = default;
    template <class U>
    HugeAlloc(const HugeAlloc<U> &) {}
    T *allocate(size_t n)
    {
        const size_t bytes = n * sizeof(T);
        if (bytes < HUGE) return static_cast<T *>(::operator new(bytes));
        const size_t sz = (bytes + HUGE - 1) / HUGE * HUGE;
        void *p = std::aligned_alloc(HUGE, sz);
        if (!p) throw std::bad_alloc();
        (void)madvise(p, sz, MADV_HUGEPAGE);
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n)
    {
        if (n * sizeof(T) < HUGE)
            ::operator delete(p);
        else
            std::free(p);
    }
};
template <class T, class U>
bool operator==(const HugeAlloc<T> &, const HugeAlloc<U> &) { return true; }
template <class T, class U>
bool operator!=(const HugeAlloc<T> &, const HugeAlloc<U> &) { return false; }
template <class T>
using hvec = std::vector<T, HugeAlloc<T>>;

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
#pragma omp parallel for schedule(static, 4096)
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

/* One level's links in the order the pass reads them: per node its in-links
 * then its out-links (weights as the reference reads them, 1 for an
 * unweighted graph), so a node's neighbourhood is one contiguous run. */
struct Link {
    int nb;
    double w;
};
struct NodeRec {
    uint64_t a0;          /* first link */
    uint32_t nin, nout;
    double wout, win;     /* weighted out / in degree */
};
struct CommTot {
    double out, in;       /* woutctot, winctot */
};

/* The packed links of one level, shared by its instances. */
struct LevelLinks {
    hvec<NodeRec> rec;
    hvec<Link> adj;

    explicit LevelLinks(const LGraph &gr)
    {
        const int n = (int)gr.nodes;
        rec.resize(n);
        adj.resize(gr.in_nb.size() + gr.out_nb.size());
        const bool win_unit = gr.in_w.empty(), wout_unit = gr.out_w.empty();
#pragma omp parallel for schedule(static, 4096)
        for (int i = 0; i < n; ++i) {
            NodeRec &r = rec[i];
            r.a0 = gr.in_begin(i) + gr.out_begin(i);
            r.nin = (uint32_t)(gr.in_end[i] - gr.in_begin(i));
            r.nout = (uint32_t)(gr.out_end[i] - gr.out_begin(i));
            r.wout = gr.wout[i];
            r.win = gr.win[i];
            uint64_t a = r.a0;
            for (uint64_t k = gr.in_begin(i); k < gr.in_end[i]; ++k) adj[a++] = {gr.in_nb[k], win_unit ? 1. : gr.in_w[k]};
            for (uint64_t k = gr.out_begin(i); k < gr.out_end[i]; ++k)
                adj[a++] = {gr.out_nb[k], wout_unit ? 1. : gr.out_w[k]};
        }
    }
};

struct Part {
    const LGraph *g;
    int size;
    double min_mod;
    const hvec<NodeRec> &rec;
    const hvec<Link> &adj;
    hvec<double> nw;
    hvec<unsigned> npos;
    unsigned nlast = 0;
    hvec<int> n2c;
    hvec<double> woutc;
    hvec<CommTot> tot;

    Part(const LGraph &gr, const LevelLinks &ll, double minm)
        : g(&gr), size((int)gr.nodes), min_mod(minm), rec(ll.rec), adj(ll.adj)
    {
        nw.assign(size, -1);
        npos.assign(size, 0);
        n2c.resize(size);
        woutc.resize(size);
        tot.resize(size);
        for (int i = 0; i < size; ++i) {
            n2c[i] = i;
            woutc[i] = gr.self[i];
            tot[i] = {gr.wout[i], gr.win[i]};
        }
    }

    /* winc (the in-link twin of woutc) is kept by the reference but read by
     * neither the gain nor the modularity, so it is not kept here. */
    void remove(int node)
    {
        woutc[node] = 0;
        tot[n2c[node]].out -= rec[node].wout;
        tot[n2c[node]].in -= rec[node].win;
        n2c[node] = -1;
    }

    void insert(int node, int comm)
    {
        n2c[node] = comm;
        const NodeRec &r = rec[node];
        const Link *out = adj.data() + r.a0 + r.nin;
        for (uint32_t i = 0; i < r.nout; ++i)
            if (n2c[out[i].nb] == comm) woutc[node] += out[i].w;
        tot[comm].out += r.wout;
        tot[comm].in += r.wout;
    }

    double gain(int node, int comm, double wic) const
    {
        const double woutn = rec[node].wout, winn = rec[node].win;
        const double wc_in = tot[comm].in, wc_out = tot[comm].out;
        const double m = g->total_weight;
        return (wic / m - (woutn / m) * (wc_in / m) - (winn / m) * (wc_out / m));
    }

    double modularity() const
    {
        double q = 0.;
        const double m = g->total_weight;
        for (int i = 0; i < size; ++i)
            if (rec[i].wout > 0) q += woutc[i] / m - (rec[i].wout / m) * (tot[n2c[i]].in / m);
        return q;
    }

    void neigh_comm(int node)
    {
        for (unsigned i = 0; i < nlast; ++i) nw[npos[i]] = -1;
        nlast = 0;
        npos[0] = (unsigned)n2c[node];
        nw[npos[0]] = 0;
        nlast = 1;
        const NodeRec &r = rec[node];
        const Link *l = adj.data() + r.a0;
        const uint32_t nl = r.nin + r.nout;
        for (uint32_t i = 0; i < nl; ++i) {
            const int nb = l[i].nb;
            if (nb == node) continue;
            const unsigned c = (unsigned)n2c[nb];
            if (nw[c] == -1) {
                nw[c] = 0.;
                npos[nlast++] = c;
            }
            nw[c] += l[i].w;
        }
    }

    /* Prefetches for the node handled 'ahead' steps later: its record, then
     * its links, then its neighbours' communities, then their sums.  A
     * prefetch never changes a value; a community that changes before the
     * node is reached only costs a miss. */
    void prefetch(const hvec<int> &order, int t) const
    {
        constexpr int A1 = 24, A2 = 12, A3 = 6, A4 = 3;
        if (t + A1 < size) __builtin_prefetch(&rec[order[t + A1]]);
        if (t + A2 < size) {
            const int v = order[t + A2];
            const NodeRec &r = rec[v];
            __builtin_prefetch(&adj[r.a0]);
            if (r.nin + r.nout > 4) __builtin_prefetch(&adj[r.a0 + r.nin + r.nout - 1]);
            __builtin_prefetch(&n2c[v]);
            __builtin_prefetch(&woutc[v]);
        }
        if (t + A3 < size) {
            const NodeRec &r = rec[order[t + A3]];
            const Link *l = adj.data() + r.a0;
            for (uint32_t i = 0; i < r.nin + r.nout; ++i) __builtin_prefetch(&n2c[l[i].nb]);
        }
        if (t + A4 < size) {
            const NodeRec &r = rec[order[t + A4]];
            const Link *l = adj.data() + r.a0;
            for (uint32_t i = 0; i < r.nin + r.nout; ++i) {
                const int c = n2c[l[i].nb];
                if (c >= 0) {
                    __builtin_prefetch(&nw[c]);
                    __builtin_prefetch(&tot[c]);
                }
            }
        }
    }

    bool one_level(std::mt19937 &eng)
    {
        hvec<int> order((size_t)size);
        std::iota(order.begin(), order.end(), 0);
        std::shuffle(order.begin(), order.end(), eng);
        bool improvement = false;
        int moves;
        double new_mod = modularity(), cur_mod;
        do {
            cur_mod = new_mod;
            moves = 0;
            for (int t = 0; t < size; ++t) {
                prefetch(order, t);
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
        g2.in_end.resize(fin);
        g2.out_end.resize(fin);
        /* each community's reduced lists go to a slot bounded by its members'
         * link count; communities are independent, so they run in parallel */
        std::vector<uint64_t> ibound((size_t)fin + 1, 0), obound((size_t)fin + 1, 0);
        for (int c = 0; c < fin; ++c) {
            uint64_t bi = 0, bo = 0;
            for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                const int node = members[k];
                bi += g->in_end[node] - g->in_begin(node);
                bo += g->out_end[node] - g->out_begin(node);
            }
            ibound[c + 1] = ibound[c] + bi;
            obound[c + 1] = obound[c] + bo;
        }
        std::vector<std::pair<int, double>> ired(ibound[fin]), ored(obound[fin]);
        std::vector<uint64_t> icount(fin), ocount(fin);
        const bool win_unit = g->in_w.empty(), wout_unit = g->out_w.empty();
        auto by_key = [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; };
        /* stable by key: each key's weights are summed in node order, then
         * link order, as the reference's per-community map accumulates them */
        auto reduce = [&](std::vector<std::pair<int, double>> &lst, std::pair<int, double> *out) {
            std::stable_sort(lst.begin(), lst.end(), by_key);
            uint64_t o = 0;
            size_t k = 0;
            while (k < lst.size()) {
                const int key = lst[k].first;
                double sum = lst[k].second;
                size_t j2 = k + 1;
                while (j2 < lst.size() && lst[j2].first == key) sum += lst[j2++].second;
                out[o++] = {key, sum};
                k = j2;
            }
            return o;
        };
#pragma omp parallel
        {
            std::vector<std::pair<int, double>> inl, outl;
#pragma omp for schedule(dynamic, 256)
            for (int c = 0; c < fin; ++c) {
                inl.clear();
                outl.clear();
                for (int k = cptr[c]; k < cptr[c + 1]; ++k) {
                    const int node = members[k];
                    for (uint64_t i2 = g->in_begin(node); i2 < g->in_end[node]; ++i2)
                        inl.push_back({renumber[n2c[g->in_nb[i2]]], win_unit ? 1. : g->in_w[i2]});
                    for (uint64_t i2 = g->out_begin(node); i2 < g->out_end[node]; ++i2)
                        outl.push_back({renumber[n2c[g->out_nb[i2]]], wout_unit ? 1. : g->out_w[i2]});
                }
                icount[c] = reduce(inl, ired.data() + ibound[c]);
                ocount[c] = reduce(outl, ored.data() + obound[c]);
            }
        }
        uint64_t ni = 0, no = 0;
        for (int c = 0; c < fin; ++c) {
            ni += icount[c];
            no += ocount[c];
            g2.in_end[c] = ni;
            g2.out_end[c] = no;
        }
        g2.in_nb.resize(ni);
        g2.in_w.resize(ni);
        g2.out_nb.resize(no);
        g2.out_w.resize(no);
#pragma omp parallel for schedule(static, 1024)
        for (int c = 0; c < fin; ++c) {
            const uint64_t di = c ? g2.in_end[c - 1] : 0, dout = c ? g2.out_end[c - 1] : 0;
            for (uint64_t k = 0; k < icount[c]; ++k) {
                g2.in_nb[di + k] = ired[ibound[c] + k].first;
                g2.in_w[di + k] = ired[ibound[c] + k].second;
            }
            for (uint64_t k = 0; k < ocount[c]; ++k) {
                g2.out_nb[dout + k] = ored[obound[c] + k].first;
                g2.out_w[dout + k] = ored[obound[c] + k].second;
            }
        }
        /* the total is summed over the reduced in-links in (community, key)
         * order, as they are emitted */
        g2.total_weight = 0;
        for (uint64_t k = 0; k < ni; ++k) g2.total_weight += g2.in_w[k];
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
        const LevelLinks links(lg);
        std::vector<Part> parts;
        parts.reserve(P);
        for (int i = 0; i < P; ++i) parts.emplace_back(lg, links, precision);
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
        const hvec<int> &lvl = parts[best].n2c;
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
