/*
 * fbc oracle -- TEST INFRASTRUCTURE ONLY (see fbc_oracle.h).
 *
 * A plain restatement of the reference algorithms, written for clarity and
 * fidelity, not speed: ordered-set Dijkstra, explicit predecessor lists and
 * the reference's floating-point operation order.  Build with
 * -ffp-contract=off.
 */
#include "fbc_oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <set>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::vector<std::vector<std::pair<int, double>>> Adj;

struct orc_graph {
    int n = 0;
    long long edges = 0;
    Adj fwd, bwd;   /* each star sorted by neighbour id, like std::map */
};

namespace {

void set_err(char *err, int errlen, const std::string &msg)
{
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), (size_t)errlen - 1);
        err[errlen - 1] = '\0';
    }
}

/* addEdge semantics, DirectedWeightedGraph.h:135-170: a repeated arc adds its
 * weight to the existing one and is counted once. */
struct GraphBuilder {
    std::vector<std::map<int, double>> fwd, bwd;
    long long edges = 0;

    void add(int from, int to, double w)
    {
        if ((int)fwd.size() <= from) fwd.resize((size_t)from + 1);
        if ((int)bwd.size() <= to) bwd.resize((size_t)to + 1);
        auto f = fwd[from].find(to);
        if (f != fwd[from].end()) {
            f->second += w;
        } else {
            fwd[from][to] = w;
            ++edges;
        }
        auto b = bwd[to].find(from);
        if (b != bwd[to].end()) b->second += w;
        else bwd[to][from] = w;
    }

    /* DirectedWeightedGraph.h:92-95: the vertex count is the forward-star
     * count (1 + largest source id) and backward stars are cut to it. */
    orc_graph *finish(char *err, int errlen)
    {
        const int n = (int)fwd.size();
        for (int v = 0; v < n; ++v)
            for (auto &e : fwd[v])
                if (e.first >= n) {
                    set_err(err, errlen, "arc target " + std::to_string(e.first) +
                                         " has no out-arc (vertex count is 1 + largest source id)");
                    return nullptr;
                }
        orc_graph *g = new orc_graph;
        g->n = n;
        g->edges = edges;
        g->fwd.resize(n);
        g->bwd.resize(n);
        for (int v = 0; v < n; ++v) {
            g->fwd[v].assign(fwd[v].begin(), fwd[v].end());
            if (v < (int)bwd.size()) g->bwd[v].assign(bwd[v].begin(), bwd[v].end());
        }
        return g;
    }
};

/* (int) of a double as x86-64 cvttsd2si does it: out-of-range -> INT_MIN.
 * The reference stores (int)sigma in VertexInfo (DijkstraClusterEvaluator.h:173). */
int to_int_x86(double x)
{
    if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT_MIN;
    return (int)x;
}

/* ---- Dijkstra with shortest-path counting --------------------------------
 * DijkstraSSBrandesBC.h:83-147 (== ExactBrandesBC.h:97-161 ==
 * DijkstraClusterEvaluator.h:102-164): a std::set ordered by (dist, id),
 * erase-before-update, predecessor lists appended in settle order, FP64 sigma.
 * 'key' gives the id used to break distance ties (the global vertex id). */
struct SingleSource {
    std::vector<double> dist, sigma;
    std::vector<int> order;
    std::vector<std::vector<int>> preds;
};

void dijkstra(const Adj &adj, const std::vector<int> *key, int src, SingleSource &s)
{
    const size_t n = adj.size();
    s.dist.assign(n, DBL_MAX);
    s.sigma.assign(n, 0.0);
    s.order.clear();
    s.preds.assign(n, std::vector<int>());
    std::vector<double> &dist = s.dist;
    auto id = [key](int v) { return key ? (*key)[v] : v; };
    auto cmp = [&dist, &id](int a, int b) {
        if (dist[a] == dist[b]) return id(a) < id(b);
        return dist[a] < dist[b];
    };
    std::set<int, decltype(cmp)> queue(cmp);
    s.sigma[src] = 1;
    dist[src] = 0;
    queue.insert(src);
    while (!queue.empty()) {
        int v = *queue.begin();
        queue.erase(queue.begin());
        s.order.push_back(v);
        for (const auto &arc : adj[v]) {
            int w = arc.first;
            double nd = dist[v] + arc.second;
            if (nd < dist[w]) {
                queue.erase(w);
                dist[w] = nd;
                queue.insert(w);
                s.preds[w].clear();
                s.sigma[w] = 0;
            }
            if (nd == dist[w]) {
                s.preds[w].push_back(v);
                s.sigma[w] += s.sigma[v];
            }
        }
    }
}

/* Backward dependency sweep, DijkstraSSBrandesBC.h:61-78: vertices in reverse
 * settle order; for each predecessor v of w (in list order)
 * delta[v] += sigma[v] / sigma[w] * (1 + delta[w]).  'visit(w, delta_w)' is
 * called for every w != src after its predecessors were updated. */
template <typename Visit>
void backward(const SingleSource &s, int src, std::vector<double> &delta, Visit visit)
{
    for (size_t i = s.order.size(); i-- > 0;) {
        int w = s.order[i];
        for (int v : s.preds[w]) {
            double c = s.sigma[v] / s.sigma[w] * (1.0 + delta[w]);
            delta[v] += c;
        }
        if (w != src) visit(w, delta[w]);
    }
}

/* libgomp's schedule(static) split of [0, n) for thread tid of nth. */
void static_chunk(long long n, int nth, int tid, long long &b, long long &e)
{
    long long q = n / nth, t = n % nth;
    if (tid < t) { ++q; t = 0; }
    b = q * tid + t;
    e = b + q;
}

/* ---- SubGraph (SubGraph.h:59-148) ----------------------------------------- */
struct Cluster {
    std::vector<int> verts;     /* as given */
    std::vector<int> borders;   /* ascending (std::set) */
    std::vector<char> is_border;/* per local index */
    Adj local;                  /* forward stars restricted to the cluster, local ids */
    long long edges = 0;
};

void make_cluster(const orc_graph &g, const std::vector<int> &verts, Cluster &c)
{
    c.verts = verts;
    std::map<int, int> local;   /* global -> local */
    for (size_t i = 0; i < verts.size(); ++i) local[verts[i]] = (int)i;
    c.local.assign(verts.size(), {});
    c.is_border.assign(verts.size(), 0);
    c.edges = 0;
    std::set<int> borders;
    for (size_t i = 0; i < verts.size(); ++i) {
        int v = verts[i];
        long long outside = 0;
        for (const auto &e : g.fwd[v]) {
            auto it = local.find(e.first);
            if (it == local.end()) ++outside;
            else c.local[i].push_back({it->second, e.second});
        }
        bool border = outside > 0;
        c.edges += (long long)g.fwd[v].size() - outside;
        for (const auto &e : g.bwd[v])
            if (local.find(e.first) == local.end()) border = true;
        if (border) {
            borders.insert(v);
            c.is_border[i] = 1;
        }
    }
    c.borders.assign(borders.begin(), borders.end());
}

/* ---- cluster evaluation (DijkstraClusterEvaluator.h:50-100,166-177) -------- */
void evaluate_cluster(const Cluster &c, std::vector<double> &bc,
                      std::vector<double> &vi_len, std::vector<int> &vi_cnt)
{
    const size_t nv = c.verts.size(), nb = c.borders.size();
    std::vector<int> border_local(nb);
    {
        std::map<int, int> local;
        for (size_t i = 0; i < nv; ++i) local[c.verts[i]] = (int)i;
        for (size_t j = 0; j < nb; ++j) border_local[j] = local[c.borders[j]];
    }
    vi_len.assign(nv * nb, 0.0);
    vi_cnt.assign(nv * nb, 0);
    /* the OpenMP array reduction of the (one-thread) inner team: a private
     * zero copy, added to bc once at the end (DijkstraClusterEvaluator.h:66) */
    std::vector<double> priv(nv, 0.0), delta(nv, 0.0);
    SingleSource s;
    for (size_t si = 0; si < nv; ++si) {
        std::fill(delta.begin(), delta.end(), 0.0);
        dijkstra(c.local, &c.verts, (int)si, s);
        for (size_t j = 0; j < nb; ++j) {
            int b = border_local[j];
            vi_len[si * nb + j] = s.dist[b] != DBL_MAX ? s.dist[b] : 0;
            vi_cnt[si * nb + j] = to_int_x86(s.sigma[b]);
        }
        backward(s, (int)si, delta, [&](int w, double dw) { priv[w] += dw; });
    }
    for (size_t i = 0; i < nv; ++i) bc[c.verts[i]] += priv[i];
}

/* ---- VertexInfo arithmetic (brandes/VertexInfo.h) -------------------------- */

/* VertexInfo<V,W>::squaredDistance of two points (VertexInfo.h:293-307),
 * sequential order: per border the length term, then the count term.
 * pow(x, 2) is evaluated as x * x.  W = double is the instance the path uses;
 * W = float is exported only so the reference's float known answer
 * (test/brandes/VertexInfo.cpp:130) can check the order of operations. */
template <typename W, typename WB>
W sqdist_points_t(const W *la, const int *ca, const WB *lb, const int *cb, int nb)
{
    W sq = 0;
    for (int i = 0; i < nb; ++i) {
        W dl = la[i] - (W)lb[i];
        sq += dl * dl;
        int dc = (int)((uint32_t)ca[i] - (uint32_t)cb[i]);
        W dcd = (W)dc;
        sq += dcd * dcd;
    }
    return sq;
}

double sqdist_points(const double *la, const int *ca, const double *lb, const int *cb, int nb)
{
    return sqdist_points_t<double, double>(la, ca, lb, cb, nb);
}

/* VertexInfo<double,double> (a mean) against a point. */
double sqdist_mean(const double *ml, const double *mc, const double *lb, const int *cb, int nb)
{
    double sq = 0;
    for (int i = 0; i < nb; ++i) {
        double dl = ml[i] - lb[i];
        sq += dl * dl;
        double dc = mc[i] - (double)cb[i];
        sq += dc * dc;
    }
    return sq;
}

/* VertexInfo::normalize (VertexInfo.h:259-278): subtract the minimum length
 * (zeros of unreached borders included). */
void normalize_row(double *len, int nb)
{
    if (nb == 0) return;
    double mn = *std::min_element(len, len + nb);
    for (int i = 0; i < nb; ++i) len[i] -= mn;
}

/* VertexInfoPivotSelector::selectPivots (VertexInfoPivotSelector.h:29-127),
 * FASTBC_BRANDES_ENABLE_PIVOT_BORDER on (main.cpp:5). */
void select_pivots(const std::vector<int> &verts, int nb, std::vector<double> &vi_len,
                   const std::vector<int> &vi_cnt, const std::vector<double> &bc_of_row,
                   const std::vector<char> &row_is_border,
                   std::vector<int> &piv_rows, std::vector<int> &piv_weight)
{
    const size_t nv = verts.size();
    std::vector<int> reps;
    std::vector<std::vector<int>> members;
    for (size_t i = 0; i < nv; ++i) {
        normalize_row(&vi_len[i * nb], nb);
        bool found = false;
        for (size_t ci = 0; ci < reps.size() && !found; ++ci) {
            size_t r = (size_t)reps[ci];
            bool eq = true;
            for (int j = 0; j < nb && eq; ++j)
                eq = vi_cnt[r * nb + j] == vi_cnt[i * nb + j] &&
                     vi_len[r * nb + j] == vi_len[i * nb + j];
            if (eq) {
                members[ci].push_back((int)i);
                found = true;
            }
        }
        if (!found) {
            reps.push_back((int)i);
            members.push_back(std::vector<int>(1, (int)i));
        }
    }
    piv_rows.clear();
    piv_weight.clear();
    for (const auto &m : members) {
        size_t j = 0;
        while (j < m.size() && row_is_border[m[j]]) ++j;
        if (j == m.size()) {
            piv_rows.push_back(m[0]);
            piv_weight.push_back((int)m.size());
            continue;
        }
        int best = m[j];
        for (; j < m.size(); ++j) {
            int r = m[j];
            if (!row_is_border[r] && bc_of_row[r] < bc_of_row[best]) best = r;
        }
        piv_rows.push_back(best);
        piv_weight.push_back((int)m.size());
    }
}

/* PlusPlusKMeans::_initPlusPlus (kmeans/PlusPlusKMeans.h:172-207). */
std::vector<int> kmeans_init(int k, int P, int nb, const double *L, const int *C)
{
    std::vector<int> cent((size_t)k);
    cent[0] = 0;
    std::vector<double> cdist((size_t)P, 0.0);
    for (int i = 1; i < k; ++i) {
        int last = cent[(size_t)i - 1];
        double p = 1.0 / i;
        double q = 1.0 - p;
        int far = 0;
        for (int v = 0; v < P; ++v) {
            double d = sqdist_points(L + (size_t)last * nb, C + (size_t)last * nb,
                                     L + (size_t)v * nb, C + (size_t)v * nb, nb);
            cdist[v] = cdist[v] * q + d * p;
            if (cdist[v] > cdist[far]) far = v;
        }
        cent[i] = far;
    }
    return cent;
}

/* PlusPlusKMeans::computeCentroids (kmeans/PlusPlusKMeans.h:70-170) with
 * _centroidVariance (:209-229).  Points are row indices. */
void kmeans(int k, int P, int nb, const double *L, const int *C, const int *weights,
            double min_variance, long long max_iteration,
            std::vector<int> &out_cent, std::vector<int> &out_w)
{
    std::vector<int> next = kmeans_init(k, P, nb, L, C);
    std::vector<int> cent(next.size());
    const size_t K = next.size();
    std::vector<std::vector<double>> mlen(K, std::vector<double>(nb)), mcnt(K, std::vector<double>(nb));
    std::vector<std::vector<int>> members(K);
    long long iteration = 0;
    for (;;) {
        ++iteration;
        cent = next;
        for (size_t c = 0; c < K; ++c) {
            std::fill(mlen[c].begin(), mlen[c].end(), 0.0);
            std::fill(mcnt[c].begin(), mcnt[c].end(), 0.0);
            members[c].clear();
        }
        /* A centroid that repeats an earlier one's point has the same
         * distance to every point and a higher index, so the first strict
         * minimum (PlusPlusKMeans.h:98-112) never picks it: scanning each
         * point's first occurrence only gives the same assignment. */
        std::vector<size_t> firsts;
        {
            std::set<int> seen;
            for (size_t c = 0; c < K; ++c)
                if (seen.insert(cent[c]).second) firsts.push_back(c);
        }
        for (int v = 0; v < P; ++v) {
            const double *lv = L + (size_t)v * nb;
            const int *cv = C + (size_t)v * nb;
            int bestc = 0;
            double bestd = sqdist_points(L + (size_t)cent[0] * nb, C + (size_t)cent[0] * nb, lv, cv, nb);
            for (size_t f = 1; f < firsts.size(); ++f) {
                const size_t c = firsts[f];
                double d = sqdist_points(L + (size_t)cent[c] * nb, C + (size_t)cent[c] * nb, lv, cv, nb);
                if (d < bestd) { bestc = (int)c; bestd = d; }
            }
            for (int i = 0; i < nb; ++i) {
                mlen[bestc][i] += lv[i];
                mcnt[bestc][i] += cv[i];
            }
            members[bestc].push_back(v);
        }
        for (size_t c = 0; c < K; ++c) {
            if (members[c].empty()) continue;
            double sz = (double)members[c].size();
            for (int i = 0; i < nb; ++i) { mlen[c][i] /= sz; mcnt[c][i] /= sz; }
            int bestv = members[c][0];
            double bestd = sqdist_mean(mlen[c].data(), mcnt[c].data(), L + (size_t)bestv * nb,
                                       C + (size_t)bestv * nb, nb);
            for (size_t j = 1; j < members[c].size(); ++j) {
                int v = members[c][j];
                double d = sqdist_mean(mlen[c].data(), mcnt[c].data(), L + (size_t)v * nb,
                                       C + (size_t)v * nb, nb);
                if (d < bestd) { bestv = v; bestd = d; }
            }
            next[c] = bestv;
        }
        double var = 0;
        for (size_t c = 0; c < K; ++c) {
            double d = sqdist_points(L + (size_t)cent[c] * nb, C + (size_t)cent[c] * nb,
                                     L + (size_t)next[c] * nb, C + (size_t)next[c] * nb, nb);
            if (d > var) var = d;
        }
        if (!(var > min_variance && iteration <= max_iteration)) break;
    }
    out_cent = cent;
    out_w.assign(K, 0);
    for (size_t c = 0; c < K; ++c)
        for (int v : members[c]) out_w[c] += weights[v];
}

/* KMeansPivotSelector::selectPivots (brandes/KMeansPivotSelector.h:79-132):
 * k = max((int)(P * kfrac), 1), then drop repeated centroids keeping the
 * first occurrence and its weight. */
void kmeans_pivots(const std::vector<int> &exact_rows, const std::vector<int> &exact_w,
                   int nb, const std::vector<double> &vi_len, const std::vector<int> &vi_cnt,
                   double kfrac, std::vector<int> &piv_rows, std::vector<int> &piv_w)
{
    const int P = (int)exact_rows.size();
    int k = std::max((int)(exact_rows.size() * kfrac), 1);
    std::vector<double> L((size_t)P * nb);
    std::vector<int> C((size_t)P * nb);
    for (int p = 0; p < P; ++p)
        for (int j = 0; j < nb; ++j) {
            L[(size_t)p * nb + j] = vi_len[(size_t)exact_rows[p] * nb + j];
            C[(size_t)p * nb + j] = vi_cnt[(size_t)exact_rows[p] * nb + j];
        }
    std::vector<int> cent, cw;
    kmeans(k, P, nb, L.data(), C.data(), exact_w.data(), 0.0, 100, cent, cw);
    std::set<int> seen;
    piv_rows.clear();
    piv_w.clear();
    for (size_t i = 0; i < cent.size(); ++i) {
        int row = exact_rows[cent[i]];
        if (seen.insert(row).second) {
            piv_rows.push_back(row);
            piv_w.push_back(cw[i]);
        }
    }
}

/* ---- Louvain (louvain/LouvainGraph.h, Partition.h, LouvainGraphPartition.h) -- */

/* LouvainGraph: cumulative in/out degree arrays and flat link lists. */
struct LGraph {
    unsigned nodes = 0;
    unsigned long links = 0;
    double total_weight = 0;
    std::vector<unsigned long> in_end, out_end;    /* cumulative */
    std::vector<int> in_nb, out_nb;
    std::vector<double> in_w, out_w;

    unsigned long in_begin(int v) const { return v == 0 ? 0 : in_end[v - 1]; }
    unsigned long out_begin(int v) const { return v == 0 ? 0 : out_end[v - 1]; }
    unsigned in_count(int v) const { return (unsigned)(in_end[v] - in_begin(v)); }
    unsigned out_count(int v) const { return (unsigned)(out_end[v] - out_begin(v)); }

    /* LouvainGraph.h:119-132 */
    double selfloop(int v) const
    {
        for (unsigned i = 0; i < out_count(v); ++i)
            if (out_nb[out_begin(v) + i] == v) return out_w.size() ? out_w[out_begin(v) + i] : 1.0;
        return 0.0;
    }
    /* LouvainGraph.h:134-162: sums in link order */
    double in_degree_w(int v) const
    {
        if (in_w.empty()) return (double)in_count(v);
        double r = 0;
        for (unsigned i = 0; i < in_count(v); ++i) r += in_w[in_begin(v) + i];
        return r;
    }
    double out_degree_w(int v) const
    {
        if (out_w.empty()) return (double)out_count(v);
        double r = 0;
        for (unsigned i = 0; i < out_count(v); ++i) r += out_w[out_begin(v) + i];
        return r;
    }
};

/* LouvainGraph(graph), LouvainGraph.h:46-77 */
LGraph lgraph_from(const orc_graph &g)
{
    LGraph L;
    L.nodes = (unsigned)g.n;
    L.links = (unsigned long)g.edges;
    L.in_end.resize(g.n);
    L.out_end.resize(g.n);
    L.total_weight = 0;
    for (int i = 0; i < g.n; ++i) {
        for (const auto &e : g.bwd[i]) {
            L.in_nb.push_back(e.first);
            L.in_w.push_back(e.second);
            L.total_weight += e.second;
        }
        L.in_end[i] = L.in_nb.size();
        for (const auto &e : g.fwd[i]) {
            L.out_nb.push_back(e.first);
            L.out_w.push_back(e.second);
        }
        L.out_end[i] = L.out_nb.size();
    }
    return L;
}

/* Partition (louvain/Partition.h:38-370) */
struct LPartition {
    LGraph g;
    int size;
    double min_modularity;
    std::vector<double> neigh_weight;
    std::vector<unsigned> neigh_pos;
    unsigned neigh_last = 0;
    std::vector<int> n2c;
    std::vector<double> woutc, winc, wout, win, woutctot, winctot;

    LPartition(const LGraph &gr, double minm) : g(gr)
    {
        size = (int)g.nodes;
        neigh_weight.assign(size, -1);
        neigh_pos.assign(size, 0);
        n2c.resize(size);
        woutc.resize(size); winc.resize(size); wout.resize(size);
        win.resize(size); woutctot.resize(size); winctot.resize(size);
        for (int i = 0; i < size; ++i) {
            n2c[i] = i;
            woutc[i] = winc[i] = g.selfloop(i);
            wout[i] = woutctot[i] = g.out_degree_w(i);
            win[i] = winctot[i] = g.in_degree_w(i);
        }
        min_modularity = minm;
    }

    /* Partition.h:94-102 */
    void remove(int node)
    {
        woutc[node] = 0;
        winc[node] = 0;
        woutctot[n2c[node]] -= g.out_degree_w(node);
        winctot[n2c[node]] -= g.in_degree_w(node);
        n2c[node] = -1;
    }

    /* Partition.h:104-120, including the out-degree added to winctot */
    void insert(int node, int comm)
    {
        n2c[node] = comm;
        for (unsigned i = 0; i < g.out_count(node); ++i)
            if (n2c[g.out_nb[g.out_begin(node) + i]] == comm)
                woutc[node] += g.out_w[g.out_begin(node) + i];
        for (unsigned i = 0; i < g.in_count(node); ++i)
            if (n2c[g.in_nb[g.in_begin(node) + i]] == comm)
                winc[node] += g.in_w[g.in_begin(node) + i];
        woutctot[comm] += g.out_degree_w(node);
        winctot[comm] += g.out_degree_w(node);
    }

    /* Partition.h:122-140 */
    double gain(int node, int comm, double wic) const
    {
        double woutn = wout[node], winn = win[node];
        double wc_in = winctot[comm], wc_out = woutctot[comm];
        double m = g.total_weight;
        return (wic / m - (woutn / m) * (wc_in / m) - (winn / m) * (wc_out / m));
    }

    /* Partition.h:142-152 */
    double modularity() const
    {
        double q = 0.;
        double m = g.total_weight;
        for (int i = 0; i < size; ++i)
            if (wout[i] > 0)
                q += woutc[i] / m - (wout[i] / m) * (winctot[n2c[i]] / m);
        return q;
    }

    /* Partition.h:154-197 */
    void neigh_comm(unsigned node)
    {
        for (unsigned i = 0; i < neigh_last; ++i) neigh_weight[neigh_pos[i]] = -1;
        neigh_last = 0;
        neigh_pos[0] = (unsigned)n2c[node];
        neigh_weight[neigh_pos[0]] = 0;
        neigh_last = 1;
        for (unsigned i = 0; i < g.in_count((int)node); ++i) {
            unsigned nb = (unsigned)g.in_nb[g.in_begin((int)node) + i];
            unsigned c = (unsigned)n2c[nb];
            double w = g.in_w.empty() ? 1. : g.in_w[g.in_begin((int)node) + i];
            if (nb != node) {
                if (neigh_weight[c] == -1) {
                    neigh_weight[c] = 0.;
                    neigh_pos[neigh_last++] = c;
                }
                neigh_weight[c] += w;
            }
        }
        for (unsigned i = 0; i < g.out_count((int)node); ++i) {
            unsigned nb = (unsigned)g.out_nb[g.out_begin((int)node) + i];
            if (nb != node) {
                unsigned c = (unsigned)n2c[nb];
                double w = g.out_w.empty() ? 1. : g.out_w[g.out_begin((int)node) + i];
                if (neigh_weight[c] == -1) {
                    neigh_weight[c] = 0.;
                    neigh_pos[neigh_last++] = c;
                }
                neigh_weight[c] += w;
            }
        }
    }

    /* Partition.h:297-370 */
    bool one_level(const std::vector<int> &order)
    {
        bool improvement = false;
        int moves;
        double new_mod = modularity();
        double cur_mod = new_mod;
        do {
            cur_mod = new_mod;
            moves = 0;
            for (int t = 0; t < size; ++t) {
                int node = order[t];
                int node_comm = n2c[node];
                neigh_comm((unsigned)node);
                remove(node);
                int best_comm = node_comm;
                double best_increase = 0.;
                for (unsigned i = 0; i < neigh_last; ++i) {
                    double inc = gain(node, (int)neigh_pos[i], neigh_weight[neigh_pos[i]]);
                    if (inc > best_increase) {
                        best_comm = (int)neigh_pos[i];
                        best_increase = inc;
                    }
                }
                insert(node, best_comm);
                if (best_comm != node_comm) ++moves;
            }
            new_mod = modularity();
            if (moves > 0) improvement = true;
        } while (moves > 0 && new_mod - cur_mod > min_modularity);
        return improvement;
    }

    /* Partition.h:288-295: libstdc++ std::shuffle driven by the instance's
     * std::mt19937. */
    bool one_level(std::mt19937 &eng)
    {
        std::vector<int> order((size_t)size);
        for (int i = 0; i < size; ++i) order[i] = i;
        std::shuffle(order.begin(), order.end(), eng);
        return one_level(order);
    }

    /* Partition.h:200-283 */
    LGraph to_graph() const
    {
        std::vector<int> renumber((size_t)size, -1);
        for (int node = 0; node < size; ++node) renumber[n2c[node]]++;
        int fin = 0;
        for (int i = 0; i < size; ++i)
            if (renumber[i] != -1) renumber[i] = fin++;
        std::vector<std::vector<int>> comm_nodes((size_t)fin);
        for (int node = 0; node < size; ++node) comm_nodes[renumber[n2c[node]]].push_back(node);

        LGraph g2;
        g2.nodes = (unsigned)comm_nodes.size();
        g2.total_weight = 0;
        g2.links = 0;
        g2.in_end.resize(comm_nodes.size());
        g2.out_end.resize(comm_nodes.size());
        for (int comm = 0; comm < (int)comm_nodes.size(); ++comm) {
            std::map<int, double> inm, outm;
            for (int node : comm_nodes[comm]) {
                for (unsigned i = 0; i < g.in_count(node); ++i) {
                    int nc = renumber[n2c[g.in_nb[g.in_begin(node) + i]]];
                    double w = g.in_w.empty() ? 1. : g.in_w[g.in_begin(node) + i];
                    auto it = inm.find(nc);
                    if (it == inm.end()) inm.insert(std::make_pair(nc, w));
                    else it->second += w;
                }
                for (unsigned i = 0; i < g.out_count(node); ++i) {
                    int nc = renumber[n2c[g.out_nb[g.out_begin(node) + i]]];
                    double w = g.out_w.empty() ? 1. : g.out_w[g.out_begin(node) + i];
                    auto it = outm.find(nc);
                    if (it == outm.end()) outm.insert(std::make_pair(nc, w));
                    else it->second += w;
                }
            }
            g2.in_end[comm] = (comm == 0) ? inm.size() : g2.in_end[comm - 1] + inm.size();
            g2.out_end[comm] = (comm == 0) ? outm.size() : g2.out_end[comm - 1] + outm.size();
            g2.links += inm.size();
            for (const auto &e : inm) {
                g2.total_weight += e.second;
                g2.in_nb.push_back(e.first);
                g2.in_w.push_back(e.second);
            }
            for (const auto &e : outm) {
                g2.out_nb.push_back(e.first);
                g2.out_w.push_back(e.second);
            }
        }
        return g2;
    }
};

/* LouvainGraphPartition::partitionGraph (LouvainGraphPartition.h:84-130) with
 * renumber_communities (:27-46) and build_result (:48-69). */
std::vector<std::vector<int>> louvain(const orc_graph &g, const std::vector<uint64_t> &seed_list,
                                      double precision)
{
    std::set<uint64_t> seeds(seed_list.begin(), seed_list.end());
    std::vector<std::mt19937> eng;
    for (uint64_t s : seeds) eng.push_back(std::mt19937((std::mt19937::result_type)s));
    const int P = (int)eng.size();
    LGraph lg = lgraph_from(g);
    std::vector<LPartition> p((size_t)P, LPartition(lg, precision));
    std::vector<int> n2c(lg.nodes);
    for (unsigned i = 0; i < lg.nodes; ++i) n2c[i] = (int)i;
    std::vector<char> improvements((size_t)P, 1);
    std::vector<double> mods((size_t)P);
    bool improvement;
    do {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < P; ++i) {
            improvements[i] = p[i].one_level(eng[i]);
            mods[i] = p[i].modularity();
        }
        int best = 0;
        for (int i = 1; i < P; ++i)
            if (mods[i] > mods[best]) best = i;
        improvement = improvements[best] != 0;
        lg = p[best].to_graph();
        {
            const std::vector<int> &lvl = p[best].n2c;
            std::vector<int> renumber(lvl.size(), -1);
            for (size_t node = 0; node < lvl.size(); ++node) renumber[lvl[node]]++;
            int fin = 0;
            for (size_t i = 0; i < lvl.size(); ++i)
                if (renumber[i] != -1) renumber[i] = fin++;
            for (size_t i = 0; i < n2c.size(); ++i) n2c[i] = renumber[lvl[n2c[i]]];
        }
        for (int i = 0; i < P; ++i) p[i] = LPartition(lg, precision);
    } while (improvement);
    int mx = 0;
    for (int c : n2c) mx = std::max(mx, c);
    std::vector<std::vector<int>> r((size_t)mx + 1);
    for (size_t i = 0; i < n2c.size(); ++i) r[n2c[i]].push_back((int)i);
    return r;
}

/* ---- clustered BC (ClusteredBrandesBC.h:63-169) ---------------------------- */
struct ClusteredRun {
    std::vector<std::vector<int>> comms;
    std::vector<Cluster> clusters;
    std::vector<std::vector<int>> piv, piv_w;      /* global vertex ids */
    std::vector<double> intra;
    double te = 0;
    double t_louvain = 0, t_cluster = 0, t_pivot = 0;
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

/* ClusteredBrandeBC::computeBC (ClusteredBrandesBC.h:63-169).  The partition
 * comes from Louvain (LouvainGraphPartition, as main.cpp composes it), or,
 * when 'given' is not null, from the caller: the reference takes any
 * IGraphPartition strategy (ClusteredBrandesBC.h:80-82), whose communities
 * are vertex lists; here community c holds the vertices v with given[v] == c
 * in ascending order, communities in ascending c, empty ones left out; a
 * negative index leaves the vertex in no community. */
int clustered(const orc_graph &g, const std::vector<uint64_t> &seeds, double precision,
              double kfrac, int threads, long long max_pivots, std::vector<double> &bc,
              ClusteredRun &run, const int *given = nullptr)
{
    auto t0 = std::chrono::steady_clock::now();
    bc.assign(g.n, 0.0);
    if (given) {
        int k = 0;
        for (int v = 0; v < g.n; ++v) k = std::max(k, given[v] + 1);
        std::vector<std::vector<int>> comms((size_t)k);
        for (int v = 0; v < g.n; ++v)
            if (given[v] >= 0) comms[(size_t)given[v]].push_back(v);
        for (auto &c : comms)
            if (!c.empty()) run.comms.push_back(std::move(c));
    } else {
        run.comms = louvain(g, seeds, precision);
    }
    run.t_louvain = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    const size_t nc = run.comms.size();
    run.clusters.resize(nc);
    run.piv.assign(nc, {});
    run.piv_w.assign(nc, {});
    /* clusters touch disjoint entries of bc, so any schedule gives the same
     * values (ClusteredBrandesBC.h:91-121) */
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (long long i = 0; i < (long long)nc; ++i) {
        Cluster &c = run.clusters[i];
        make_cluster(g, run.comms[i], c);
        if (c.borders.empty()) continue;   /* FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED */
        std::vector<double> vi_len;
        std::vector<int> vi_cnt;
        evaluate_cluster(c, bc, vi_len, vi_cnt);
        const int nb = (int)c.borders.size();
        std::vector<double> bc_row(c.verts.size());
        for (size_t r = 0; r < c.verts.size(); ++r) bc_row[r] = bc[c.verts[r]];
        std::vector<int> rows, w;
        select_pivots(c.verts, nb, vi_len, vi_cnt, bc_row, c.is_border, rows, w);
        if (kfrac > 0) {
            std::vector<int> krows, kw;
            kmeans_pivots(rows, w, nb, vi_len, vi_cnt, kfrac, krows, kw);
            rows.swap(krows);
            w.swap(kw);
        }
        for (size_t p = 0; p < rows.size(); ++p) {
            run.piv[i].push_back(c.verts[rows[p]]);
            run.piv_w[i].push_back(w[p]);
        }
    }
    run.intra = bc;
    run.t_cluster = seconds_since(t0);
    t0 = std::chrono::steady_clock::now();
    long long npiv = 0;
    for (size_t i = 0; i < nc; ++i) {
        npiv += (long long)run.piv[i].size();
        if (!run.clusters[i].borders.empty())
            run.te += (double)run.clusters[i].verts.size() * (double)run.clusters[i].edges;
    }
    run.te += (double)npiv * (double)g.edges;

    long long done = 0;
    for (size_t c = 0; c < nc && done < max_pivots; ++c) {
        const std::vector<int> &piv = run.piv[c];
        long long np = (long long)piv.size();
        if (done + np > max_pivots) np = max_pivots - done;
        done += np;
        std::vector<std::vector<double>> part((size_t)threads, std::vector<double>());
#pragma omp parallel num_threads(threads)
        {
#ifdef _OPENMP
            int tid = omp_get_thread_num(), nth = omp_get_num_threads();
#else
            int tid = 0, nth = 1;
#endif
            long long b, e;
            static_chunk(np, nth, tid, b, e);
            std::vector<double> priv((size_t)g.n, 0.0), delta((size_t)g.n);
            SingleSource s;
            for (long long p = b; p < e; ++p) {
                std::fill(delta.begin(), delta.end(), 0.0);
                dijkstra(g.fwd, nullptr, piv[p], s);
                std::vector<double> dep((size_t)g.n, 0.0);
                backward(s, piv[p], delta, [&](int w, double dw) { dep[w] += dw; });
                const double wp = (double)run.piv_w[c][p];
                for (int v = 0; v < g.n; ++v) priv[v] += dep[v] * wp;
                for (int v : run.clusters[c].verts) priv[v] -= run.intra[v] * wp;
            }
            part[tid].swap(priv);
        }
        for (int t = 0; t < threads; ++t)
            if (!part[t].empty())
                for (int v = 0; v < g.n; ++v) bc[v] += part[t][v];
    }
    run.t_pivot = seconds_since(t0);
    return (int)npiv;
}

} // namespace

/* ---- C API --------------------------------------------------------------- */
extern "C" {

orc_graph *orc_graph_from_file(const char *path, char *err, int errlen)
{
    std::ifstream in(path);
    if (!in.is_open()) {
        set_err(err, errlen, std::string("cannot open ") + path);
        return nullptr;
    }
    GraphBuilder b;
    for (;;) {
        long long s, d;
        double w;
        if (!(in >> s)) {
            if (in.eof()) break;
            set_err(err, errlen, "malformed edge list");
            return nullptr;
        }
        if (!(in >> d >> w)) {
            set_err(err, errlen, "malformed edge list");
            return nullptr;
        }
        /* DirectedWeightedGraph.h:78-81 */
        if (w <= 0) {
            set_err(err, errlen, "Edge weight must be greater than zero");
            return nullptr;
        }
        if (s < 0 || d < 0 || s > INT_MAX || d > INT_MAX) {
            set_err(err, errlen, "vertex id out of range");
            return nullptr;
        }
        b.add((int)s, (int)d, w);
    }
    return b.finish(err, errlen);
}

orc_graph *orc_graph_from_arcs(uint64_t m, const uint32_t *src, const uint32_t *dst,
                               const double *w, char *err, int errlen)
{
    GraphBuilder b;
    for (uint64_t i = 0; i < m; ++i) {
        if (w[i] <= 0) {
            set_err(err, errlen, "Edge weight must be greater than zero");
            return nullptr;
        }
        if (src[i] > (uint32_t)INT_MAX || dst[i] > (uint32_t)INT_MAX) {
            set_err(err, errlen, "vertex id out of range");
            return nullptr;
        }
        b.add((int)src[i], (int)dst[i], w[i]);
    }
    return b.finish(err, errlen);
}

void orc_graph_free(orc_graph *g) { delete g; }
int orc_graph_n(const orc_graph *g) { return g->n; }
long long orc_graph_edges(const orc_graph *g) { return g->edges; }

int orc_graph_csr(const orc_graph *g, uint64_t *rowptr, uint32_t *col, double *w)
{
    uint64_t k = 0;
    for (int v = 0; v < g->n; ++v) {
        rowptr[v] = k;
        for (const auto &e : g->fwd[v]) {
            col[k] = (uint32_t)e.first;
            w[k] = e.second;
            ++k;
        }
    }
    rowptr[g->n] = k;
    return 0;
}

int orc_graph_bstar(const orc_graph *g, int v, int cap, int *src, double *w)
{
    if (v < 0 || v >= g->n) return -1;
    int i = 0;
    for (const auto &e : g->bwd[v]) {
        if (i < cap) { src[i] = e.first; w[i] = e.second; }
        ++i;
    }
    return i;
}

int orc_sssp(const orc_graph *g, int src, double *dist, double *sigma, int *order)
{
    if (src < 0 || src >= g->n) return -1;
    SingleSource s;
    dijkstra(g->fwd, nullptr, src, s);
    for (int v = 0; v < g->n; ++v) {
        dist[v] = s.dist[v] == DBL_MAX ? INFINITY : s.dist[v];
        sigma[v] = s.sigma[v];
        order[v] = v < (int)s.order.size() ? s.order[v] : -1;
    }
    return (int)s.order.size();
}

int orc_single_source_dep(const orc_graph *g, int src, double *dep)
{
    if (src < 0 || src >= g->n) return -1;
    SingleSource s;
    dijkstra(g->fwd, nullptr, src, s);
    std::vector<double> delta((size_t)g->n, 0.0);
    for (int v = 0; v < g->n; ++v) dep[v] = 0;
    backward(s, src, delta, [&](int w, double dw) { dep[w] += dw; });
    return 0;
}

int orc_exact_bc_range(const orc_graph *g, int s0, int s1, int threads, double *bc)
{
    if (threads < 1) threads = 1;
    if (s0 < 0) s0 = 0;
    if (s1 > g->n) s1 = g->n;
    const long long ns = s1 > s0 ? s1 - s0 : 0;
    std::vector<std::vector<double>> part((size_t)threads);
#pragma omp parallel num_threads(threads)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num(), nth = omp_get_num_threads();
#else
        int tid = 0, nth = 1;
#endif
        long long b, e;
        static_chunk(ns, nth, tid, b, e);
        std::vector<double> priv((size_t)g->n, 0.0), delta((size_t)g->n);
        SingleSource s;
        for (long long i = b; i < e; ++i) {
            int src = (int)(s0 + i);
            std::fill(delta.begin(), delta.end(), 0.0);
            dijkstra(g->fwd, nullptr, src, s);
            backward(s, src, delta, [&](int w, double dw) { priv[w] += dw; });
        }
        part[tid].swap(priv);
    }
    for (int v = 0; v < g->n; ++v) bc[v] = 0;
    for (int t = 0; t < threads; ++t)
        if (!part[t].empty())
            for (int v = 0; v < g->n; ++v) bc[v] += part[t][v];
    return 0;
}

int orc_exact_bc(const orc_graph *g, int threads, double *bc)
{
    return orc_exact_bc_range(g, 0, g->n, threads, bc);
}

int orc_subgraph(const orc_graph *g, int nv, const int *verts, int *nborders,
                 int *borders, long long *edges)
{
    for (int i = 0; i < nv; ++i)
        if (verts[i] < 0 || verts[i] >= g->n) return -1;
    Cluster c;
    make_cluster(*g, std::vector<int>(verts, verts + nv), c);
    *nborders = (int)c.borders.size();
    for (size_t i = 0; i < c.borders.size(); ++i) borders[i] = c.borders[i];
    *edges = c.edges;
    return 0;
}

int orc_cluster_eval(const orc_graph *g, int nv, const int *verts, double *bc,
                     int *nborders, int *borders, double *vi_len, int *vi_cnt)
{
    for (int i = 0; i < nv; ++i)
        if (verts[i] < 0 || verts[i] >= g->n) return -1;
    Cluster c;
    make_cluster(*g, std::vector<int>(verts, verts + nv), c);
    std::vector<double> bcv(bc, bc + g->n), L;
    std::vector<int> C;
    evaluate_cluster(c, bcv, L, C);
    std::copy(bcv.begin(), bcv.end(), bc);
    *nborders = (int)c.borders.size();
    std::copy(c.borders.begin(), c.borders.end(), borders);
    std::copy(L.begin(), L.end(), vi_len);
    std::copy(C.begin(), C.end(), vi_cnt);
    return 0;
}

int orc_select_pivots(int nv, const int *verts, int nb, double *vi_len, const int *vi_cnt,
                      const double *bc_of_vert, int nborders, const int *borders,
                      int *pivots, int *weights)
{
    std::vector<int> v(verts, verts + nv);
    std::vector<double> L(vi_len, vi_len + (size_t)nv * nb);
    std::vector<int> C(vi_cnt, vi_cnt + (size_t)nv * nb);
    std::vector<double> bcr(bc_of_vert, bc_of_vert + nv);
    std::set<int> bset(borders, borders + nborders);
    std::vector<char> isb((size_t)nv);
    for (int i = 0; i < nv; ++i) isb[i] = bset.count(verts[i]) ? 1 : 0;
    std::vector<int> rows, w;
    select_pivots(v, nb, L, C, bcr, isb, rows, w);
    std::copy(L.begin(), L.end(), vi_len);
    for (size_t i = 0; i < rows.size(); ++i) {
        pivots[i] = verts[rows[i]];
        weights[i] = w[i];
    }
    return (int)rows.size();
}

int orc_kmeans(int k, int P, int nb, const double *vi_len, const int *vi_cnt,
               const int *weights, double min_variance, int max_iteration,
               int *centroids, int *out_weights)
{
    if (k < 1 || P < 1) return -1;
    std::vector<int> cent, w;
    kmeans(k, P, nb, vi_len, vi_cnt, weights, min_variance, max_iteration, cent, w);
    for (size_t i = 0; i < cent.size(); ++i) {
        centroids[i] = cent[i];
        out_weights[i] = w[i];
    }
    return (int)cent.size();
}

float orc_vi_sqdist_f32(int nb, const float *len_a, const int *cnt_a, const int *len_b, const int *cnt_b)
{
    return sqdist_points_t<float, int>(len_a, cnt_a, len_b, cnt_b, nb);
}

double orc_vi_sqdist(int nb, const double *len_a, const int *cnt_a, const double *len_b, const int *cnt_b)
{
    return sqdist_points(len_a, cnt_a, len_b, cnt_b, nb);
}

int orc_louvain(const orc_graph *g, int nseeds, const uint64_t *seeds, double precision,
                int *comm_of)
{
    if (nseeds < 1) return -1;
    auto comms = louvain(*g, std::vector<uint64_t>(seeds, seeds + nseeds), precision);
    for (size_t c = 0; c < comms.size(); ++c)
        for (int v : comms[c]) comm_of[v] = (int)c;
    return (int)comms.size();
}

int orc_clustered_bc(const orc_graph *g, int nseeds, const uint64_t *seeds, double precision,
                     double kfrac, int threads, double *bc, int *comm_of, double *intra_bc,
                     int *pivots, int *weights, int *pivot_cluster, double *te)
{
    if (nseeds < 1 || threads < 1) return -1;
    std::vector<double> out;
    ClusteredRun run;
    int np = clustered(*g, std::vector<uint64_t>(seeds, seeds + nseeds), precision, kfrac,
                       threads, LLONG_MAX, out, run);
    std::copy(out.begin(), out.end(), bc);
    if (comm_of)
        for (size_t c = 0; c < run.comms.size(); ++c)
            for (int v : run.comms[c]) comm_of[v] = (int)c;
    if (intra_bc) std::copy(run.intra.begin(), run.intra.end(), intra_bc);
    int k = 0;
    for (size_t c = 0; c < run.piv.size(); ++c)
        for (size_t p = 0; p < run.piv[c].size(); ++p, ++k) {
            if (pivots) pivots[k] = run.piv[c][p];
            if (weights) weights[k] = run.piv_w[c][p];
            if (pivot_cluster) pivot_cluster[k] = (int)c;
        }
    if (te) *te = run.te;
    return np;
}

int orc_clustered_bc_partition(const orc_graph *g, const int *comm_in, double kfrac, int threads, double *bc,
                               double *intra_bc, int *pivots, int *weights, int *pivot_cluster)
{
    if (!comm_in || threads < 1) return -1;
    std::vector<double> out;
    ClusteredRun run;
    int np = clustered(*g, {}, 0.0, kfrac, threads, LLONG_MAX, out, run, comm_in);
    std::copy(out.begin(), out.end(), bc);
    if (intra_bc) std::copy(run.intra.begin(), run.intra.end(), intra_bc);
    int k = 0;
    for (size_t c = 0; c < run.piv.size(); ++c)
        for (size_t p = 0; p < run.piv[c].size(); ++p, ++k) {
            if (pivots) pivots[k] = run.piv[c][p];
            if (weights) weights[k] = run.piv_w[c][p];
            if (pivot_cluster) pivot_cluster[k] = (int)c;
        }
    return np;
}

int orc_kmeans_pivots(int P, int nb, const double *vi_len, const int *vi_cnt, const int *weights, double kfrac,
                      int *pivots, int *out_weights)
{
    if (P < 1 || nb < 0) return -1;
    std::vector<int> rows((size_t)P), w(weights, weights + P), pr, pw;
    for (int p = 0; p < P; ++p) rows[(size_t)p] = p;
    std::vector<double> L(vi_len, vi_len + (size_t)P * nb);
    std::vector<int> C(vi_cnt, vi_cnt + (size_t)P * nb);
    kmeans_pivots(rows, w, nb, L, C, kfrac, pr, pw);
    std::copy(pr.begin(), pr.end(), pivots);
    std::copy(pw.begin(), pw.end(), out_weights);
    return (int)pr.size();
}

int orc_clustered_pivots(const orc_graph *g, int nseeds, const uint64_t *seeds, double precision,
                         double kfrac, int threads, int *comm_of, double *intra_bc, int *pivots,
                         int *weights, int *pivot_cluster)
{
    if (nseeds < 1 || threads < 1) return -1;
    std::vector<double> out;
    ClusteredRun run;
    int np = clustered(*g, std::vector<uint64_t>(seeds, seeds + nseeds), precision, kfrac, threads, 0, out,
                       run);
    for (size_t c = 0; c < run.comms.size(); ++c)
        for (int v : run.comms[c]) comm_of[v] = (int)c;
    std::copy(run.intra.begin(), run.intra.end(), intra_bc);
    int k = 0;
    for (size_t c = 0; c < run.piv.size(); ++c)
        for (size_t p = 0; p < run.piv[c].size(); ++p, ++k) {
            pivots[k] = run.piv[c][p];
            weights[k] = run.piv_w[c][p];
            pivot_cluster[k] = (int)c;
        }
    return np;
}

int orc_clustered_bc_sampled(const orc_graph *g, int nseeds, const uint64_t *seeds,
                             double precision, double kfrac, int threads, int max_pivots,
                             double *bc, double *phase_seconds)
{
    if (nseeds < 1 || threads < 1) return -1;
    std::vector<double> out;
    ClusteredRun run;
    int np = clustered(*g, std::vector<uint64_t>(seeds, seeds + nseeds), precision, kfrac,
                       threads, max_pivots, out, run);
    std::copy(out.begin(), out.end(), bc);
    if (phase_seconds) {
        phase_seconds[0] = run.t_louvain;
        phase_seconds[1] = run.t_cluster;
        phase_seconds[2] = run.t_pivot;
    }
    return np;
}

} // extern "C"
