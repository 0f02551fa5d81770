/*
 * Edge-list loading with DirectedWeightedGraph semantics
 * (DirectedWeightedGraph.h:67-96 loader, :135-170 addEdge, :172-182 vertices):
 *   - '<src> <dst> <weight>' triples, ids as int, weight as double;
 *   - a weight <= 0 (or an unreadable weight) throws
 *     std::invalid_argument("Edge weight must be greater than zero");
 *   - an unreadable id ends the input, as the reference's stream loop does;
 *   - repeated arcs are summed in file order and counted once;
 *   - the vertex count is 1 + the largest source id.
 * The text is parsed in one pass over a buffer instead of iostream
 * extraction; arcs are grouped by a stable sort instead of std::map inserts.
 */
#include "fbc_host.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <numeric>
#include <stdexcept>

namespace fbchost {

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

/* Parses an int token; returns false when the token is not an int. */
bool parse_int(const char *&p, const char *end, long long &out)
{
    while (p < end && is_space(*p)) ++p;
    if (p >= end) return false;
    const char *q = p;
    bool neg = false;
    if (*q == '+' || *q == '-') {
        neg = *q == '-';
        ++q;
    }
    if (q >= end || *q < '0' || *q > '9') return false;
    long long v = 0;
    while (q < end && *q >= '0' && *q <= '9') {
        v = v * 10 + (*q - '0');
        if (v > (long long)INT_MAX + 1) return false;
        ++q;
    }
    v = neg ? -v : v;
    if (v < INT_MIN || v > INT_MAX) return false;
    out = v;
    p = q;
    return true;
}

} // namespace

void finish_graph(Graph &g)
{
    static std::atomic<uint64_t> next_id{1};
    g.id = next_id++;
    const int n = g.n;
    /* backward CSR: counting sort of the forward arcs by target; rows end up
     * sorted by source because sources are visited in ascending order */
    g.rrp.assign((size_t)n + 1, 0);
    for (size_t e = 0; e < g.col.size(); ++e) g.rrp[(size_t)g.col[e] + 1]++;
    for (int v = 0; v < n; ++v) g.rrp[v + 1] += g.rrp[v];
    g.rcol.assign(g.col.size(), 0);
    g.rw.assign(g.col.size(), 0.0);
    std::vector<uint64_t> fill(g.rrp.begin(), g.rrp.end() - 1);
    for (int u = 0; u < n; ++u)
        for (uint64_t e = g.rp[u]; e < g.rp[u + 1]; ++e) {
            const uint64_t pos = fill[g.col[e]]++;
            g.rcol[pos] = u;
            g.rw[pos] = g.w[e];
        }
    g.edges = (long long)g.col.size();
}

Graph from_arcs(size_t m, const int *src, const int *dst, const double *w)
{
    return from_arcs_n(-1, m, src, dst, w);
}

Graph from_arcs_n(int n, size_t m, const int *src, const int *dst, const double *w)
{
    Graph g;
    int maxsrc = -1;
    for (size_t i = 0; i < m; ++i) {
        if (src[i] < 0 || dst[i] < 0) throw std::invalid_argument("negative vertex id");
        maxsrc = std::max(maxsrc, src[i]);
    }
    if (n >= 0 && maxsrc >= n) throw std::invalid_argument("arc source beyond the vertex count");
    g.n = n >= 0 ? n : maxsrc + 1;
    for (size_t i = 0; i < m; ++i)
        if (dst[i] >= g.n)
            throw std::invalid_argument("arc target " + std::to_string(dst[i]) +
                                        " has no out-arc (the vertex count is 1 + the largest source id)");
    g.in_wdeg.assign((size_t)std::max(g.n, 0), 0.0);
    g.out_wdeg.assign((size_t)std::max(g.n, 0), 0.0);
    /* file-order sums of repeated arcs (DirectedWeightedGraph.h:154-165) */
    std::vector<size_t> order(m);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return src[a] != src[b] ? src[a] < src[b] : dst[a] < dst[b];
    });
    g.rp.assign((size_t)g.n + 1, 0);
    g.col.reserve(m);
    g.w.reserve(m);
    for (size_t i = 0; i < m;) {
        const size_t a = order[i];
        double sum = w[a];
        size_t j = i + 1;
        while (j < m && src[order[j]] == src[a] && dst[order[j]] == dst[a]) sum += w[order[j++]];
        g.col.push_back(dst[a]);
        g.w.push_back(sum);
        g.rp[(size_t)src[a] + 1]++;
        i = j;
    }
    for (int v = 0; v < g.n; ++v) g.rp[v + 1] += g.rp[v];
    for (size_t i = 0; i < m; ++i) {
        g.total_weight += w[i];
        g.in_wdeg[dst[i]] += w[i];
        g.out_wdeg[src[i]] += w[i];
    }
    finish_graph(g);
    return g;
}

Graph load_edge_list(std::istream &in)
{
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const char *p = text.data(), *end = text.data() + text.size();
    std::vector<int> src, dst;
    std::vector<double> w;
    for (;;) {
        long long s, d;
        if (!parse_int(p, end, s)) break;
        if (!parse_int(p, end, d)) break;
        while (p < end && is_space(*p)) ++p;
        if (p >= end) break;                      /* triple cut by the end of input */
        char *q = nullptr;
        errno = 0;
        const double x = std::strtod(p, &q);
        if (q == p || (q < end && !is_space(*q)) || !(x > 0))
            throw std::invalid_argument("Edge weight must be greater than zero");
        p = q;
        if (s < 0 || d < 0) throw std::invalid_argument("negative vertex id");
        src.push_back((int)s);
        dst.push_back((int)d);
        w.push_back(x);
    }
    return from_arcs(src.size(), src.data(), dst.data(), w.data());
}

Graph load_edge_list_file(const std::string &path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("cannot open " + path);
    return load_edge_list(in);
}

} // namespace fbchost
