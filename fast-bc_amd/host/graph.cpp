/*
 * Edge-list loading with DirectedWeightedGraph semantics
 * (DirectedWeightedGraph.h:67-96 loader, :135-170 addEdge, :172-182 vertices):
 *   - '<src> <dst> <weight>' triples, ids as int, weight as double;
 *   - a weight <= 0 (or an unreadable weight) throws
 *     std::invalid_argument("Edge weight must be greater than zero");
 *   - an unreadable id ends the input, as the reference's stream loop does;
 *   - repeated arcs are summed in file order and counted once;
 *   - the vertex count is 1 + the largest source id.
 * The text is read whole and parsed in parallel chunks cut at line ends
 * (merged in file order, with the sequential parse's stop and error rules;
 * a triple split across a chunk boundary sends the file to the sequential
 * parse); arcs are grouped by a counting sort on the source, stable, then by
 * target within each row, instead of std::map inserts.
 */
#include "fbc_host.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <numeric>
#include <stdexcept>
#include <sys/stat.h>

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
    /* file-order sums of repeated arcs (DirectedWeightedGraph.h:154-165):
     * arcs bucketed by source in file order, each bucket stably sorted by
     * target, equal targets summed in file order */
    const size_t nn = (size_t)std::max(g.n, 0);
    std::vector<uint64_t> bstart(nn + 1, 0);
    for (size_t i = 0; i < m; ++i) bstart[(size_t)src[i] + 1]++;
    for (size_t v = 0; v < nn; ++v) bstart[v + 1] += bstart[v];
    std::vector<uint32_t> idx(m);
    {
        std::vector<uint64_t> fill(bstart.begin(), bstart.end() - 1);
        for (size_t i = 0; i < m; ++i) idx[fill[src[i]]++] = (uint32_t)i;
    }
    std::vector<uint32_t> ucount(nn, 0);
#pragma omp parallel for schedule(dynamic, 4096)
    for (size_t v = 0; v < nn; ++v) {
        uint32_t *b = idx.data() + bstart[v], *e = idx.data() + bstart[v + 1];
        if (e - b > 32)
            std::stable_sort(b, e, [&](uint32_t x, uint32_t y) { return dst[x] < dst[y]; });
        else
            for (uint32_t *p = b + 1; p < e; ++p) {   /* insertion sort, stable */
                const uint32_t x = *p;
                uint32_t *q = p;
                while (q > b && dst[q[-1]] > dst[x]) {
                    *q = q[-1];
                    --q;
                }
                *q = x;
            }
        uint32_t u = 0;
        for (uint32_t *p = b; p < e; ++p) u += p == b || dst[p[-1]] != dst[*p];
        ucount[v] = u;
    }
    g.rp.assign(nn + 1, 0);
    for (size_t v = 0; v < nn; ++v) g.rp[v + 1] = g.rp[v] + ucount[v];
    g.col.resize(g.rp[nn]);
    g.w.resize(g.rp[nn]);
#pragma omp parallel for schedule(dynamic, 4096)
    for (size_t v = 0; v < nn; ++v) {
        const uint32_t *b = idx.data() + bstart[v], *e = idx.data() + bstart[v + 1];
        uint64_t o = g.rp[v];
        for (const uint32_t *p = b; p < e;) {
            const int t = dst[*p];
            double sum = w[*p];
            const uint32_t *q = p + 1;
            while (q < e && dst[*q] == t) sum += w[*q++];
            g.col[o] = t;
            g.w[o] = sum;
            ++o;
            p = q;
        }
    }
    for (size_t i = 0; i < m; ++i) {
        g.total_weight += w[i];
        g.in_wdeg[dst[i]] += w[i];
        g.out_wdeg[src[i]] += w[i];
    }
    finish_graph(g);
    return g;
}

namespace {

enum ChunkEnd { CHUNK_DONE, CHUNK_STOP, CHUNK_PARTIAL, CHUNK_BAD_WEIGHT, CHUNK_NEG };

/* Triples of [p, end) in order; stops as the sequential loop does: a
 * non-int id ends the input (CHUNK_STOP), whitespace only is CHUNK_DONE, ids
 * with no weight before 'end' are CHUNK_PARTIAL. */
ChunkEnd parse_chunk(const char *p, const char *end, std::vector<int> &src, std::vector<int> &dst,
                     std::vector<double> &w)
{
    for (;;) {
        long long s, d;
        const char *start = p;
        if (!parse_int(p, end, s)) {
            while (start < end && is_space(*start)) ++start;
            return start >= end ? CHUNK_DONE : CHUNK_STOP;
        }
        if (!parse_int(p, end, d)) {
            const char *q = p;
            while (q < end && is_space(*q)) ++q;
            return q >= end ? CHUNK_PARTIAL : CHUNK_STOP;
        }
        while (p < end && is_space(*p)) ++p;
        if (p >= end) return CHUNK_PARTIAL;
        /* plain digit tokens convert exactly; anything else goes to strtod */
        const char *q = p;
        uint64_t iv = 0;
        while (q < end && *q >= '0' && *q <= '9' && q - p < 15) iv = iv * 10 + (uint64_t)(*q++ - '0');
        double x;
        if (q > p && (q >= end || is_space(*q))) {
            x = (double)iv;
        } else {
            char *r = nullptr;
            errno = 0;
            x = std::strtod(p, &r);
            if (r == p || (r < end && !is_space(*r))) return CHUNK_BAD_WEIGHT;
            q = r;
        }
        if (!(x > 0)) return CHUNK_BAD_WEIGHT;
        if (s < 0 || d < 0) return CHUNK_NEG;
        p = q;
        src.push_back((int)s);
        dst.push_back((int)d);
        w.push_back(x);
    }
}

Graph graph_from_text(const std::string &text)
{
    const char *base = text.data(), *end = base + text.size();
    const int nchunk = text.size() < (1u << 20) ? 1 : 64;
    std::vector<const char *> cut(nchunk + 1);
    cut[0] = base;
    cut[nchunk] = end;
    for (int k = 1; k < nchunk; ++k) {
        const char *c = base + text.size() * k / nchunk;
        while (c < end && *c != '\n') ++c;
        cut[k] = c < end ? c + 1 : end;
        if (cut[k] < cut[k - 1]) cut[k] = cut[k - 1];
    }
    std::vector<std::vector<int>> cs(nchunk), cd(nchunk);
    std::vector<std::vector<double>> cw(nchunk);
    std::vector<ChunkEnd> ce(nchunk);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nchunk; ++k) ce[k] = parse_chunk(cut[k], cut[k + 1], cs[k], cd[k], cw[k]);
    /* the chunks that count, in order, and where their triples go */
    int used = 0;
    for (int k = 0; k < nchunk; ++k) {
        if (ce[k] == CHUNK_PARTIAL && k + 1 < nchunk) {
            /* a triple across the cut: parse the whole text in one piece */
            std::vector<int> src, dst;
            std::vector<double> w;
            const ChunkEnd e = parse_chunk(base, end, src, dst, w);
            if (e == CHUNK_BAD_WEIGHT) throw std::invalid_argument("Edge weight must be greater than zero");
            if (e == CHUNK_NEG) throw std::invalid_argument("negative vertex id");
            return from_arcs(src.size(), src.data(), dst.data(), w.data());
        }
        used = k + 1;
        if (ce[k] == CHUNK_BAD_WEIGHT) throw std::invalid_argument("Edge weight must be greater than zero");
        if (ce[k] == CHUNK_NEG) throw std::invalid_argument("negative vertex id");
        if (ce[k] == CHUNK_STOP || ce[k] == CHUNK_PARTIAL) break;   /* the input ends here */
    }
    std::vector<size_t> off((size_t)used + 1, 0);
    for (int k = 0; k < used; ++k) off[k + 1] = off[k] + cs[k].size();
    std::vector<int> src(off[used]), dst(off[used]);
    std::vector<double> w(off[used]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < used; ++k) {
        std::copy(cs[k].begin(), cs[k].end(), src.begin() + off[k]);
        std::copy(cd[k].begin(), cd[k].end(), dst.begin() + off[k]);
        std::copy(cw[k].begin(), cw[k].end(), w.begin() + off[k]);
    }
    return from_arcs(src.size(), src.data(), dst.data(), w.data());
}

} // namespace

Graph load_edge_list(std::istream &in)
{
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return graph_from_text(text);
}

namespace {

/* The text file a cache belongs to: size, modification time, and an FNV-1a
 * hash of its first and last MiB. */
struct TextPrint {
    uint64_t size = 0;
    int64_t mtime_ns = 0;
    uint64_t hash = 0;
    bool operator==(const TextPrint &o) const { return size == o.size && mtime_ns == o.mtime_ns && hash == o.hash; }
};

bool text_print(const std::string &path, TextPrint &tp)
{
    struct stat st;
    if (::stat(path.c_str(), &st) != 0) return false;
    tp.size = (uint64_t)st.st_size;
    tp.mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000ll + st.st_mtim.tv_nsec;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    uint64_t h = 1469598103934665603ull;
    std::vector<unsigned char> buf(1 << 20);
    auto mix = [&](size_t r) {
        for (size_t i = 0; i < r; ++i) h = (h ^ buf[i]) * 1099511628211ull;
    };
    mix(std::fread(buf.data(), 1, buf.size(), f));
    if (tp.size > buf.size() && std::fseek(f, -(long)buf.size(), SEEK_END) == 0) mix(std::fread(buf.data(), 1, buf.size(), f));
    std::fclose(f);
    tp.hash = h;
    return true;
}

constexpr char CACHE_MAGIC[8] = {'F', 'B', 'C', 'C', 'S', 'R', '1', '\0'};

struct CacheHeader {
    char magic[8];
    TextPrint text;
    int64_t n, edges;
    double total_weight;
};

template <typename T>
bool read_vec(FILE *f, std::vector<T> &v, size_t count)
{
    v.resize(count);
    return std::fread(v.data(), sizeof(T), count, f) == count;
}

template <typename T>
bool write_vec(FILE *f, const std::vector<T> &v)
{
    return std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
}

bool read_cache(const std::string &cache, const TextPrint &tp, Graph &g)
{
    FILE *f = std::fopen(cache.c_str(), "rb");
    if (!f) return false;
    CacheHeader h;
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, CACHE_MAGIC, 8) == 0 && h.text == tp &&
              h.n >= 0 && h.n < INT_MAX && h.edges >= 0;
    if (ok) {
        g.n = (int)h.n;
        g.total_weight = h.total_weight;
        ok = read_vec(f, g.rp, (size_t)h.n + 1) && read_vec(f, g.col, (size_t)h.edges) &&
             read_vec(f, g.w, (size_t)h.edges) && read_vec(f, g.in_wdeg, (size_t)h.n) &&
             read_vec(f, g.out_wdeg, (size_t)h.n) && std::fgetc(f) == EOF;
        ok = ok && g.rp[0] == 0 && g.rp[(size_t)h.n] == (uint64_t)h.edges;
        for (int64_t v = 0; ok && v < h.n; ++v) ok = g.rp[v] <= g.rp[v + 1];
        for (size_t e = 0; ok && e < g.col.size(); ++e) ok = g.col[e] >= 0 && g.col[e] < g.n;
    }
    std::fclose(f);
    if (ok) finish_graph(g);
    return ok;
}

void write_cache(const std::string &cache, const TextPrint &tp, const Graph &g)
{
    const std::string tmp = cache + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;   /* a cache that cannot be written is simply not kept */
    CacheHeader h;
    std::memcpy(h.magic, CACHE_MAGIC, 8);
    h.text = tp;
    h.n = g.n;
    h.edges = g.edges;
    h.total_weight = g.total_weight;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && write_vec(f, g.rp) && write_vec(f, g.col) && write_vec(f, g.w) &&
              write_vec(f, g.in_wdeg) && write_vec(f, g.out_wdeg);
    ok = std::fclose(f) == 0 && ok;
    if (ok)
        std::rename(tmp.c_str(), cache.c_str());
    else
        std::remove(tmp.c_str());
}

} // namespace

Graph load_edge_list_cached(const std::string &path, const std::string &cache_path, bool *hit)
{
    TextPrint tp;
    if (!text_print(path, tp)) throw std::runtime_error("cannot open " + path);
    Graph g;
    if (read_cache(cache_path, tp, g)) {
        if (hit) *hit = true;
        return g;
    }
    if (hit) *hit = false;
    g = load_edge_list_file(path);
    write_cache(cache_path, tp, g);
    return g;
}

Graph load_edge_list_file(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string text;
    char buf[1 << 16];
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long size = std::ftell(f);
        if (size > 0) text.reserve((size_t)size);
        std::rewind(f);
    }
    for (size_t r; (r = std::fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, r);
    std::fclose(f);
    return graph_from_text(text);
}

} // namespace fbchost
