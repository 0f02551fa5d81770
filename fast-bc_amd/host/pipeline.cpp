/*
 * Orchestration of the clustered betweenness path on one GPU
 * (ClusteredBrandeBC::computeBC, brandes/ClusteredBrandesBC.h:63-169):
 *
 *   1. Louvain on the host (bit-identical partition);
 *   2. cluster views + every (cluster, source) Brandes in one persistent
 *      launch (fbc_cluster_eval_resident): intra-cluster BC in the
 *      reference's FP64 order, VertexInfo rows left in device memory;
 *   3. pivot selection: exact classes of every cluster on the device from
 *      the resident VertexInfo rows (fbc_cluster_classes), k-means over
 *      them on the device (fbc_cluster_kmeans); the host selector takes over
 *      on a hash collision, the host k-means for sets the device declines;
 *   4. every pivot's full-graph Brandes in one persistent launch
 *      (fbc_pivot_phase), the weighted accumulation and the intra-cluster
 *      correction in FP64.
 *
 * Borderless clusters are skipped as with FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED
 * (main.cpp:6, ClusteredBrandesBC.h:99-120).
 */
#include "fbc_host.h"

#include <algorithm>
#include <memory>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <set>
#include <stdexcept>

namespace fbchost {

namespace {

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* Repeated centroids are dropped keeping the first occurrence and its
 * weight; prow / pw are replaced by the k-means pivots
 * (KMeansPivotSelector.h:103-129). */
void apply_centroids(uint32_t ncl, const std::vector<std::vector<int>> &cent, const std::vector<std::vector<int>> &cw,
                     std::vector<std::vector<int>> &prow, std::vector<std::vector<int>> &pw)
{
    for (uint32_t c = 0; c < ncl; ++c) {
        if (prow[c].empty()) continue;
        std::vector<int> rows, w;
        std::set<int> seen;
        for (size_t i = 0; i < cent[c].size(); ++i) {
            const int r = prow[c][cent[c][i]];
            if (seen.insert(r).second) {
                rows.push_back(r);
                w.push_back(cw[c][i]);
            }
        }
        prow[c].swap(rows);
        pw[c].swap(w);
    }
}

/* KMeansPivotSelector (KMeansPivotSelector.h:79-132) for every cluster:
 * k = max((int)(P * kfrac), 1) over the P exact pivots.  Sets whose distances
 * are exact in FP64 run as one GPU batch (fbc_kmeans_batch); the others on the
 * host.  Repeated centroids are dropped keeping the first occurrence and its
 * weight; prow / pw are replaced by the k-means pivots. */
void kmeans_stage(Device &dev, uint32_t ncl, const std::vector<uint32_t> &bptr, double kfrac,
                  std::vector<std::vector<int>> &prow, std::vector<std::vector<int>> &pw,
                  const std::vector<std::vector<double>> &plen, const std::vector<std::vector<int>> &pcnt,
                  PhaseTimes &t)
{
    const double t0 = now();
    /* a set goes to the GPU when its rows are integers small enough for exact
     * FP64 distances (which also makes them fit int32) and k is supported */
    std::vector<char> on_gpu(ncl, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t c = 0; c < ncl; ++c) {
        if (prow[c].empty()) continue;
        const uint32_t nb = bptr[c + 1] - bptr[c];
        double mx = 0;
        bool integral = true;
        for (double x : plen[c]) {
            mx = std::max(mx, std::fabs(x));
            integral &= x == std::floor(x);
        }
        for (int x : pcnt[c]) mx = std::max(mx, std::fabs((double)x));
        const int kc = std::max((int)(prow[c].size() * kfrac), 1);
        on_gpu[c] = integral && 2.0 * 2 * nb * mx * mx < 9007199254740992.0 && kc <= FBC_KMEANS_MAX_K;
    }
    std::vector<uint32_t> gsets, hsets;
    for (uint32_t c = 0; c < ncl; ++c)
        if (!prow[c].empty()) (on_gpu[c] ? gsets : hsets).push_back(c);
    std::vector<std::vector<int>> cent(ncl), cw(ncl);
    if (!gsets.empty()) {
        std::vector<uint32_t> np, nf, kk;
        std::vector<size_t> foff(gsets.size() + 1, 0), woff(gsets.size() + 1, 0);
        for (size_t i = 0; i < gsets.size(); ++i) {
            const uint32_t c = gsets[i];
            const uint32_t P = (uint32_t)prow[c].size(), nb = bptr[c + 1] - bptr[c];
            np.push_back(P);
            nf.push_back(2 * nb);
            kk.push_back((uint32_t)std::max((int)(P * kfrac), 1));
            foff[i + 1] = foff[i] + (size_t)P * 2 * nb;
            woff[i + 1] = woff[i] + P;
        }
        /* feature rows: length then count per border */
        std::vector<int32_t> feat(foff.back());
        std::vector<int32_t> w(woff.back());
#pragma omp parallel for schedule(dynamic, 1)
        for (size_t i = 0; i < gsets.size(); ++i) {
            const uint32_t c = gsets[i], P = np[i], nb = nf[i] / 2;
            int32_t *f = feat.data() + foff[i];
            for (size_t e = 0; e < (size_t)P * nb; ++e) {
                f[2 * e] = (int32_t)plen[c][e];
                f[2 * e + 1] = pcnt[c][e];
            }
            std::copy(pw[c].begin(), pw[c].end(), w.begin() + woff[i]);
        }
        const size_t ktot = std::accumulate(kk.begin(), kk.end(), (size_t)0);
        std::vector<uint32_t> oc(ktot);
        std::vector<int32_t> ow(ktot);
        t.kmeans_prep = now() - t0;
        dev.check(fbc_kmeans_batch(dev.ctx(), (uint32_t)gsets.size(), np.data(), nf.data(), kk.data(), feat.data(),
                                   w.data(), 0.0, 100, oc.data(), ow.data()));
        fbc_last_stats(dev.ctx(), &t.kmeans_gpu_ms, nullptr);
        size_t o = 0;
        for (size_t i = 0; i < gsets.size(); ++i) {
            cent[gsets[i]].assign(oc.begin() + o, oc.begin() + o + kk[i]);
            cw[gsets[i]].assign(ow.begin() + o, ow.begin() + o + kk[i]);
            o += kk[i];
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t i = 0; i < hsets.size(); ++i) {
        const uint32_t c = hsets[i], nb = bptr[c + 1] - bptr[c];
        const int P = (int)prow[c].size();
        kmeans(std::max((int)(P * kfrac), 1), P, (int)nb, plen[c].data(), pcnt[c].data(), pw[c].data(), 0.0, 100,
               cent[c], cw[c]);
    }
    apply_centroids(ncl, cent, cw, prow, pw);
    t.kmeans = now() - t0;
    t.kmeans_gpu_sets = gsets.size();
    t.kmeans_host_sets = hsets.size();
}

/* KMeansPivotSelector on the classes the device grouped
 * (fbc_cluster_kmeans); clusters it declines (features too large for exact
 * FP64 distances, too many centroids) run the host k-means on their class
 * rows. */
void kmeans_device_stage(Device &dev, fbc_clusters *cl, uint32_t ncl, const std::vector<uint32_t> &bptr,
                         double kfrac, std::vector<std::vector<int>> &prow, std::vector<std::vector<int>> &pw,
                         PhaseTimes &t)
{
    const double t0 = now();
    size_t total = 0;
    for (uint32_t c = 0; c < ncl; ++c) total += prow[c].size();
    std::vector<uint32_t> kptr((size_t)ncl + 1, 0), oc(std::max<size_t>(total, 1));
    std::vector<int32_t> ow(std::max<size_t>(total, 1));
    dev.check(fbc_cluster_kmeans(dev.ctx(), cl, kfrac, 0.0, 100, kptr.data(), oc.data(), ow.data()));
    std::vector<std::vector<int>> cent(ncl), cw(ncl);
    std::vector<uint32_t> hsets;
    for (uint32_t c = 0; c < ncl; ++c) {
        if (prow[c].empty()) continue;
        if (kptr[c + 1] > kptr[c]) {
            cent[c].assign(oc.begin() + kptr[c], oc.begin() + kptr[c + 1]);
            cw[c].assign(ow.begin() + kptr[c], ow.begin() + kptr[c + 1]);
        } else {
            hsets.push_back(c);
        }
    }
    if (kptr[ncl] > 0) fbc_last_stats(dev.ctx(), &t.kmeans_gpu_ms, nullptr);
    /* the declined clusters' points, then their k-means in parallel */
    std::vector<std::vector<double>> L(hsets.size());
    std::vector<std::vector<int>> C(hsets.size());
    for (size_t i = 0; i < hsets.size(); ++i) {
        const uint32_t c = hsets[i], nb = bptr[c + 1] - bptr[c];
        const size_t P = prow[c].size();
        std::vector<int32_t> f(P * 2 * nb);
        dev.check(fbc_cluster_class_features(dev.ctx(), cl, c, f.data()));
        L[i].resize(P * nb);
        C[i].resize(P * nb);
        for (size_t e = 0; e < P * nb; ++e) {
            L[i][e] = (double)f[2 * e];
            C[i][e] = f[2 * e + 1];
        }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t i = 0; i < hsets.size(); ++i) {
        const uint32_t c = hsets[i], nb = bptr[c + 1] - bptr[c];
        const int P = (int)prow[c].size();
        kmeans(std::max((int)(P * kfrac), 1), P, (int)nb, L[i].data(), C[i].data(), pw[c].data(), 0.0, 100, cent[c],
               cw[c]);
    }
    apply_centroids(ncl, cent, cw, prow, pw);
    t.kmeans = now() - t0;
    t.kmeans_gpu_sets = 0;
    for (uint32_t c = 0; c < ncl; ++c) t.kmeans_gpu_sets += kptr[c + 1] > kptr[c];
    t.kmeans_host_sets = hsets.size();
}

/* The host selector on rows copied out of the device: the fallback when the
 * device's class grouping meets a hash collision. */
void host_select(Device &dev, fbc_clusters *cl, uint32_t ncl, const std::vector<uint32_t> &cptr,
                 const std::vector<uint32_t> &verts, const std::vector<uint32_t> &bptr,
                 const std::vector<uint32_t> &borders, const std::vector<uint64_t> &vioff,
                 const std::vector<double> &intra, double kfrac, std::vector<std::vector<int>> &prow,
                 std::vector<std::vector<int>> &pw, PhaseTimes &t)
{
    /* uninitialised: the download writes every entry */
    std::unique_ptr<uint32_t[]> vl(new uint32_t[std::max<uint64_t>(vioff[ncl], 1)]);
    std::unique_ptr<int32_t[]> vc(new int32_t[std::max<uint64_t>(vioff[ncl], 1)]);
    dev.check(fbc_cluster_rows(dev.ctx(), cl, vl.get(), vc.get()));
    std::vector<std::vector<double>> plen(ncl);
    std::vector<std::vector<int>> pcnt(ncl);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t c = 0; c < ncl; ++c) {
        const int nv = (int)(cptr[c + 1] - cptr[c]);
        const int nb = (int)(bptr[c + 1] - bptr[c]);
        if (nb == 0) continue;
        std::vector<double> len((size_t)nv * nb);
        std::vector<int> cnt((size_t)nv * nb);
        for (size_t i = 0; i < (size_t)nv * nb; ++i) {
            len[i] = (double)vl[vioff[c] + i];
            cnt[i] = vc[vioff[c] + i];
        }
        std::vector<char> isb((size_t)nv, 0);
        for (int j = 0; j < nb; ++j) {
            const uint32_t b = borders[bptr[c] + j];
            /* borders are ascending, like the cluster's vertex list */
            const uint32_t *first = &verts[cptr[c]], *last = &verts[cptr[c + 1]];
            isb[std::lower_bound(first, last, b) - first] = 1;
        }
        std::vector<double> bcr((size_t)nv);
        for (int r = 0; r < nv; ++r) bcr[r] = intra[verts[cptr[c] + r]];
        select_pivots(nv, nb, len.data(), cnt.data(), bcr.data(), isb.data(), prow[c], pw[c]);
        if (kfrac > 0) {
            plen[c].resize(prow[c].size() * nb);
            pcnt[c].resize(prow[c].size() * nb);
            for (size_t p = 0; p < prow[c].size(); ++p)
                for (int j = 0; j < nb; ++j) {
                    plen[c][p * nb + j] = len[(size_t)prow[c][p] * nb + j];
                    pcnt[c][p * nb + j] = cnt[(size_t)prow[c][p] * nb + j];
                }
        }
    }
    if (kfrac > 0) kmeans_stage(dev, ncl, bptr, kfrac, prow, pw, plen, pcnt, t);
}

} // namespace

Device::Device(int device)
{
    const int rc = fbc_create(device, &ctx_);
    if (rc) throw std::runtime_error("fbc_create failed (status " + std::to_string(rc) +
                                     "): a gfx950 (MI355X) device is required");
    /* FBC_OPTIONS="name=value,..." sets fbc_set_option knobs on the context */
    if (const char *env = std::getenv("FBC_OPTIONS")) {
        std::string all(env);
        size_t at = 0;
        while (at < all.size()) {
            size_t end = all.find(',', at);
            if (end == std::string::npos) end = all.size();
            const std::string item = all.substr(at, end - at);
            at = end + 1;
            if (item.empty()) continue;
            const size_t eq = item.find('=');
            if (eq == std::string::npos) throw std::invalid_argument("FBC_OPTIONS entry without '=': " + item);
            char *stop = nullptr;
            const std::string val = item.substr(eq + 1);
            const long long v = std::strtoll(val.c_str(), &stop, 10);
            if (val.empty() || *stop) throw std::invalid_argument("FBC_OPTIONS value is not an integer: " + item);
            const int orc = fbc_set_option(ctx_, item.substr(0, eq).c_str(), v);
            if (orc) {
                const std::string msg = fbc_last_error(ctx_);
                fbc_destroy(ctx_);
                ctx_ = nullptr;
                throw std::invalid_argument("FBC_OPTIONS: " + msg);
            }
        }
    }
}

Device::~Device()
{
    if (graph_) fbc_graph_free(graph_);
    if (ctx_) fbc_destroy(ctx_);
}

Device &Device::shared()
{
    static Device dev(0);
    return dev;
}

void Device::check(int rc) const
{
    if (rc) throw std::runtime_error(std::string("fbc error ") + std::to_string(rc) + ": " + fbc_last_error(ctx_));
}

fbc_graph *Device::graph(const Graph &g)
{
    if (cached_ == g.id && graph_) return graph_;
    if (graph_) {
        fbc_graph_free(graph_);
        graph_ = nullptr;
        cached_ = 0;
    }
    std::vector<uint32_t> col(g.col.size()), w(g.w.size());
    for (size_t e = 0; e < g.w.size(); ++e) {
        const double x = g.w[e];
        if (!(x >= 1 && x <= 4294967295.0 && x == std::floor(x)))
            throw std::invalid_argument("the GPU path needs integer arc weights in [1, 2^32-1]");
        w[e] = (uint32_t)x;
        col[e] = (uint32_t)g.col[e];
    }
    check(fbc_graph_upload(ctx_, (uint32_t)g.n, g.col.size(), g.rp.data(), col.data(), w.data(), &graph_));
    cached_ = g.id;
    return graph_;
}

std::vector<double> exact_bc(Device &dev, const Graph &g, int prec)
{
    std::vector<double> bc((size_t)g.n, 0.0);
    if (g.n == 0) return bc;
    std::lock_guard<std::mutex> lk(dev.lock());
    dev.check(fbc_exact_bc(dev.ctx(), dev.graph(g), prec, bc.data()));
    return bc;
}

std::vector<double> single_source_dep(Device &dev, const Graph &g, int src, int prec)
{
    std::vector<double> dep((size_t)g.n, 0.0);
    if (src < 0 || src >= g.n) throw std::out_of_range("source vertex out of range");
    std::lock_guard<std::mutex> lk(dev.lock());
    dev.check(fbc_single_source_dep(dev.ctx(), dev.graph(g), (uint32_t)src, prec, dep.data()));
    return dep;
}

/* delta, half-team limit and how they were chosen, and whether the
 * LDS-view kernel ran, for the last launch (fbc_last_counters entries 14-16
 * and 18) */
static void last_tuning(Device &dev, uint64_t out[4])
{
    uint64_t c[32] = {};
    const int k = fbc_last_counters(dev.ctx(), c, 32);
    for (int i = 0; i < 3; ++i) out[i] = k >= 17 ? c[14 + i] : 0;
    out[3] = k >= 19 ? c[18] : 0;
}

ClusteredRun clustered_bc(Device &dev, const Graph &g, const ClusteredOptions &opt)
{
    ClusteredRun run;
    const double t_start = now();
    run.bc.assign((size_t)g.n, 0.0);
    if (g.n == 0) return run;

    double t0 = now();
    std::vector<std::vector<int>> comms;
    if (!opt.partition.empty()) {
        /* community c = the vertices whose index is c, ascending; empty
         * indices dropped; a negative index leaves the vertex in no community
         * (a partition strategy need not cover every vertex) */
        if (opt.partition.size() != (size_t)g.n) throw std::invalid_argument("partition size differs from the graph's");
        int k = 0;
        for (int c : opt.partition) k = std::max(k, c + 1);
        std::vector<std::vector<int>> by((size_t)k);
        for (int v = 0; v < g.n; ++v)
            if (opt.partition[(size_t)v] >= 0) by[(size_t)opt.partition[(size_t)v]].push_back(v);
        for (auto &c : by)
            if (!c.empty()) comms.push_back(std::move(c));
    } else {
        if (opt.progress) opt.progress(0, "Computing clusters with Louvain algorithm...");
        comms = louvain(g, opt.seeds, opt.precision);
    }
    run.t.louvain = now() - t0;
    const uint32_t ncl = (uint32_t)comms.size();
    run.t.clusters = ncl;
    auto report = [&](int level, const std::string &msg) {
        if (opt.progress) opt.progress(level, msg);
    };
    report(0, "Graph partitioned in " + std::to_string(ncl) + " clusters");
    report(0, "Evaluating intra cluster BC...");

    std::lock_guard<std::mutex> lk(dev.lock());
    fbc_graph *dg = dev.graph(g);

    t0 = now();
    std::vector<uint32_t> cptr((size_t)ncl + 1, 0), verts;
    verts.reserve((size_t)g.n);
    run.comm.assign((size_t)g.n, -1);
    for (uint32_t c = 0; c < ncl; ++c) {
        for (int v : comms[c]) {
            verts.push_back((uint32_t)v);
            run.comm[v] = (int)c;
        }
        cptr[c + 1] = (uint32_t)verts.size();
    }
    std::vector<uint32_t> bptr((size_t)ncl + 1);
    fbc_clusters *cl = nullptr;
    dev.check(fbc_clusters_create(dev.ctx(), dg, ncl, cptr.data(), verts.data(), &cl, bptr.data()));
    struct Guard {
        fbc_clusters *c;
        ~Guard() { fbc_clusters_free(c); }
    } guard{cl};
    std::vector<uint32_t> borders(bptr[ncl]);
    dev.check(fbc_clusters_borders(cl, borders.data()));
    std::vector<uint64_t> vioff((size_t)ncl + 1, 0);
    for (uint32_t c = 0; c < ncl; ++c)
        vioff[c + 1] = vioff[c] + (uint64_t)(cptr[c + 1] - cptr[c]) * (bptr[c + 1] - bptr[c]);
    const double t_views = now() - t0;
    run.intra.assign((size_t)g.n, 0.0);
    const double t_eval = now();
    /* the VertexInfo rows stay on the device for the class selection */
    dev.check(fbc_cluster_eval_resident(dev.ctx(), cl, run.intra.data()));
    fbc_last_stats(dev.ctx(), &run.t.cluster_kernel_ms, &run.t.cluster_te);
    last_tuning(dev, run.t.cluster_tuning);
    run.t.cluster = now() - t0;
    {
        char buf[200];
        std::snprintf(buf, sizeof buf,
                      "Intra-cluster BC evaluated (%.3f s: views %.3f s, evaluation %.3f s of which kernels %.3f s)",
                      run.t.cluster, t_views, now() - t_eval, run.t.cluster_kernel_ms * 1e-3);
        report(1, buf);
    }

    /* exact pivots per cluster (the selectors only read their own cluster),
     * on the device; a hash collision there sends the selection to the host */
    t0 = now();
    std::vector<std::vector<int>> piv(ncl), pw(ncl), prow(ncl);
    std::vector<uint32_t> cls_ptr((size_t)ncl + 1, 0), cls_row(std::max<size_t>(verts.size(), 1)),
        cls_w(std::max<size_t>(verts.size(), 1));
    const int crc = fbc_cluster_classes(dev.ctx(), cl, run.intra.data(), cls_ptr.data(), cls_row.data(), cls_w.data());
    if (crc != 0 && crc != FBC_E_UNSUPPORTED) dev.check(crc);
    if (crc == 0) {
        for (uint32_t c = 0; c < ncl; ++c) {
            prow[c].assign(cls_row.begin() + cls_ptr[c], cls_row.begin() + cls_ptr[c + 1]);
            pw[c].assign(cls_w.begin() + cls_ptr[c], cls_w.begin() + cls_ptr[c + 1]);
        }
        char buf[160];
        std::snprintf(buf, sizeof buf, "exact pivot classes: %u on the device (%.3f s)", cls_ptr[ncl], now() - t0);
        report(1, buf);
        if (opt.kfrac > 0) {
            kmeans_device_stage(dev, cl, ncl, bptr, opt.kfrac, prow, pw, run.t);
            std::snprintf(buf, sizeof buf, "k-means: %zu sets on the GPU, %zu on the host (%.3f s, GPU kernels %.3f s)",
                          run.t.kmeans_gpu_sets, run.t.kmeans_host_sets, run.t.kmeans, run.t.kmeans_gpu_ms * 1e-3);
            report(1, buf);
        }
    } else {
        report(1, std::string("class grouping on the host: ") + fbc_last_error(dev.ctx()));
        host_select(dev, cl, ncl, cptr, verts, bptr, borders, vioff, run.intra, opt.kfrac, prow, pw, run.t);
    }
    for (uint32_t c = 0; c < ncl; ++c)
        for (size_t p = 0; p < prow[c].size(); ++p) piv[c].push_back((int)verts[cptr[c] + prow[c][p]]);
    for (uint32_t c = 0; c < ncl; ++c) {
        if (bptr[c + 1] > bptr[c]) ++run.t.bordered;
        for (size_t p = 0; p < piv[c].size(); ++p) {
            run.pivots.push_back((uint32_t)piv[c][p]);
            run.weights.push_back((uint32_t)pw[c][p]);
            run.pivot_cluster.push_back(c);
        }
    }
    run.t.pivot_count = run.pivots.size();
    run.t.pivots = now() - t0;

    report(0, "Computing global BC from " + std::to_string(run.pivots.size()) + " pivots...");

    /* the pivot terms are additive, so the pivots go in chunks (bounded
     * launches, progress between them) */
    t0 = now();
    run.bc = run.intra;
    if (opt.form != FBC_FORM_SELECT_ONLY) {
        const size_t np = run.pivots.size();
        const size_t chunk = opt.form == FBC_FORM_PLAIN ? 256 : (size_t)1 << 16;
        for (size_t p0 = 0; p0 < np; p0 += chunk) {
            const uint32_t k = (uint32_t)std::min(chunk, np - p0);
            if (opt.form == FBC_FORM_PLAIN)
                dev.check(fbc_pivot_phase_plain(dev.ctx(), cl, run.pivots.data() + p0, run.weights.data() + p0,
                                                run.pivot_cluster.data() + p0, k, run.intra.data(),
                                                run.bc.data()));
            else
                dev.check(fbc_pivot_phase(dev.ctx(), cl, run.pivots.data() + p0, run.weights.data() + p0,
                                          run.pivot_cluster.data() + p0, k, run.intra.data(), opt.prec,
                                          run.bc.data()));
            double ms = 0, te = 0;
            fbc_last_stats(dev.ctx(), &ms, &te);
            if (p0 == 0) last_tuning(dev, run.t.pivot_tuning);
            run.t.pivot_kernel_ms += ms;
            run.t.pivot_te += te;
            if (np > chunk) report(1, "pivots done: " + std::to_string(p0 + k) + " / " + std::to_string(np));
        }
    }
    run.t.pivot_phase = now() - t0;
    run.t.total = now() - t_start;
    return run;
}

} // namespace fbchost
