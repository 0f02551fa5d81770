/*
 * Orchestration of the clustered betweenness path on one GPU
 * (ClusteredBrandeBC::computeBC, brandes/ClusteredBrandesBC.h:63-169):
 *
 *   1. Louvain on the host (bit-identical partition);
 *   2. cluster views + every (cluster, source) Brandes in one persistent
 *      launch (fbc_cluster_eval): intra-cluster BC in the reference's FP64
 *      order, VertexInfo rows;
 *   3. pivot selection per cluster on the host (OpenMP over clusters);
 *   4. every pivot's full-graph Brandes in one persistent launch
 *      (fbc_pivot_phase), the weighted accumulation and the intra-cluster
 *      correction in FP64.
 *
 * Borderless clusters are skipped as with FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED
 * (main.cpp:6, ClusteredBrandesBC.h:99-120).
 */
#include "fbc_host.h"

#include <chrono>
#include <cmath>
#include <stdexcept>

namespace fbchost {

namespace {

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

Device::Device(int device)
{
    const int rc = fbc_create(device, &ctx_);
    if (rc) throw std::runtime_error("fbc_create failed (status " + std::to_string(rc) +
                                     "): a gfx950 (MI355X) device is required");
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

ClusteredRun clustered_bc(Device &dev, const Graph &g, const ClusteredOptions &opt)
{
    ClusteredRun run;
    const double t_start = now();
    run.bc.assign((size_t)g.n, 0.0);
    if (g.n == 0) return run;

    double t0 = now();
    const std::vector<std::vector<int>> comms = louvain(g, opt.seeds, opt.precision);
    run.t.louvain = now() - t0;
    const uint32_t ncl = (uint32_t)comms.size();
    run.t.clusters = ncl;

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
    std::vector<uint64_t> vl(vioff[ncl]), vc(vioff[ncl]);
    run.intra.assign((size_t)g.n, 0.0);
    dev.check(fbc_cluster_eval(dev.ctx(), cl, run.intra.data(), vl.data(), vc.data()));
    fbc_last_stats(dev.ctx(), &run.t.cluster_kernel_ms, &run.t.cluster_te);
    run.t.cluster = now() - t0;

    /* pivots per cluster (the selectors only read their own cluster) */
    t0 = now();
    std::vector<std::vector<int>> piv(ncl), pw(ncl);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t c = 0; c < ncl; ++c) {
        const int nv = (int)(cptr[c + 1] - cptr[c]);
        const int nb = (int)(bptr[c + 1] - bptr[c]);
        if (nb == 0) continue;
        std::vector<double> len((size_t)nv * nb);
        std::vector<int> cnt((size_t)nv * nb);
        for (size_t i = 0; i < (size_t)nv * nb; ++i) {
            len[i] = (double)vl[vioff[c] + i];
            cnt[i] = vi_count_from_sigma(vc[vioff[c] + i]);
        }
        std::vector<char> isb((size_t)nv, 0);
        for (int j = 0; j < nb; ++j) {
            const uint32_t b = borders[bptr[c] + j];
            /* borders are ascending, like the cluster's vertex list */
            const uint32_t *first = &verts[cptr[c]], *last = &verts[cptr[c + 1]];
            isb[std::lower_bound(first, last, b) - first] = 1;
        }
        std::vector<double> bcr((size_t)nv);
        for (int r = 0; r < nv; ++r) bcr[r] = run.intra[verts[cptr[c] + r]];
        std::vector<int> rows, w;
        select_pivots(nv, nb, len.data(), cnt.data(), bcr.data(), isb.data(), rows, w);
        if (opt.kfrac > 0) {
            std::vector<int> kr, kw;
            kmeans_pivots(rows, w, nb, len.data(), cnt.data(), opt.kfrac, kr, kw);
            rows.swap(kr);
            w.swap(kw);
        }
        for (size_t p = 0; p < rows.size(); ++p) {
            piv[c].push_back((int)verts[cptr[c] + rows[p]]);
            pw[c].push_back(w[p]);
        }
    }
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

    t0 = now();
    run.bc = run.intra;
    dev.check(fbc_pivot_phase(dev.ctx(), cl, run.pivots.data(), run.weights.data(), run.pivot_cluster.data(),
                              (uint32_t)run.pivots.size(), run.intra.data(), opt.prec, run.bc.data()));
    fbc_last_stats(dev.ctx(), &run.t.pivot_kernel_ms, &run.t.pivot_te);
    run.t.pivot_phase = now() - t0;
    run.t.total = now() - t_start;
    return run;
}

} // namespace fbchost
