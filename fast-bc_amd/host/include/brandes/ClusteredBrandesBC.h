/*
 * Clustered Brandes (brandes/ClusteredBrandesBC.h:19-169).
 *
 * ClusteredBrandeBC keeps the reference's composition of strategies: Louvain
 * partition, per-cluster evaluation and pivot selection (in parallel over
 * clusters), then one single-source Brandes per pivot:
 *     BC[v] += w_p * dep_p(v);   BC[v] -= w_p * intra(v) for v in p's cluster.
 * With the GPU-backed strategies of this library every cluster and every pivot
 * is its own launch (the plain form).
 *
 * GpuClusteredBrandesBC computes the same result with the batched path:
 * every (cluster, source) task in one persistent launch and every pivot in
 * another (fbchost::clustered_bc).
 */
#ifndef FASTBC_AMD_CLUSTEREDBRANDESBC_H
#define FASTBC_AMD_CLUSTEREDBRANDESBC_H

#include "../GraphBridge.h"
#include "../IGraphPartition.h"
#include "../SubGraph.h"
#include "GpuDevice.h"
#include "IBrandesBC.h"
#include "IClusterEvaluator.h"
#include "IPivotSelector.h"
#include "ISSBrandesBC.h"

#include <random>
#include <set>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class ClusteredBrandeBC : public IBrandesBC<V, W> {
public:
    ClusteredBrandeBC(std::shared_ptr<IGraphPartition<V, W>> gp, std::shared_ptr<IClusterEvaluator<V, W>> ce,
                      std::shared_ptr<ISSBrandesBC<V, W>> ssb, std::shared_ptr<IPivotSelector<V, W>> ps)
        : _gp(gp), _ce(ce), _ssb(ssb), _ps(ps)
    {
    }

    std::vector<W> computeBC(const std::shared_ptr<const IGraph<V, W>> graph) override
    {
        const size_t n = graph->vertices().size();
        std::vector<W> bc(n, (W)0);
        std::vector<std::shared_ptr<VertexInfo<V, W>>> vi(n, nullptr);
        std::vector<std::vector<V>> comms =
            _gp->partitionGraph(std::static_pointer_cast<const IDegreeGraph<V, W>>(graph));
        std::vector<std::shared_ptr<ISubGraph<V, W>>> cluster(comms.size());
        std::vector<std::pair<std::vector<V>, std::vector<V>>> piv(comms.size());
#pragma omp parallel for schedule(dynamic, 1)
        for (long long i = 0; i < (long long)comms.size(); ++i) {
            cluster[i] = std::make_shared<SubGraph<V, W>>(comms[i], graph);
            if (cluster[i]->borders().empty()) continue;   /* FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED */
            _ce->evaluateCluster(bc, vi, cluster[i]);
            piv[i] = _ps->selectPivots(bc, vi, cluster[i]->vertices(), cluster[i]->borders());
        }
        const std::vector<W> intra(bc);
        for (size_t c = 0; c < cluster.size(); ++c)
            for (size_t p = 0; p < piv[c].first.size(); ++p) {
                const std::vector<W> dep = _ssb->singleSourceBrandes(piv[c].first[p], graph);
                const W wp = (W)piv[c].second[p];
                for (size_t v = 0; v < n; ++v) bc[v] += dep[v] * wp;
                for (const V &v : cluster[c]->vertices()) bc[v] -= intra[v] * wp;
            }
        return bc;
    }

private:
    std::shared_ptr<IGraphPartition<V, W>> _gp;
    std::shared_ptr<IClusterEvaluator<V, W>> _ce;
    std::shared_ptr<ISSBrandesBC<V, W>> _ssb;
    std::shared_ptr<IPivotSelector<V, W>> _ps;
};

template <typename V, typename W>
class GpuClusteredBrandesBC : public IBrandesBC<V, W> {
public:
    /* kFrac <= 0: VertexInfoPivotSelector, else KMeansPivotSelector(kFrac) */
    GpuClusteredBrandesBC(const std::set<std::mt19937::result_type> &seeds, double precision = 0.01,
                          double kFrac = 0.0, int prec = FBC_PREC_F64)
    {
        _opt.seeds.assign(seeds.begin(), seeds.end());
        _opt.precision = precision;
        _opt.kfrac = kFrac;
        _opt.prec = prec;
    }

    /* The batched path with the caller's partition strategy in place of
     * Louvain, as ClusteredBrandeBC takes one (ClusteredBrandesBC.h:37-44,
     * 80-82). */
    GpuClusteredBrandesBC(std::shared_ptr<IGraphPartition<V, W>> gp, double kFrac = 0.0, int prec = FBC_PREC_F64)
        : _gp(gp)
    {
        _opt.kfrac = kFrac;
        _opt.prec = prec;
    }

    std::vector<W> computeBC(const std::shared_ptr<const IGraph<V, W>> graph) override
    {
        auto g = bridge::host_graph_ptr<V, W>(*graph);
        if (_gp) {
            const std::vector<std::vector<V>> comms =
                _gp->partitionGraph(std::static_pointer_cast<const IDegreeGraph<V, W>>(graph));
            _opt.partition.assign((size_t)g->n, -1);
            for (size_t c = 0; c < comms.size(); ++c)
                for (const V &v : comms[c]) _opt.partition.at((size_t)v) = (int)c;
        }
        _last = fbchost::clustered_bc(gpuDevice(), *g, _opt);
        return std::vector<W>(_last.bc.begin(), _last.bc.end());
    }

    /* partition, pivots and phase times of the last computeBC */
    const fbchost::ClusteredRun &lastRun() const { return _last; }

private:
    std::shared_ptr<IGraphPartition<V, W>> _gp;
    fbchost::ClusteredOptions _opt;
    fbchost::ClusteredRun _last;
};

} // namespace brandes
} // namespace fastbc

#endif
