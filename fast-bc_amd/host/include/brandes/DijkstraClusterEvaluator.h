/*
 * DijkstraClusterEvaluator (brandes/DijkstraClusterEvaluator.h:50-180): every
 * vertex of the cluster as a source on the cluster's own arcs, in one GPU
 * launch; intra-cluster BC is added to clusterBC in the reference's FP64 order
 * and a VertexInfo (path length and (int) count to each ascending border,
 * length 0 when unreached) is stored for every cluster vertex.
 */
#ifndef FASTBC_AMD_DIJKSTRACLUSTEREVALUATOR_H
#define FASTBC_AMD_DIJKSTRACLUSTEREVALUATOR_H

#include "../GraphBridge.h"
#include "GpuDevice.h"
#include "IClusterEvaluator.h"

#include <algorithm>
#include <stdexcept>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class DijkstraClusterEvaluator : public IClusterEvaluator<V, W> {
public:
    void evaluateCluster(std::vector<W> &clusterBC, std::vector<std::shared_ptr<VertexInfo<V, W>>> &globalVI,
                         std::shared_ptr<const ISubGraph<V, W>> cluster) override
    {
        const fbchost::Graph &g = bridge::host_graph<V, W>(*cluster->referenceGraph());
        const std::vector<V> &vs = cluster->vertices();
        if (!std::is_sorted(vs.begin(), vs.end()))
            throw std::invalid_argument("cluster vertices must be ascending (as Louvain communities are)");
        std::vector<uint32_t> ptr = {0, (uint32_t)vs.size()}, verts(vs.begin(), vs.end());
        fbchost::Device &dev = gpuDevice();
        std::lock_guard<std::mutex> lk(dev.lock());
        fbc_graph *dg = dev.graph(g);
        uint32_t bptr[2];
        fbc_clusters *cl = nullptr;
        dev.check(fbc_clusters_create(dev.ctx(), dg, 1, ptr.data(), verts.data(), &cl, bptr));
        const uint32_t nb = bptr[1];
        std::vector<uint32_t> len((size_t)vs.size() * nb);
        std::vector<int32_t> cnt((size_t)vs.size() * nb);
        /* the cluster's sum starts from zero and is then added to clusterBC,
         * as the reference's private reduction copy is (DijkstraClusterEvaluator.h:66) */
        std::vector<double> bc(clusterBC.size(), 0.0);
        const int rc = fbc_cluster_eval(dev.ctx(), cl, bc.data(), len.data(), cnt.data());
        fbc_clusters_free(cl);
        dev.check(rc);
        for (size_t r = 0; r < vs.size(); ++r) {
            clusterBC[vs[r]] += (W)bc[vs[r]];
            auto vi = std::make_shared<VertexInfo<V, W>>((int)nb);
            for (uint32_t j = 0; j < nb; ++j) {
                vi->setBorderSPLength((int)j, (W)len[r * nb + j]);
                vi->setBorderSPCount((int)j, (V)cnt[r * nb + j]);
            }
            globalVI[vs[r]] = vi;
        }
    }
};

} // namespace brandes
} // namespace fastbc

#endif
