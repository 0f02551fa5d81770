/*
 * LouvainGraphPartition (louvain/LouvainGraphPartition.h:15-131): one Louvain
 * instance per seed (ascending), partition bit-identical to the reference's;
 * the engine is fbchost::louvain.
 */
#ifndef FASTBC_AMD_LOUVAINGRAPHPARTITION_H
#define FASTBC_AMD_LOUVAINGRAPHPARTITION_H

#include "../IGraphPartition.h"
#include "../GraphBridge.h"

#include <random>
#include <set>

namespace fastbc {
namespace louvain {

template <typename V, typename W>
class LouvainGraphPartition : public IGraphPartition<V, W> {
public:
    LouvainGraphPartition(const std::set<std::mt19937::result_type> &seeds, double precision = 0.01)
        : _seeds(seeds.begin(), seeds.end()), _precision(precision)
    {
    }

    std::vector<std::vector<V>> partitionGraph(std::shared_ptr<const IDegreeGraph<V, W>> graph) override
    {
        const fbchost::Graph &g = bridge::host_graph<V, W>(*graph);
        auto comms = fbchost::louvain(g, _seeds, _precision);
        std::vector<std::vector<V>> r(comms.size());
        for (size_t c = 0; c < comms.size(); ++c) r[c].assign(comms[c].begin(), comms[c].end());
        return r;
    }

private:
    std::vector<uint64_t> _seeds;
    double _precision;
};

} // namespace louvain
} // namespace fastbc

#endif
