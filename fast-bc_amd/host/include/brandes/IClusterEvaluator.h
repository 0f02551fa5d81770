/* IClusterEvaluator (brandes/IClusterEvaluator.h:33-36): clusterBC and
 * globalVI are sized to the whole graph; only the cluster's entries change and
 * BC is added. */
#ifndef FASTBC_AMD_ICLUSTEREVALUATOR_H
#define FASTBC_AMD_ICLUSTEREVALUATOR_H

#include "../ISubGraph.h"
#include "VertexInfo.h"

#include <memory>
#include <vector>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class IClusterEvaluator {
public:
    virtual ~IClusterEvaluator() = default;
    virtual void evaluateCluster(std::vector<W> &clusterBC,
                                 std::vector<std::shared_ptr<VertexInfo<V, W>>> &globalVI,
                                 std::shared_ptr<const ISubGraph<V, W>> cluster) = 0;
};

} // namespace brandes
} // namespace fastbc

#endif
