/* IGraphPartition (IGraphPartition.h:19). */
#ifndef FASTBC_AMD_IGRAPHPARTITION_H
#define FASTBC_AMD_IGRAPHPARTITION_H

#include "IDegreeGraph.h"

#include <memory>
#include <vector>

namespace fastbc {

template <typename V, typename W>
class IGraphPartition {
public:
    virtual ~IGraphPartition() = default;
    /* communities, each an ascending vertex list, ordered by community id */
    virtual std::vector<std::vector<V>> partitionGraph(std::shared_ptr<const IDegreeGraph<V, W>> graph) = 0;
};

} // namespace fastbc

#endif
