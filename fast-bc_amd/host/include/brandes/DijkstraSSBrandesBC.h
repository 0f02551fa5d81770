/*
 * DijkstraSSBrandesBC (brandes/DijkstraSSBrandesBC.h:45-147): the dependency
 * of every vertex on the shortest paths from one source, one GPU launch per
 * call (the plain form).  Calls from several threads are serialized.
 */
#ifndef FASTBC_AMD_DIJKSTRASSBRANDESBC_H
#define FASTBC_AMD_DIJKSTRASSBRANDESBC_H

#include "../GraphBridge.h"
#include "GpuDevice.h"
#include "ISSBrandesBC.h"

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class DijkstraSSBrandesBC : public ISSBrandesBC<V, W> {
public:
    std::vector<W> singleSourceBrandes(V source, std::shared_ptr<const IGraph<V, W>> graph) override
    {
        auto g = bridge::host_graph_ptr<V, W>(*graph);
        std::vector<double> d = fbchost::single_source_dep(gpuDevice(), *g, (int)source, gpuPrecision());
        return std::vector<W>(d.begin(), d.end());
    }
};

} // namespace brandes
} // namespace fastbc

#endif
