/*
 * ExactBrandesBC (brandes/ExactBrandesBC.h:46-95): all-sources Brandes, every
 * source in one persistent gfx950 launch (fbc_exact_bc).
 */
#ifndef FASTBC_AMD_EXACTBRANDESBC_H
#define FASTBC_AMD_EXACTBRANDESBC_H

#include "../GraphBridge.h"
#include "GpuDevice.h"
#include "IBrandesBC.h"

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class ExactBrandesBC : public IBrandesBC<V, W> {
public:
    std::vector<W> computeBC(const std::shared_ptr<const IGraph<V, W>> graph) override
    {
        auto g = bridge::host_graph_ptr<V, W>(*graph);
        std::vector<double> bc = fbchost::exact_bc(gpuDevice(), *g, gpuPrecision());
        return std::vector<W>(bc.begin(), bc.end());
    }
};

} // namespace brandes
} // namespace fastbc

#endif
