/* IBrandesBC (brandes/IBrandesBC.h:24). */
#ifndef FASTBC_AMD_IBRANDESBC_H
#define FASTBC_AMD_IBRANDESBC_H

#include "../IGraph.h"

#include <memory>
#include <vector>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class IBrandesBC {
public:
    virtual ~IBrandesBC() = default;
    /* betweenness of every vertex of a complete graph (ids 0..n-1) */
    virtual std::vector<W> computeBC(const std::shared_ptr<const IGraph<V, W>> graph) = 0;
};

} // namespace brandes
} // namespace fastbc

#endif
