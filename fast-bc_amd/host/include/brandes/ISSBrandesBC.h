/* ISSBrandesBC (brandes/ISSBrandesBC.h:26-28). */
#ifndef FASTBC_AMD_ISSBRANDESBC_H
#define FASTBC_AMD_ISSBRANDESBC_H

#include "../IGraph.h"

#include <memory>
#include <vector>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class ISSBrandesBC {
public:
    virtual ~ISSBrandesBC() = default;
    /* dependency of every vertex on the shortest paths from source */
    virtual std::vector<W> singleSourceBrandes(V source, std::shared_ptr<const IGraph<V, W>> graph) = 0;
};

} // namespace brandes
} // namespace fastbc

#endif
