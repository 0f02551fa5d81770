/* ISubGraph (ISubGraph.h:21-40): a cluster view with its border vertices. */
#ifndef FASTBC_AMD_ISUBGRAPH_H
#define FASTBC_AMD_ISUBGRAPH_H

#include "IGraph.h"

#include <memory>
#include <set>

namespace fastbc {

template <typename V, typename W>
class ISubGraph : public IGraph<V, W> {
public:
    /* vertices with an arc leaving or entering the vertex set, ascending */
    virtual const std::set<V> &borders() const = 0;
    virtual bool isBorder(V vertex) const = 0;
    virtual std::shared_ptr<const IGraph<V, W>> referenceGraph() const = 0;
};

} // namespace fastbc

#endif
