/* IPivotSelector (brandes/IPivotSelector.h:34-38): normalizes the VertexInfo
 * of the given vertices in place; returns (pivots, class cardinalities). */
#ifndef FASTBC_AMD_IPIVOTSELECTOR_H
#define FASTBC_AMD_IPIVOTSELECTOR_H

#include "VertexInfo.h"

#include <memory>
#include <set>
#include <utility>
#include <vector>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class IPivotSelector {
public:
    virtual ~IPivotSelector() = default;
    virtual std::pair<std::vector<V>, std::vector<V>> selectPivots(
        const std::vector<W> &globalBC, const std::vector<std::shared_ptr<VertexInfo<V, W>>> &verticesInfo,
        const std::vector<V> &vertices, const std::set<V> &borders) = 0;
};

} // namespace brandes
} // namespace fastbc

#endif
