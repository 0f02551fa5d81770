/*
 * VertexInfoPivotSelector (brandes/VertexInfoPivotSelector.h:29-127): classes
 * of equal normalized VertexInfo in first-appearance order; per class the
 * non-border member of smallest BC (first wins), or the first member when all
 * are borders; weight = class size.  Normalizes the VertexInfo in place.
 */
#ifndef FASTBC_AMD_VERTEXINFOPIVOTSELECTOR_H
#define FASTBC_AMD_VERTEXINFOPIVOTSELECTOR_H

#include "../../fbc_host.h"
#include "IPivotSelector.h"

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class VertexInfoPivotSelector : public IPivotSelector<V, W> {
public:
    std::pair<std::vector<V>, std::vector<V>> selectPivots(
        const std::vector<W> &globalBC, const std::vector<std::shared_ptr<VertexInfo<V, W>>> &verticesInfo,
        const std::vector<V> &vertices, const std::set<V> &borders) override
    {
        const int nv = (int)vertices.size();
        const int nb = nv ? verticesInfo[vertices[0]]->borders() : 0;
        std::vector<double> len((size_t)nv * nb), bc((size_t)nv);
        std::vector<int> cnt((size_t)nv * nb);
        std::vector<char> isb((size_t)nv);
        for (int r = 0; r < nv; ++r) {
            auto &vi = verticesInfo[vertices[r]];
            vi->normalize();
            for (int j = 0; j < nb; ++j) {
                len[(size_t)r * nb + j] = (double)vi->getBorderSPLength(j);
                cnt[(size_t)r * nb + j] = (int)vi->getBorderSPCount(j);
            }
            bc[r] = (double)globalBC[vertices[r]];
            isb[r] = borders.count(vertices[r]) ? 1 : 0;
        }
        std::vector<int> rows, w;
        fbchost::select_pivots(nv, nb, len.data(), cnt.data(), bc.data(), isb.data(), rows, w);
        std::pair<std::vector<V>, std::vector<V>> out;
        for (size_t i = 0; i < rows.size(); ++i) {
            out.first.push_back(vertices[rows[i]]);
            out.second.push_back((V)w[i]);
        }
        return out;
    }
};

} // namespace brandes
} // namespace fastbc

#endif
