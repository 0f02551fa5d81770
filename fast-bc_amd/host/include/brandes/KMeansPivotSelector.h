/*
 * KMeansPivotSelector (brandes/KMeansPivotSelector.h:55-132): exact pivots
 * first, then k = max((int)(P * kFrac), 1) k-means super-classes; repeated
 * centroids are dropped keeping the first occurrence and its weight.
 */
#ifndef FASTBC_AMD_KMEANSPIVOTSELECTOR_H
#define FASTBC_AMD_KMEANSPIVOTSELECTOR_H

#include "../kmeans/IKMeans.h"
#include "IPivotSelector.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>

namespace fastbc {
namespace brandes {

template <typename V, typename W>
class KMeansPivotSelector : public IPivotSelector<V, W> {
public:
    KMeansPivotSelector(std::shared_ptr<IPivotSelector<V, W>> exactPivotSelector,
                        std::shared_ptr<kmeans::IKMeans<V, W>> kmeans, double kFrac, W stopVariance = 0,
                        size_t maxIteration = 100)
        : _exact(exactPivotSelector), _kmeans(kmeans), _kFrac(kFrac), _stop(stopVariance), _maxIt(maxIteration)
    {
        if (_kFrac < 0.0 || _kFrac > 1.0) throw std::invalid_argument("Given KFrac parameter is out of bounds");
        if (maxIteration < 100)
            std::fprintf(stderr, "Given max iteration for kmeans pivot selection is low (%zu)\n", maxIteration);
    }

    std::pair<std::vector<V>, std::vector<V>> selectPivots(
        const std::vector<W> &globalBC, const std::vector<std::shared_ptr<VertexInfo<V, W>>> &verticesInfo,
        const std::vector<V> &vertices, const std::set<V> &borders) override
    {
        const auto exact = _exact->selectPivots(globalBC, verticesInfo, vertices, borders);
        const int k = std::max((int)(exact.first.size() * _kFrac), 1);
        auto pw = _kmeans->computeCentroids(k, exact.first, exact.second, verticesInfo, _stop, _maxIt);
        std::pair<std::vector<V>, std::vector<V>> out;
        std::set<V> seen;
        for (size_t i = 0; i < pw.first.size(); ++i)
            if (seen.insert(pw.first[i]).second) {
                out.first.push_back(pw.first[i]);
                out.second.push_back(pw.second[i]);
            }
        return out;
    }

private:
    std::shared_ptr<IPivotSelector<V, W>> _exact;
    std::shared_ptr<kmeans::IKMeans<V, W>> _kmeans;
    const double _kFrac;
    const W _stop;
    const size_t _maxIt;
};

} // namespace brandes
} // namespace fastbc

#endif
