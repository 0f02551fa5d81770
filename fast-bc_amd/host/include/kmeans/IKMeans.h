/* IKMeans (kmeans/IKMeans.h:32-38). */
#ifndef FASTBC_AMD_IKMEANS_H
#define FASTBC_AMD_IKMEANS_H

#include "../brandes/VertexInfo.h"

#include <memory>
#include <utility>
#include <vector>

namespace fastbc {
namespace kmeans {

template <typename V, typename W>
class IKMeans {
public:
    virtual ~IKMeans() = default;
    /* k centroids among 'vertices' and their weights (sum of member weights) */
    virtual std::pair<std::vector<V>, std::vector<V>> computeCentroids(
        int k, const std::vector<V> &vertices, const std::vector<V> &weights,
        const std::vector<std::shared_ptr<brandes::VertexInfo<V, W>>> &vertexInfo, W minVariance = 0,
        size_t maxIteration = 100) = 0;
};

} // namespace kmeans
} // namespace fastbc

#endif
