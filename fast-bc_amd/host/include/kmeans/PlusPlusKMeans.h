/*
 * PlusPlusKMeans (kmeans/PlusPlusKMeans.h:70-229): deterministic farthest-point
 * initialisation, assignment to the first nearest centroid, the member nearest
 * to each mean as the new centroid; returns the centroids of the last
 * assignment with the summed weights of their members.
 */
#ifndef FASTBC_AMD_PLUSPLUSKMEANS_H
#define FASTBC_AMD_PLUSPLUSKMEANS_H

#include "../../fbc_host.h"
#include "IKMeans.h"

namespace fastbc {
namespace kmeans {

template <typename V, typename W>
class PlusPlusKMeans : public IKMeans<V, W> {
public:
    std::pair<std::vector<V>, std::vector<V>> computeCentroids(
        int k, const std::vector<V> &vertices, const std::vector<V> &weights,
        const std::vector<std::shared_ptr<brandes::VertexInfo<V, W>>> &vertexInfo, W minVariance = 0,
        size_t maxIteration = 100) override
    {
        const int P = (int)vertices.size();
        const int nb = P ? vertexInfo[vertices[0]]->borders() : 0;
        std::vector<double> L((size_t)P * nb);
        std::vector<int> C((size_t)P * nb), w(weights.begin(), weights.end());
        for (int p = 0; p < P; ++p)
            for (int j = 0; j < nb; ++j) {
                L[(size_t)p * nb + j] = (double)vertexInfo[vertices[p]]->getBorderSPLength(j);
                C[(size_t)p * nb + j] = (int)vertexInfo[vertices[p]]->getBorderSPCount(j);
            }
        std::vector<int> cent, cw;
        fbchost::kmeans(k, P, nb, L.data(), C.data(), w.data(), (double)minVariance, maxIteration, cent, cw);
        std::pair<std::vector<V>, std::vector<V>> out;
        for (size_t i = 0; i < cent.size(); ++i) {
            out.first.push_back(vertices[cent[i]]);
            out.second.push_back((V)cw[i]);
        }
        return out;
    }
};

} // namespace kmeans
} // namespace fastbc

#endif
