/* Shared between kmeans.hip and the C-ABI layer (not a public header). */
#ifndef FBC_KMEANS_H
#define FBC_KMEANS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbcdev {

/* A batch of independent k-means problems (one per cluster); all arrays are
 * concatenations indexed through the per-set offsets.  Sets whose active flag
 * is 0 are skipped by every kernel. */
struct KmParams {
    uint32_t ncl;
    const uint32_t *P, *D, *K;          /* points, features, centroids per set */
    const uint64_t *xoff;               /* row block of the set in X (P * D doubles) */
    const uint64_t *poff;               /* per-point arrays */
    const uint64_t *koff;               /* per-centroid arrays */
    const uint64_t *moff;               /* member-sum block (K * D doubles) */
    const uint64_t *goff;               /* Gram block of the set in G (P * P doubles) */
    const int32_t *X;                   /* integer features, row-major */
    double *xnorm;                      /* squared norm of every row (exact; the Gram diagonal) */
    const double *G;                    /* Gram matrices X X^T of the active sets (exact) */
    const int32_t *weights;             /* per point */
    double *cdist;                      /* per point, initialisation */
    double *mdist;                      /* per member slot: distance to its centroid's mean */
    double *kvar;                       /* per centroid: distance from its centroid to the next */
    uint32_t *cent, *next, *cnt, *start;/* per centroid */
    uint32_t *assign, *members;         /* per point */
    double *msum;                       /* per centroid and feature: sum over members */
    int32_t *cweights;                  /* per centroid, final */
    uint32_t *active;                   /* per set: 1 while iterating */
    uint32_t *uk;                       /* per centroid slot: the set's distinct centroids, by first index */
    uint32_t *nu;                       /* per set: how many distinct centroids */
    uint32_t *firstk;                   /* per point: scratch of the distinct-centroid pass, all ~0 between uses */
    double min_variance;
};

/* Largest centroid count whose member bookkeeping fits the update kernel's LDS. */
constexpr uint32_t KM_MAX_K = 12288;

/* Gram tiles: tile t covers rows [64 * ti, +64) x [64 * tj, +64) of set s,
 * tj >= ti, packed as (s, ti, tj) triples. */
hipError_t km_launch_gram(const KmParams &p, double *G, uint32_t ntiles, const uint32_t *tiles,
                          hipStream_t s);
hipError_t km_launch_init(const KmParams &p, hipStream_t s);
/* Work tiles of one iteration (device arrays, built by the host per group):
 *   pt   (set, first point) of 256 points       -- assignment
 *   sum  (set, first slot, first feature) of 64 member slots x 256 features
 *   mt   (set, first slot) of 64 member slots    -- member distances
 *   kt   (set, first centroid) of 256 centroids  -- nearest members */
struct KmTiles {
    uint32_t npt, nsum, nmt, nkt;
    const uint32_t *pt_cl, *pt_first, *sum, *mt_cl, *mt_first, *kt_cl, *kt_first;
};
hipError_t km_launch_iteration(const KmParams &p, const KmTiles &t, uint32_t iteration,
                               uint32_t max_iteration, hipStream_t s);

} // namespace fbcdev

#endif
