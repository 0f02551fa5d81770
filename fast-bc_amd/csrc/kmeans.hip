/*
 * k-means super-classes of pivots on gfx950: PlusPlusKMeans::computeCentroids
 * (kmeans/PlusPlusKMeans.h:70-229) for many clusters at once, with the
 * reference's results.
 *
 * Points are VertexInfo rows: feature 2j = path length to border j, 2j+1 =
 * path count (VertexInfo.h:293-307 sums, per border, the length term then the
 * count term; the feature order keeps that order).  Point-to-point distances
 * are exact integers when every |feature| is small enough (checked by the
 * host: 2 * D * max^2 < 2^53), so the assignment step ranks centroids by
 * ||c||^2 - 2 c.x, an exact integer in FP64 whatever the summation order, and
 * ties go to the first centroid as in the reference.  Mean-to-member
 * distances are not integers; they are summed in the reference's sequential
 * order by one thread each.
 *
 * Kernels (one workgroup per cluster unless noted):
 *   k_km_init     farthest-point initialisation (PlusPlusKMeans.h:172-207)
 *   k_km_assign   tiled point x centroid dot products + first-wins argmin
 *                 (one workgroup per 64-point tile)
 *   k_km_update   members, means, nearest member to each mean, the variance
 *                 test (PlusPlusKMeans.h:115-151) and the final weights
 */
#include "fbc_kmeans.h"

namespace fbcdev {

constexpr int KWG = 256;
constexpr int TP = 64;      /* points per assignment tile */
constexpr int TC = 64;      /* centroids per inner tile */
constexpr int TD = 16;      /* features per LDS stage */

/* sum over features of (a - b)^2 in feature order */
__device__ __forceinline__ double sq_rows(const double *a, const double *b, uint32_t D)
{
    double sq = 0;
    for (uint32_t j = 0; j < D; ++j) {
        const double d = a[j] - b[j];
        sq += d * d;
    }
    return sq;
}

__global__ __launch_bounds__(KWG) void k_km_init(KmParams p)
{
    const uint32_t c = blockIdx.x;
    const uint32_t P = p.P[c], D = p.D[c], K = p.K[c];
    const double *X = p.X + p.xoff[c];
    double *cd = p.cdist + p.poff[c];
    uint32_t *cent = p.cent + p.koff[c];
    __shared__ double s_val[KWG];
    __shared__ uint32_t s_idx[KWG];
    for (uint32_t v = threadIdx.x; v < P; v += KWG) cd[v] = 0.0;
    if (threadIdx.x == 0) cent[0] = 0;
    __syncthreads();
    for (uint32_t i = 1; i < K; ++i) {
        const uint32_t last = cent[i - 1];
        const double pr = 1.0 / i;
        const double q = 1.0 - pr;
        double best = -1.0;
        uint32_t bidx = 0xFFFFFFFFu;
        for (uint32_t v = threadIdx.x; v < P; v += KWG) {
            const double d = sq_rows(X + (uint64_t)last * D, X + (uint64_t)v * D, D);
            const double nv = cd[v] * q + d * pr;
            cd[v] = nv;
            if (nv > best) {            /* v ascending within a thread: first wins */
                best = nv;
                bidx = v;
            }
        }
        s_val[threadIdx.x] = best;
        s_idx[threadIdx.x] = bidx;
        __syncthreads();
        for (int s = KWG / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                const double ov = s_val[threadIdx.x + s];
                const uint32_t oi = s_idx[threadIdx.x + s];
                if (ov > s_val[threadIdx.x] || (ov == s_val[threadIdx.x] && oi < s_idx[threadIdx.x])) {
                    s_val[threadIdx.x] = ov;
                    s_idx[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        /* largest value, smallest index among equals: the reference's scan
         * from point 0 with a strict '>' */
        if (threadIdx.x == 0) cent[i] = s_idx[0];
        __syncthreads();
    }
}

/* grid: one workgroup per 64-point tile; tile_cl / tile_p0 map tiles */
__global__ __launch_bounds__(KWG) void k_km_assign(KmParams p, const uint32_t *tile_cl, const uint32_t *tile_p0)
{
    const uint32_t c = tile_cl[blockIdx.x];
    if (!p.active[c]) return;
    const uint32_t p0 = tile_p0[blockIdx.x];
    const uint32_t P = p.P[c], D = p.D[c], K = p.K[c];
    const double *X = p.X + p.xoff[c];
    const double *xn = p.xnorm + p.poff[c];
    const uint32_t *cent = p.cent + p.koff[c];
    __shared__ double sx[TD][TP + 1];
    __shared__ double sc[TD][TC + 1];
    __shared__ double s_best[16][TP];
    __shared__ uint32_t s_bidx[16][TP];
    const int tx = threadIdx.x % 16;   /* centroid sub-block */
    const int ty = threadIdx.x / 16;   /* point sub-block */
    double best[4];
    uint32_t bidx[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        best[a] = 0;
        bidx[a] = 0xFFFFFFFFu;
    }
    for (uint32_t c0 = 0; c0 < K; c0 += TC) {
        double acc[4][4] = {};
        for (uint32_t d0 = 0; d0 < D; d0 += TD) {
            for (int e = threadIdx.x; e < TD * TP; e += KWG) {
                const int dd = e / TP, pp = e % TP;
                const uint32_t v = p0 + pp, f = d0 + dd;
                sx[dd][pp] = (v < P && f < D) ? X[(uint64_t)v * D + f] : 0.0;
            }
            for (int e = threadIdx.x; e < TD * TC; e += KWG) {
                const int dd = e / TC, cc = e % TC;
                const uint32_t k = c0 + cc, f = d0 + dd;
                sc[dd][cc] = (k < K && f < D) ? X[(uint64_t)cent[k] * D + f] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int dd = 0; dd < TD; ++dd) {
                double xv[4], cv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    xv[a] = sx[dd][ty * 4 + a];
                    cv[a] = sc[dd][tx * 4 + a];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_fma(xv[a], cv[b], acc[a][b]);
            }
            __syncthreads();
        }
        /* score = ||c||^2 - 2 c.x (exact); centroids ascending within a thread */
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t k = c0 + tx * 4 + b;
            if (k >= K) continue;
            const double cn = xn[cent[k]];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double s = cn - 2.0 * acc[a][b];
                if (bidx[a] == 0xFFFFFFFFu || s < best[a]) {
                    best[a] = s;
                    bidx[a] = k;
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        s_best[tx][ty * 4 + a] = best[a];
        s_bidx[tx][ty * 4 + a] = bidx[a];
    }
    __syncthreads();
    if (threadIdx.x < TP) {
        const uint32_t v = p0 + threadIdx.x;
        if (v < P) {
            double b = 0;
            uint32_t bi = 0xFFFFFFFFu;
            for (int t = 0; t < 16; ++t) {
                const uint32_t i = s_bidx[t][threadIdx.x];
                if (i == 0xFFFFFFFFu) continue;
                const double s = s_best[t][threadIdx.x];
                if (bi == 0xFFFFFFFFu || s < b || (s == b && i < bi)) {
                    b = s;
                    bi = i;
                }
            }
            p.assign[p.poff[c] + v] = bi;
        }
    }
}

__global__ __launch_bounds__(KWG) void k_km_update(KmParams p, uint32_t iteration, uint32_t max_iteration)
{
    const uint32_t c = blockIdx.x;
    if (!p.active[c]) return;
    const uint32_t P = p.P[c], D = p.D[c], K = p.K[c];
    const double *X = p.X + p.xoff[c];
    const uint32_t *asg = p.assign + p.poff[c];
    uint32_t *cent = p.cent + p.koff[c];
    uint32_t *next = p.next + p.koff[c];
    uint32_t *cnt = p.cnt + p.koff[c];
    uint32_t *start = p.start + p.koff[c];
    uint32_t *members = p.members + p.poff[c];
    double *mean = p.mean + p.moff[c];
    __shared__ double s_var[KWG];

    /* members of each centroid in point order (stable counting sort) */
    if (threadIdx.x == 0) {
        for (uint32_t k = 0; k < K; ++k) cnt[k] = 0;
        for (uint32_t v = 0; v < P; ++v) cnt[asg[v]]++;
        uint32_t run = 0;
        for (uint32_t k = 0; k < K; ++k) {
            start[k] = run;
            run += cnt[k];
        }
        for (uint32_t k = 0; k < K; ++k) cnt[k] = 0;
        for (uint32_t v = 0; v < P; ++v) {
            const uint32_t k = asg[v];
            members[start[k] + cnt[k]++] = v;
        }
    }
    __syncthreads();
    /* means: exact integer sums divided once by the member count */
    for (uint64_t e = threadIdx.x; e < (uint64_t)K * D; e += KWG) {
        const uint32_t k = (uint32_t)(e / D), f = (uint32_t)(e % D);
        const uint32_t n = cnt[k];
        if (!n) continue;
        double s = 0;
        for (uint32_t i = 0; i < n; ++i) s += X[(uint64_t)members[start[k] + i] * D + f];
        mean[e] = s / (double)n;
    }
    __syncthreads();
    /* nearest member to each mean, reference order and first-wins */
    for (uint32_t k = threadIdx.x; k < K; k += KWG) {
        const uint32_t n = cnt[k];
        if (!n) {
            next[k] = cent[k];
            continue;
        }
        const double *m = mean + (uint64_t)k * D;
        uint32_t bv = members[start[k]];
        double bd = sq_rows(m, X + (uint64_t)bv * D, D);
        for (uint32_t i = 1; i < n; ++i) {
            const uint32_t v = members[start[k] + i];
            const double d = sq_rows(m, X + (uint64_t)v * D, D);
            if (d < bd) {
                bv = v;
                bd = d;
            }
        }
        next[k] = bv;
    }
    __syncthreads();
    double var = 0;
    for (uint32_t k = threadIdx.x; k < K; k += KWG) {
        const double d = sq_rows(X + (uint64_t)cent[k] * D, X + (uint64_t)next[k] * D, D);
        if (d > var) var = d;
    }
    s_var[threadIdx.x] = var;
    __syncthreads();
    for (int s = KWG / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && s_var[threadIdx.x + s] > s_var[threadIdx.x]) s_var[threadIdx.x] = s_var[threadIdx.x + s];
        __syncthreads();
    }
    const bool go_on = s_var[0] > p.min_variance && iteration <= max_iteration;
    if (!go_on) {
        /* final: the centroids of this assignment and their members' weights */
        const int32_t *w = p.weights + p.poff[c];
        int32_t *cw = p.cweights + p.koff[c];
        for (uint32_t k = threadIdx.x; k < K; k += KWG) {
            int32_t s = 0;
            for (uint32_t i = 0; i < cnt[k]; ++i) s += w[members[start[k] + i]];
            cw[k] = s;
        }
        if (threadIdx.x == 0) p.active[c] = 0;
    } else {
        for (uint32_t k = threadIdx.x; k < K; k += KWG) cent[k] = next[k];
    }
}

hipError_t km_launch_init(const KmParams &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_km_init, dim3(p.ncl), dim3(KWG), 0, s, p);
    return hipGetLastError();
}

hipError_t km_launch_iteration(const KmParams &p, uint32_t ntiles, const uint32_t *tile_cl,
                               const uint32_t *tile_p0, uint32_t iteration, uint32_t max_iteration,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_km_assign, dim3(ntiles), dim3(KWG), 0, s, p, tile_cl, tile_p0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_update, dim3(p.ncl), dim3(KWG), 0, s, p, iteration, max_iteration);
    return hipGetLastError();
}

} // namespace fbcdev
