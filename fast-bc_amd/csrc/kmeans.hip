/*
 * k-means super-classes of pivots on gfx950: PlusPlusKMeans::computeCentroids
 * (kmeans/PlusPlusKMeans.h:70-229) for many clusters at once, with the
 * reference's results.
 *
 * Points are VertexInfo rows: feature 2j = path length to border j, 2j+1 =
 * path count (VertexInfo.h:293-307 sums, per border, the length term then the
 * count term; the feature order keeps that order).  Every feature is an
 * integer and the host checks 2 * D * max|x|^2 < 2^53, so the Gram matrix
 * G = X X^T and every point-to-point distance ||x_a||^2 + ||x_b||^2 - 2 G[a][b]
 * are exact integers in FP64 whatever the summation order.  G is computed once
 * per set; the farthest-point initialisation, the assignment and the
 * variance test then read distances from it (one coalesced row per centroid)
 * instead of recomputing dot products in every step.  Ties go to the first
 * point / centroid as in the reference.  Mean-to-member distances are not
 * integers; each is summed in the reference's sequential feature order by one
 * thread, and the nearest member is the first strict minimum in point order.
 *
 * Kernels:
 *   k_km_gram    64 x 64 tiles of G (upper triangle, mirrored)
 *   k_km_init    farthest-point initialisation (PlusPlusKMeans.h:172-207),
 *                one 1024-thread workgroup per set
 *   k_km_unique  the distinct centroids of every set, by first index (the
 *                farthest-point start and members' moves repeat points, and a
 *                repeat never wins the first-wins assignment)
 *   k_km_assign  first-wins nearest centroid of every point over the distinct
 *                centroids, 256 points per workgroup
 *   k_km_members member lists grouped by centroid, one 1024-thread
 *                workgroup per set
 *   k_km_msum    member sums per centroid (the means' numerators), coalesced
 *   k_km_dist    member-to-mean distances (PlusPlusKMeans.h:130-146)
 *   k_km_pick    nearest member to each mean
 *   k_km_finish  the variance test (PlusPlusKMeans.h:150-151), then the
 *                final weights or the next centroids
 */
#include "fbc_kmeans.h"

namespace fbcdev {

constexpr int GWG = 256;    /* Gram tile workgroup */
constexpr int GT = 64;      /* Gram tile edge */
constexpr int GD = 16;      /* features per LDS stage */
constexpr int SWG = 1024;   /* per-set workgroups */
constexpr int NWAVE = SWG / 64;
constexpr int AWG = 256;    /* points per assignment workgroup */
constexpr int SNAP_M = 64;  /* members per LDS stage of the snap */
constexpr int SNAP_F = 64;  /* features per LDS stage of the snap */
constexpr int KM_INIT_REG = 16;   /* points per thread kept in registers by the initialisation */

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

/* exact squared distance between points a and b of a set */
__device__ __forceinline__ double gram_dist(const double *Gs, const double *xn, uint32_t P, uint32_t a,
                                            uint32_t b)
{
    return (xn[a] + xn[b]) - 2.0 * Gs[(uint64_t)a * P + b];
}

/* Exclusive prefix sum over the SWG threads of a workgroup; returns the
 * thread's offset, *total the sum. */
__device__ uint32_t block_scan(uint32_t x, uint32_t *s_wave, uint32_t *total)
{
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < NWAVE ? s_wave[lane] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < NWAVE) s_wave[lane] = w;   /* inclusive wave totals */
    }
    __syncthreads();
    const uint32_t base = wave ? s_wave[wave - 1] : 0;
    *total = s_wave[NWAVE - 1];
    __syncthreads();
    return base + inc - x;
}

/* FP64 matrix-core block: D(16x16) += A(16x4) B(4x16).  Lane l supplies
 * A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[l / 16 + 4r][l % 16]
 * in element r (checked on gfx950 against a host product). */
typedef double dbl4 __attribute__((ext_vector_type(4)));

/* grid: one workgroup per 64 x 64 tile (s, ti, tj), tj >= ti.  Wave w
 * computes the 32 x 32 quarter (w / 2, w % 2) of the tile as 2 x 2 blocks of
 * FP64 MFMA; the features stream through LDS GD at a time.  Every product
 * and partial sum is an integer below 2^53 (the exactness precondition), so
 * the matrix cores' summation order gives the exact Gram entries. */
__global__ __launch_bounds__(GWG) void k_km_gram(KmParams p, double *G, const uint32_t *tiles)
{
    const uint32_t s = tiles[3 * blockIdx.x], ti = tiles[3 * blockIdx.x + 1], tj = tiles[3 * blockIdx.x + 2];
    const uint32_t P = p.P[s], D = p.D[s];
    const int32_t *X = p.X + p.xoff[s];
    double *Gs = G + p.goff[s];
    __shared__ double sa[GD][GT + 1];
    __shared__ double sb[GD][GT + 1];
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int li = lane % 16, lk = lane / 16;
    const int wr = (w / 2) * 32, wc = (w % 2) * 32;
    const uint32_t r0 = ti * GT, c0 = tj * GT;
    dbl4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (uint32_t d0 = 0; d0 < D; d0 += GD) {
        for (int e = threadIdx.x; e < GD * GT; e += GWG) {
            const int dd = e % GD, rr = e / GD;
            const uint32_t f = d0 + dd, ra = r0 + rr, rb = c0 + rr;
            sa[dd][rr] = (ra < P && f < D) ? (double)X[(uint64_t)ra * D + f] : 0.0;
            sb[dd][rr] = (rb < P && f < D) ? (double)X[(uint64_t)rb * D + f] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GD; kk += 4) {
            const double a0 = sa[kk + lk][wr + li], a1 = sa[kk + lk][wr + 16 + li];
            const double b0 = sb[kk + lk][wc + li], b1 = sb[kk + lk][wc + 16 + li];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = r0 + wr + a * 16 + lk + 4 * r, col = c0 + wc + b * 16 + li;
                if (row < P && col < P) {
                    const double v = acc[a][b][r];
                    Gs[(uint64_t)row * P + col] = v;
                    Gs[(uint64_t)col * P + row] = v;
                    if (row == col) p.xnorm[p.poff[s] + row] = v;
                }
            }
}

/* Farthest-point initialisation of one set, one workgroup.  Each thread
 * keeps the running means of its points (v = tid + j * SWG) in registers
 * when the set fits (REG), else in cdist. */
template <bool REG>
__device__ __forceinline__ void km_init_set(const KmParams &p, uint32_t c)
{
    constexpr int NJ = REG ? KM_INIT_REG : 1;
    const uint32_t P = p.P[c], K = p.K[c];
    const double *xn = p.xnorm + p.poff[c];
    const double *Gs = p.G + p.goff[c];
    double *cd = p.cdist + p.poff[c];
    uint32_t *cent = p.cent + p.koff[c];
    __shared__ double s_val[NWAVE];
    __shared__ uint32_t s_idx[NWAVE];
    __shared__ uint32_t s_last;
    double cdr[NJ], xnr[NJ];
    if (REG) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const uint32_t v = threadIdx.x + j * SWG;
            cdr[j] = 0.0;
            xnr[j] = v < P ? xn[v] : 0.0;
        }
    } else {
        for (uint32_t v = threadIdx.x; v < P; v += SWG) cd[v] = 0.0;
    }
    if (threadIdx.x == 0) {
        cent[0] = 0;
        s_last = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    for (uint32_t i = 1; i < K; ++i) {
        const uint32_t last = s_last;
        const double pr = 1.0 / i;
        const double q = 1.0 - pr;
        const double *row = Gs + (uint64_t)last * P;
        const double xl = xn[last];
        double best = -1.0;
        uint32_t bidx = 0xFFFFFFFFu;
        if (REG) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const uint32_t v = threadIdx.x + j * SWG;
                if (v < P) {
                    const double d = (xl + xnr[j]) - 2.0 * row[v];
                    cdr[j] = cdr[j] * q + d * pr;
                    if (cdr[j] > best) {   /* v ascending within a thread: first wins */
                        best = cdr[j];
                        bidx = v;
                    }
                }
            }
        } else {
            for (uint32_t v = threadIdx.x; v < P; v += SWG) {
                const double d = (xl + xn[v]) - 2.0 * row[v];
                const double nv = cd[v] * q + d * pr;
                cd[v] = nv;
                if (nv > best) {
                    best = nv;
                    bidx = v;
                }
            }
        }
        /* largest value, smallest index among equals: the reference's scan
         * from point 0 with a strict '>' */
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(best, o, 64);
            const uint32_t oi = __shfl_xor(bidx, o, 64);
            if (ov > best || (ov == best && oi < bidx)) {
                best = ov;
                bidx = oi;
            }
        }
        if (lane == 0) {
            s_val[wave] = best;
            s_idx[wave] = bidx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = s_val[0];
            uint32_t bi = s_idx[0];
            for (int w = 1; w < NWAVE; ++w)
                if (s_val[w] > b || (s_val[w] == b && s_idx[w] < bi)) {
                    b = s_val[w];
                    bi = s_idx[w];
                }
            cent[i] = bi;
            s_last = bi;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(SWG) void k_km_init(KmParams p)
{
    const uint32_t c = blockIdx.x;
    if (!p.active[c]) return;
    if (p.P[c] <= (uint32_t)SWG * KM_INIT_REG)
        km_init_set<true>(p, c);
    else
        km_init_set<false>(p, c);
}

/* The distinct centroids of each set (one workgroup per set): centroid k
 * is kept when no smaller index holds the same point; uk lists the kept
 * indices ascending, nu counts them.  A repeated centroid scores exactly as
 * its first occurrence, so the first-wins assignment never picks it and the
 * assignment may scan the distinct ones only. */
__global__ __launch_bounds__(SWG) void k_km_unique(KmParams p)
{
    const uint32_t c = blockIdx.x;
    if (!p.active[c]) return;
    const uint32_t K = p.K[c];
    const uint32_t *cent = p.cent + p.koff[c];
    uint32_t *first = p.firstk + p.poff[c];
    uint32_t *uk = p.uk + p.koff[c];
    __shared__ uint32_t s_wave[NWAVE];
    for (uint32_t k = threadIdx.x; k < K; k += SWG)
        __hip_atomic_fetch_min(&first[cent[k]], k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const uint32_t per = (K + SWG - 1) / SWG, k0 = threadIdx.x * per;
    uint32_t part = 0;
    for (uint32_t k = k0; k < k0 + per && k < K; ++k)
        part += __hip_atomic_load(&first[cent[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == k;
    uint32_t total;
    uint32_t pos = block_scan(part, s_wave, &total);
    for (uint32_t k = k0; k < k0 + per && k < K; ++k)
        if (__hip_atomic_load(&first[cent[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == k) uk[pos++] = k;
    if (threadIdx.x == 0) p.nu[c] = total;
    __syncthreads();
    /* back to ~0 for the next pass */
    for (uint32_t k = threadIdx.x; k < K; k += SWG)
        __hip_atomic_store(&first[cent[k]], 0xFFFFFFFFu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* grid: one workgroup per 256-point tile; tile_cl / tile_p0 map tiles */
__global__ __launch_bounds__(AWG) void k_km_assign(KmParams p, const uint32_t *tile_cl, const uint32_t *tile_p0)
{
    const uint32_t c = tile_cl[blockIdx.x];
    if (!p.active[c]) return;
    const uint32_t P = p.P[c], NU = p.nu[c];
    const double *xn = p.xnorm + p.poff[c];
    const double *Gs = p.G + p.goff[c];
    const uint32_t *cent = p.cent + p.koff[c];
    const uint32_t *uk = p.uk + p.koff[c];
    __shared__ uint32_t s_u[AWG], s_k[AWG];
    __shared__ double s_n[AWG];
    const uint32_t v = tile_p0[blockIdx.x] + threadIdx.x;
    const bool valid = v < P;
    double best = 0;
    uint32_t bk = 0xFFFFFFFFu;
    for (uint32_t j0 = 0; j0 < NU; j0 += AWG) {
        __syncthreads();
        if (j0 + threadIdx.x < NU) {
            const uint32_t k = uk[j0 + threadIdx.x];
            const uint32_t u = cent[k];
            s_k[threadIdx.x] = k;
            s_u[threadIdx.x] = u;
            s_n[threadIdx.x] = xn[u];
        }
        __syncthreads();
        const uint32_t kn = NU - j0 < (uint32_t)AWG ? NU - j0 : (uint32_t)AWG;
        if (!valid) continue;
        /* score ||c||^2 - 2 c.x ranks the centroids as ||x - c||^2 does (exact);
         * centroids ascending: the first of equals is kept */
#pragma unroll 8
        for (uint32_t j = 0; j < kn; ++j) {
            const double sc = s_n[j] - 2.0 * Gs[(uint64_t)s_u[j] * P + v];
            if (bk == 0xFFFFFFFFu || sc < best) {
                best = sc;
                bk = s_k[j];
            }
        }
    }
    if (valid) p.assign[p.poff[c] + v] = bk;
}

/* Member lists grouped by centroid (one workgroup per set).  The order inside
 * a group is irrelevant: sums are exact and the nearest member is chosen by
 * (distance, index). */
__global__ __launch_bounds__(SWG) void k_km_members(KmParams p)
{
    const uint32_t c = blockIdx.x;
    if (!p.active[c]) return;
    const uint32_t P = p.P[c], K = p.K[c];
    const uint32_t *asg = p.assign + p.poff[c];
    uint32_t *cnt = p.cnt + p.koff[c];
    uint32_t *start = p.start + p.koff[c];
    uint32_t *members = p.members + p.poff[c];
    __shared__ uint32_t s_cnt[KM_MAX_K];
    __shared__ uint32_t s_start[KM_MAX_K];
    __shared__ uint32_t s_wave[NWAVE];
    for (uint32_t k = threadIdx.x; k < K; k += SWG) s_cnt[k] = 0;
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < P; v += SWG) atomicAdd(&s_cnt[asg[v]], 1u);
    __syncthreads();
    /* member sums start from zero (rows of centroids without members are
     * never read) */
    {
        const uint32_t D = p.D[c];
        double *msum = p.msum + p.moff[c];
        const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
        for (uint32_t k = wave; k < K; k += NWAVE)
            if (s_cnt[k])
                for (uint32_t f = lane; f < D; f += 64) msum[(uint64_t)k * D + f] = 0.0;
    }
    {
        const uint32_t per = (K + SWG - 1) / SWG, k0 = threadIdx.x * per;
        uint32_t part = 0;
        for (uint32_t k = k0; k < k0 + per && k < K; ++k) part += s_cnt[k];
        uint32_t total;
        uint32_t run = block_scan(part, s_wave, &total);
        for (uint32_t k = k0; k < k0 + per && k < K; ++k) {
            s_start[k] = run;
            start[k] = run;
            cnt[k] = s_cnt[k];
            run += s_cnt[k];
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += SWG) s_cnt[k] = 0;
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < P; v += SWG) {
        const uint32_t k = asg[v];
        members[s_start[k] + atomicAdd(&s_cnt[k], 1u)] = v;
    }
}

/* Member sums per centroid: a workgroup takes 64 member slots (grouped by
 * centroid) x 256 features; each thread walks the slots of its feature with
 * coalesced row loads and adds every run of one centroid into msum with an
 * FP64 atomic.  The sums are integers below 2^53, so any order is exact.
 * grid: (set, first slot, first feature) triples. */
__global__ __launch_bounds__(AWG) void k_km_msum(KmParams p, const uint32_t *tiles)
{
    const uint32_t c = tiles[3 * blockIdx.x];
    if (!p.active[c]) return;
    const uint32_t m0 = tiles[3 * blockIdx.x + 1], f = tiles[3 * blockIdx.x + 2] + threadIdx.x;
    const uint32_t P = p.P[c], D = p.D[c];
    const int32_t *X = p.X + p.xoff[c];
    const uint32_t *asg = p.assign + p.poff[c];
    const uint32_t *members = p.members + p.poff[c];
    double *msum = p.msum + p.moff[c];
    __shared__ uint32_t s_v[SNAP_M], s_k[SNAP_M];
    const uint32_t nm = P - m0 < (uint32_t)SNAP_M ? P - m0 : (uint32_t)SNAP_M;
    if (threadIdx.x < nm) {
        const uint32_t v = members[m0 + threadIdx.x];
        s_v[threadIdx.x] = v;
        s_k[threadIdx.x] = asg[v];
    }
    __syncthreads();
    if (f >= D) return;
    /* the features of MSB slots are loaded together before they are summed
     * (an atomic between two loads would otherwise keep them apart); a slot
     * past nm reads slot 0's row again and is not summed */
    constexpr uint32_t MSB = 16;
    double acc = 0;
    uint32_t k = s_k[0];
    for (uint32_t i0 = 0; i0 < nm; i0 += MSB) {
        int32_t xv[MSB];
#pragma unroll
        for (uint32_t j = 0; j < MSB; ++j) xv[j] = X[(uint64_t)s_v[i0 + j < nm ? i0 + j : 0] * D + f];
#pragma unroll
        for (uint32_t j = 0; j < MSB; ++j) {
            if (i0 + j >= nm) break;
            if (s_k[i0 + j] != k) {
                unsafeAtomicAdd(&msum[(uint64_t)k * D + f], acc);
                acc = 0;
                k = s_k[i0 + j];
            }
            acc += (double)xv[j];
        }
    }
    unsafeAtomicAdd(&msum[(uint64_t)k * D + f], acc);
}

/* Distance of 64 member slots to their centroid's mean (sum / count, the
 * reference's mean), summed in the reference's sequential feature order by
 * one thread per slot; differences of 64 slots x 64 features are staged
 * through LDS with coalesced row loads.  grid: (set, first slot) pairs. */
__global__ __launch_bounds__(AWG) void k_km_dist(KmParams p, const uint32_t *tile_cl, const uint32_t *tile_m0)
{
    const uint32_t c = tile_cl[blockIdx.x];
    if (!p.active[c]) return;
    const uint32_t m0 = tile_m0[blockIdx.x];
    const uint32_t P = p.P[c], D = p.D[c];
    const int32_t *X = p.X + p.xoff[c];
    const uint32_t *asg = p.assign + p.poff[c];
    const uint32_t *cnt = p.cnt + p.koff[c];
    const uint32_t *members = p.members + p.poff[c];
    const double *msum = p.msum + p.moff[c];
    double *md = p.mdist + p.poff[c];
    __shared__ double s_d[SNAP_M][SNAP_F + 1];
    __shared__ uint32_t s_v[SNAP_M], s_k[SNAP_M];
    __shared__ double s_n[SNAP_M];
    const uint32_t nm = P - m0 < (uint32_t)SNAP_M ? P - m0 : (uint32_t)SNAP_M;
    if (threadIdx.x < nm) {
        const uint32_t v = members[m0 + threadIdx.x], k = asg[v];
        s_v[threadIdx.x] = v;
        s_k[threadIdx.x] = k;
        s_n[threadIdx.x] = (double)cnt[k];
    }
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    double sq = 0;
    for (uint32_t f0 = 0; f0 < D; f0 += SNAP_F) {
        const uint32_t f = f0 + lane;
        if (f < D) {
            /* a wave's PW members: all their loads first (a member past nm
             * reads member 0's again and is not stored) */
            constexpr uint32_t PW = SNAP_M / (AWG / 64);
            double mv[PW];
            int32_t xv[PW];
#pragma unroll
            for (uint32_t j = 0; j < PW; ++j) {
                const uint32_t mi = wave + j * (AWG / 64), mm = mi < nm ? mi : 0;
                mv[j] = msum[(uint64_t)s_k[mm] * D + f];
                xv[j] = X[(uint64_t)s_v[mm] * D + f];
            }
#pragma unroll
            for (uint32_t j = 0; j < PW; ++j) {
                const uint32_t mi = wave + j * (AWG / 64);
                if (mi < nm) s_d[mi][lane] = mv[j] / s_n[mi] - (double)xv[j];
            }
        }
        __syncthreads();
        if (threadIdx.x < nm) {
            const uint32_t nf = D - f0 < (uint32_t)SNAP_F ? D - f0 : (uint32_t)SNAP_F;
            for (uint32_t j = 0; j < nf; ++j) {
                const double d = s_d[threadIdx.x][j];
                sq += d * d;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < nm) md[m0 + threadIdx.x] = sq;
}

/* Nearest member to each mean -- smallest distance, then smallest index (the
 * reference scans members in point order with a strict '<') -- and the
 * distance from the current centroid to it.  grid: (set, first centroid). */
__global__ __launch_bounds__(AWG) void k_km_pick(KmParams p, const uint32_t *tile_cl, const uint32_t *tile_k0)
{
    const uint32_t c = tile_cl[blockIdx.x];
    if (!p.active[c]) return;
    const uint32_t k = tile_k0[blockIdx.x] + threadIdx.x;
    const uint32_t P = p.P[c], K = p.K[c];
    if (k >= K) return;
    const double *xn = p.xnorm + p.poff[c];
    const double *Gs = p.G + p.goff[c];
    const uint32_t *cent = p.cent + p.koff[c];
    const uint32_t n = p.cnt[p.koff[c] + k], b0 = p.start[p.koff[c] + k];
    const uint32_t *members = p.members + p.poff[c];
    const double *md = p.mdist + p.poff[c];
    uint32_t bv = cent[k];
    if (n) {
        /* the smallest (distance, index) pair: a minimum, so the members'
         * loads go PB at a time (a slot past n reads the first again) */
        constexpr uint32_t PB = 8;
        bv = members[b0];
        double bd = md[b0];
        for (uint32_t i0 = 1; i0 < n; i0 += PB) {
            uint32_t vv[PB];
            double dd[PB];
#pragma unroll
            for (uint32_t j = 0; j < PB; ++j) {
                const uint32_t i = i0 + j < n ? i0 + j : 0;
                vv[j] = members[b0 + i];
                dd[j] = md[b0 + i];
            }
#pragma unroll
            for (uint32_t j = 0; j < PB; ++j)
                if (dd[j] < bd || (dd[j] == bd && vv[j] < bv)) {
                    bv = vv[j];
                    bd = dd[j];
                }
        }
    }
    p.next[p.koff[c] + k] = bv;
    p.kvar[p.koff[c] + k] = gram_dist(Gs, xn, P, cent[k], bv);
}

/* The variance test (PlusPlusKMeans.h:150-151), then the final weights or the
 * next centroids (one workgroup per set). */
__global__ __launch_bounds__(AWG) void k_km_finish(KmParams p, uint32_t iteration, uint32_t max_iteration)
{
    const uint32_t c = blockIdx.x;
    if (!p.active[c]) return;
    const uint32_t K = p.K[c];
    const double *kvar = p.kvar + p.koff[c];
    __shared__ double s_var[AWG / 64];
    double var = 0;
    for (uint32_t k = threadIdx.x; k < K; k += AWG) var = kvar[k] > var ? kvar[k] : var;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(var, o, 64);
        if (ov > var) var = ov;
    }
    if (threadIdx.x % 64 == 0) s_var[threadIdx.x / 64] = var;
    __syncthreads();
    double vmax = s_var[0];
    for (int w = 1; w < AWG / 64; ++w) vmax = s_var[w] > vmax ? s_var[w] : vmax;
    const bool go_on = vmax > p.min_variance && iteration <= max_iteration;
    uint32_t *cent = p.cent + p.koff[c];
    const uint32_t *next = p.next + p.koff[c];
    if (!go_on) {
        /* final: the centroids of this assignment and their members' weights */
        const int32_t *w = p.weights + p.poff[c];
        const uint32_t *cnt = p.cnt + p.koff[c];
        const uint32_t *start = p.start + p.koff[c];
        const uint32_t *members = p.members + p.poff[c];
        int32_t *cw = p.cweights + p.koff[c];
        for (uint32_t k = threadIdx.x; k < K; k += AWG) {
            int32_t s = 0;
            for (uint32_t i = 0; i < cnt[k]; ++i) s += w[members[start[k] + i]];
            cw[k] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) p.active[c] = 0;
    } else {
        for (uint32_t k = threadIdx.x; k < K; k += AWG) cent[k] = next[k];
    }
}

hipError_t km_launch_gram(const KmParams &p, double *G, uint32_t ntiles, const uint32_t *tiles, hipStream_t s)
{
    if (!ntiles) return hipSuccess;
    hipLaunchKernelGGL(k_km_gram, dim3(ntiles), dim3(GWG), 0, s, p, G, tiles);
    return hipGetLastError();
}

hipError_t km_launch_init(const KmParams &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_km_init, dim3(p.ncl), dim3(SWG), 0, s, p);
    return hipGetLastError();
}

hipError_t km_launch_iteration(const KmParams &p, const KmTiles &t, uint32_t iteration,
                               uint32_t max_iteration, hipStream_t s)
{
    hipLaunchKernelGGL(k_km_unique, dim3(p.ncl), dim3(SWG), 0, s, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_assign, dim3(t.npt), dim3(AWG), 0, s, p, t.pt_cl, t.pt_first);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_members, dim3(p.ncl), dim3(SWG), 0, s, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_msum, dim3(t.nsum), dim3(AWG), 0, s, p, t.sum);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_dist, dim3(t.nmt), dim3(AWG), 0, s, p, t.mt_cl, t.mt_first);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_pick, dim3(t.nkt), dim3(AWG), 0, s, p, t.kt_cl, t.kt_first);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_km_finish, dim3(p.ncl), dim3(AWG), 0, s, p, iteration, max_iteration);
    return hipGetLastError();
}

} // namespace fbcdev
