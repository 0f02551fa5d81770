/*
 * Exact pivot classes on gfx950: VertexInfoPivotSelector::selectPivots
 * (VertexInfoPivotSelector.h:28-127) for every cluster at once, on the
 * VertexInfo rows the cluster evaluation leaves in device memory, so that the
 * rows (one entry per vertex and border, gigabytes on large graphs) never
 * cross PCIe.
 *
 * The reference compares every vertex's normalized row (lengths minus their
 * minimum, counts) with the representative of every class found so far.
 * Here each row gets a 64-bit hash of its normalized entries and cluster; a
 * stable radix sort of (hash, row) puts the members of a class next to each
 * other in ascending row order, and each run of equal hashes is checked entry
 * by entry against its first row.  A mismatch (a hash collision) is reported
 * and the caller falls back to the host selector, so the classes are always
 * the reference's: same members, same order of first appearance (classes are
 * emitted in the order of their first rows), same pivot (the first non-border
 * member with the smallest intra-cluster BC, else the first member) and the
 * class size as weight.
 *
 * Kernels:
 *   k_class_hash     one wavefront per row: minimum length, then the sum of
 *                    per-entry hashes (order-free, so lanes combine freely)
 *   k_class_borders  border flags per row
 *   k_class_runs     one thread per run start: verification, size, pivot
 *   k_class_compact  class list in row order through a prefix sum
 *   k_class_fmax     largest |feature| per cluster (the k-means exactness test)
 *   k_class_gather   k-means points of the selected clusters
 */
#include "fbc_classes.h"

#include <hipcub/hipcub.hpp>

namespace fbcdev {

namespace {

constexpr int CWG = 256;
constexpr int CWAVES = CWG / 64;

__device__ __forceinline__ uint64_t smix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x)
{
    for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o, 64));
    return x;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x)
{
    for (int o = 32; o > 0; o >>= 1) x += (uint64_t)__shfl_xor((unsigned long long)x, o, 64);
    return x;
}

__device__ __forceinline__ uint64_t wave_max64(uint64_t x)
{
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t y = (uint64_t)__shfl_xor((unsigned long long)x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

__global__ __launch_bounds__(CWG) void k_class_hash(RowSet rs, uint64_t key_mask, uint64_t *key, uint32_t *val,
                                                    uint32_t *rmin)
{
    const uint64_t r = (uint64_t)blockIdx.x * CWAVES + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (r >= rs.nrows) return;
    const uint32_t c = rs.vcl[r];
    const uint32_t nb = rs.bptr[c + 1] - rs.bptr[c];
    if (nb == 0) {
        if (lane == 0) {
            key[r] = CLASS_KEY_NONE;
            val[r] = (uint32_t)r;
            rmin[r] = 0;
        }
        return;
    }
    const uint64_t base = rs.vioff[c] + (r - rs.voff[c]) * nb;
    uint32_t mn = 0xFFFFFFFFu;
    for (uint32_t j = lane; j < nb; j += 64) mn = min(mn, rs.vi_len[base + j]);
    mn = wave_min(mn);
    uint64_t h = 0;
    for (uint32_t j = lane; j < nb; j += 64) {
        const uint64_t x = (uint64_t)(rs.vi_len[base + j] - mn) | ((uint64_t)j << 32);
        const uint64_t y = (uint64_t)(uint32_t)rs.vi_cnt[base + j];
        h += smix(smix(x) ^ (y * 0x9e3779b97f4a7c15ull));
    }
    h = wave_sum64(h);
    if (lane == 0) {
        uint64_t k = (h + smix(0x5bd1e995ull + c)) & key_mask;
        if (k == CLASS_KEY_NONE) k = CLASS_KEY_NONE - 1;
        key[r] = k;
        val[r] = (uint32_t)r;
        rmin[r] = mn;
    }
}

__global__ __launch_bounds__(CWG) void k_class_borders(RowSet rs, uint8_t *is_border)
{
    const uint32_t c = blockIdx.x;
    const uint32_t b0 = rs.bptr[c], b1 = rs.bptr[c + 1];
    for (uint32_t k = b0 + threadIdx.x; k < b1; k += CWG) is_border[rs.voff[c] + rs.blocal[k]] = 1;
}

__global__ __launch_bounds__(CWG) void k_class_runs(RowSet rs, const uint64_t *skey, const uint32_t *sval,
                                                    const uint32_t *rmin, const uint8_t *is_border,
                                                    const double *intra, uint32_t *is_head, uint32_t *size,
                                                    uint32_t *pivot, uint32_t *collision)
{
    const uint64_t p = (uint64_t)blockIdx.x * CWG + threadIdx.x;
    if (p >= rs.nrows) return;
    const uint64_t k = skey[p];
    if (k == CLASS_KEY_NONE) return;
    if (p > 0 && skey[p - 1] == k) return;   /* not the first row of its run */
    const uint32_t h = sval[p];
    const uint32_t c = rs.vcl[h];
    const uint32_t nb = rs.bptr[c + 1] - rs.bptr[c];
    const uint64_t hb = rs.vioff[c] + (h - rs.voff[c]) * nb;
    const uint32_t hmin = rmin[h];
    uint32_t count = 0, best = 0xFFFFFFFFu;
    double best_bc = 0.0;
    for (uint64_t q = p; q < rs.nrows && skey[q] == k; ++q) {
        const uint32_t r = sval[q];
        if (q > p) {
            bool eq = rs.vcl[r] == c;
            if (eq) {
                const uint64_t rb = rs.vioff[c] + (r - rs.voff[c]) * nb;
                const uint32_t m = rmin[r];
                for (uint32_t j = 0; j < nb && eq; ++j)
                    eq = rs.vi_cnt[rb + j] == rs.vi_cnt[hb + j] && rs.vi_len[rb + j] - m == rs.vi_len[hb + j] - hmin;
            }
            if (!eq) {
                atomicOr(collision, 1u);
                return;
            }
        }
        ++count;
        if (!is_border[r]) {
            const double b = intra[rs.gid[r]];
            if (best == 0xFFFFFFFFu || b < best_bc) {
                best = r;
                best_bc = b;
            }
        }
    }
    is_head[h] = 1;
    size[h] = count;
    pivot[h] = best == 0xFFFFFFFFu ? h : best;
}

__global__ __launch_bounds__(CWG) void k_class_compact(RowSet rs, const uint32_t *is_head, const uint32_t *pos,
                                                       const uint32_t *size, const uint32_t *pivot,
                                                       uint32_t *cls_row, uint32_t *cls_size,
                                                       uint32_t *cls_cluster, uint32_t *cls_grow,
                                                       uint32_t *cls_ptr)
{
    const uint64_t r = (uint64_t)blockIdx.x * CWG + threadIdx.x;
    if (r < rs.nrows && is_head[r]) {
        const uint32_t i = pos[r], c = rs.vcl[r];
        cls_row[i] = (uint32_t)(pivot[r] - rs.voff[c]);
        cls_size[i] = size[r];
        cls_cluster[i] = c;
        cls_grow[i] = pivot[r];
    }
    if (r <= rs.ncl) {
        const uint32_t total = rs.nrows ? pos[rs.nrows - 1] + is_head[rs.nrows - 1] : 0;
        const uint64_t first = r < rs.ncl ? rs.voff[r] : rs.nrows;
        cls_ptr[r] = first < rs.nrows ? pos[first] : total;
    }
}

__global__ __launch_bounds__(CWG) void k_class_fmax(RowSet rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                                    const uint32_t *cls_grow, const uint32_t *rmin,
                                                    unsigned long long *fmax)
{
    const uint64_t i = (uint64_t)blockIdx.x * CWAVES + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (i >= nclasses) return;
    const uint32_t c = cls_cluster[i], r = cls_grow[i];
    const uint32_t nb = rs.bptr[c + 1] - rs.bptr[c];
    const uint64_t base = rs.vioff[c] + (r - rs.voff[c]) * nb;
    const uint32_t m = rmin[r];
    uint64_t mx = 0;
    for (uint32_t j = lane; j < nb; j += 64) {
        const uint64_t l = rs.vi_len[base + j] - m;
        const int64_t cn = rs.vi_cnt[base + j];
        const uint64_t a = (uint64_t)(cn < 0 ? -cn : cn);
        mx = l > mx ? l : mx;
        mx = a > mx ? a : mx;
    }
    mx = wave_max64(mx);
    if (lane == 0 && mx) atomicMax(&fmax[c], (unsigned long long)mx);
}

__global__ __launch_bounds__(CWG) void k_class_gather(RowSet rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                                      const uint32_t *cls_grow, const uint32_t *rmin,
                                                      const uint32_t *cls_ptr, const uint64_t *xoff, int32_t *X)
{
    const uint64_t i = (uint64_t)blockIdx.x * CWAVES + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (i >= nclasses) return;
    const uint32_t c = cls_cluster[i];
    if (xoff[c] == ~uint64_t(0)) return;
    const uint32_t r = cls_grow[i];
    const uint32_t nb = rs.bptr[c + 1] - rs.bptr[c];
    const uint64_t base = rs.vioff[c] + (r - rs.voff[c]) * nb;
    const uint32_t m = rmin[r];
    int32_t *x = X + xoff[c] + (i - cls_ptr[c]) * 2ull * nb;
    for (uint32_t j = lane; j < nb; j += 64) {
        x[2 * j] = (int32_t)(rs.vi_len[base + j] - m);
        x[2 * j + 1] = rs.vi_cnt[base + j];
    }
}

uint32_t blocks_for(uint64_t items, uint64_t per_block) { return (uint32_t)((items + per_block - 1) / per_block); }

} // namespace

hipError_t classes_launch_hash(const RowSet &rs, uint64_t key_mask, uint64_t *key, uint32_t *val, uint32_t *rmin,
                               hipStream_t s)
{
    if (!rs.nrows) return hipSuccess;
    hipLaunchKernelGGL(k_class_hash, dim3(blocks_for(rs.nrows, CWAVES)), dim3(CWG), 0, s, rs, key_mask, key, val,
                       rmin);
    return hipGetLastError();
}

hipError_t classes_launch_borders(const RowSet &rs, uint8_t *is_border, hipStream_t s)
{
    hipError_t e = hipMemsetAsync(is_border, 0, rs.nrows ? rs.nrows : 1, s);
    if (e != hipSuccess || !rs.ncl) return e;
    hipLaunchKernelGGL(k_class_borders, dim3(rs.ncl), dim3(CWG), 0, s, rs, is_border);
    return hipGetLastError();
}

hipError_t classes_launch_runs(const RowSet &rs, const uint64_t *skey, const uint32_t *sval, const uint32_t *rmin,
                               const uint8_t *is_border, const double *intra, uint32_t *is_head, uint32_t *size,
                               uint32_t *pivot, uint32_t *collision, hipStream_t s)
{
    hipError_t e = hipMemsetAsync(is_head, 0, (rs.nrows ? rs.nrows : 1) * sizeof(uint32_t), s);
    if (e == hipSuccess) e = hipMemsetAsync(collision, 0, sizeof(uint32_t), s);
    if (e != hipSuccess || !rs.nrows) return e;
    hipLaunchKernelGGL(k_class_runs, dim3(blocks_for(rs.nrows, CWG)), dim3(CWG), 0, s, rs, skey, sval, rmin,
                       is_border, intra, is_head, size, pivot, collision);
    return hipGetLastError();
}

hipError_t classes_launch_compact(const RowSet &rs, const uint32_t *is_head, const uint32_t *pos,
                                  const uint32_t *size, const uint32_t *pivot, uint32_t *cls_row, uint32_t *cls_size,
                                  uint32_t *cls_cluster, uint32_t *cls_grow, uint32_t *cls_ptr, hipStream_t s)
{
    const uint64_t items = rs.nrows > (uint64_t)rs.ncl + 1 ? rs.nrows : (uint64_t)rs.ncl + 1;
    hipLaunchKernelGGL(k_class_compact, dim3(blocks_for(items, CWG)), dim3(CWG), 0, s, rs, is_head, pos, size,
                       pivot, cls_row, cls_size, cls_cluster, cls_grow, cls_ptr);
    return hipGetLastError();
}

hipError_t classes_launch_feature_max(const RowSet &rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                      const uint32_t *cls_grow, const uint32_t *rmin, unsigned long long *fmax,
                                      hipStream_t s)
{
    hipError_t e = hipMemsetAsync(fmax, 0, (rs.ncl ? rs.ncl : 1) * sizeof(unsigned long long), s);
    if (e != hipSuccess || !nclasses) return e;
    hipLaunchKernelGGL(k_class_fmax, dim3(blocks_for(nclasses, CWAVES)), dim3(CWG), 0, s, rs, nclasses,
                       cls_cluster, cls_grow, rmin, fmax);
    return hipGetLastError();
}

hipError_t classes_launch_gather(const RowSet &rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                 const uint32_t *cls_grow, const uint32_t *rmin, const uint32_t *cls_ptr,
                                 const uint64_t *xoff, int32_t *X, hipStream_t s)
{
    if (!nclasses) return hipSuccess;
    hipLaunchKernelGGL(k_class_gather, dim3(blocks_for(nclasses, CWAVES)), dim3(CWG), 0, s, rs, nclasses,
                       cls_cluster, cls_grow, rmin, cls_ptr, xoff, X);
    return hipGetLastError();
}

hipError_t classes_sort_bytes(uint64_t n, size_t *bytes)
{
    *bytes = 0;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                              (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n);
}

hipError_t classes_sort(void *tmp, size_t bytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                        uint32_t *vout, uint64_t n, hipStream_t s)
{
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)n, 0, 64, s);
}

hipError_t classes_scan_bytes(uint64_t n, size_t *bytes)
{
    *bytes = 0;
    return hipcub::DeviceScan::ExclusiveSum(nullptr, *bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n);
}

hipError_t classes_scan(void *tmp, size_t bytes, const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, s);
}

} // namespace fbcdev
