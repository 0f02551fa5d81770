/* Shared between classes.hip and the C-ABI layer (not a public header). */
#ifndef FBC_CLASSES_H
#define FBC_CLASSES_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbcdev {

/* The VertexInfo rows of every cluster, resident on the device, and the
 * cluster layout needed to walk them.  Rows are indexed by the concatenated
 * vertex order (cluster-major, a cluster's vertices ascending). */
struct RowSet {
    uint64_t nrows;             /* vertices over all clusters */
    uint32_t ncl;
    const uint32_t *vcl;        /* row -> cluster */
    const uint64_t *voff;       /* cluster -> first row */
    const uint32_t *bptr;       /* cluster -> border range (ncl + 1) */
    const uint32_t *blocal;     /* local row of each border */
    const uint64_t *vioff;      /* cluster -> first VertexInfo entry */
    const uint32_t *vi_len;     /* shortest-path length to each border (0: unreached) */
    const int32_t *vi_cnt;      /* (int) shortest-path count to each border */
    const uint32_t *gid;        /* row -> global vertex id */
};

/* Key of a row whose cluster has no border (such clusters get no pivots). */
constexpr uint64_t CLASS_KEY_NONE = ~uint64_t(0);

/* key[r] = hash of row r's normalized VertexInfo (lengths minus their
 * minimum, counts) mixed with its cluster, masked by key_mask (a test hook
 * that provokes collisions); rmin[r] = the minimum length; val[r] = r. */
hipError_t classes_launch_hash(const RowSet &rs, uint64_t key_mask, uint64_t *key, uint32_t *val,
                               uint32_t *rmin, hipStream_t s);

/* Marks border rows (is_border[row] = 1; the array is cleared first). */
hipError_t classes_launch_borders(const RowSet &rs, uint8_t *is_border, hipStream_t s);

/* Over rows sorted by key (stable, so equal keys list their rows in
 * ascending order): every run of equal keys is one class provided all its
 * rows equal the run's first row; a mismatch (a hash collision) sets
 * *collision.  The class is recorded at its first row h: is_head[h] = 1,
 * size[h] = run length, pivot[h] = the first non-border member with the
 * smallest intra bc (the first member when all are borders). */
hipError_t classes_launch_runs(const RowSet &rs, const uint64_t *skey, const uint32_t *sval,
                               const uint32_t *rmin, const uint8_t *is_border, const double *intra,
                               uint32_t *is_head, uint32_t *size, uint32_t *pivot, uint32_t *collision,
                               hipStream_t s);

/* Classes in row order: pos = exclusive prefix sum of is_head.  Writes the
 * class list (pivot's local row, size, cluster, pivot's row) and
 * cls_ptr[c] = first class of c, cls_ptr[ncl] = total. */
hipError_t classes_launch_compact(const RowSet &rs, const uint32_t *is_head, const uint32_t *pos,
                                  const uint32_t *size, const uint32_t *pivot, uint32_t *cls_row,
                                  uint32_t *cls_size, uint32_t *cls_cluster, uint32_t *cls_grow,
                                  uint32_t *cls_ptr, hipStream_t s);

/* Largest |feature| over the classes of each cluster (fmax[c], cleared first). */
hipError_t classes_launch_feature_max(const RowSet &rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                      const uint32_t *cls_grow, const uint32_t *rmin,
                                      unsigned long long *fmax, hipStream_t s);

/* k-means points: class i of cluster c (xoff[c] != ~0) becomes row
 * i - cls_ptr[c] of the block at X + xoff[c]: feature 2j = length to border
 * j minus the row minimum, 2j+1 = count. */
hipError_t classes_launch_gather(const RowSet &rs, uint64_t nclasses, const uint32_t *cls_cluster,
                                 const uint32_t *cls_grow, const uint32_t *rmin, const uint32_t *cls_ptr,
                                 const uint64_t *xoff, int32_t *X, hipStream_t s);

/* Scratch bytes of the key sort and of the prefix sum over n entries. */
hipError_t classes_sort_bytes(uint64_t n, size_t *bytes);
hipError_t classes_sort(void *tmp, size_t bytes, const uint64_t *kin, uint64_t *kout, const uint32_t *vin,
                        uint32_t *vout, uint64_t n, hipStream_t s);
hipError_t classes_scan_bytes(uint64_t n, size_t *bytes);
hipError_t classes_scan(void *tmp, size_t bytes, const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t s);

} // namespace fbcdev

#endif
