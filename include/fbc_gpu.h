/*
 * fbc_gpu.h -- C ABI of the MI355X (gfx950) betweenness-centrality kernels.
 *
 * The reference (15831944/fast-bc) is a header-only C++17 template library
 * with no ABI; its hot path is reached through four virtual interfaces.  Each
 * entry point below replaces the body of one of them and is called by the C++
 * host layer in fast-bc_amd/host (which keeps the reference's class names):
 *
 *   ISSBrandesBC::singleSourceBrandes     brandes/ISSBrandesBC.h:26-28
 *       -> fbc_single_source_dep / fbc_dep_batch
 *   IBrandesBC::computeBC (ExactBrandesBC) brandes/IBrandesBC.h:24, brandes/ExactBrandesBC.h:46-95
 *       -> fbc_exact_bc
 *   IClusterEvaluator::evaluateCluster     brandes/IClusterEvaluator.h:33-36
 *       -> fbc_clusters_create + fbc_cluster_eval (all clusters in one launch)
 *   ClusteredBrandeBC pivot loop           brandes/ClusteredBrandesBC.h:141-166
 *       -> fbc_pivot_phase (all pivots in one launch)
 *       -> fbc_pivot_phase_plain (one launch per pivot, the baseline form)
 *
 * Conventions
 *   - Return 0 on success or a negative FBC_E_* status; fbc_last_error()
 *     gives a message.  No C++ exceptions cross this boundary.
 *   - Host buffers are owned by the caller; device memory by the context.
 *   - A context is not thread-safe; use one per host thread.
 *   - Calls block until their results are in the caller's buffers.
 *   - Graphs are CSR with 0-based uint32 vertex ids and positive integer
 *     uint32 weights (the reference accepts any positive double weight; the
 *     GPU path keeps path lengths exact and therefore needs integers).
 *   - Distances and shortest-path counts are exact integers: path lengths in
 *     uint32 inside the kernels (checked), sigma in uint64 with an overflow
 *     check (FBC_E_SIGMA_OVERFLOW).
 */
#ifndef FBC_GPU_H
#define FBC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FBC_OK = 0,
    FBC_E_ARG = -1,            /* bad argument */
    FBC_E_HIP = -2,            /* HIP runtime error */
    FBC_E_OOM = -3,            /* device or host allocation failed */
    FBC_E_DEVICE = -4,         /* device is not a gfx950 (MI355X) */
    FBC_E_SIGMA_OVERFLOW = -5, /* a shortest-path count exceeded 2^64-1 */
    FBC_E_DIST_OVERFLOW = -6,  /* a path length exceeded 2^32-2 */
    FBC_E_UNSUPPORTED = -7
};

/* Arithmetic of the dependency (delta) sums. */
enum {
    FBC_PREC_F64 = 0,          /* FP64, the reference's operation order per vertex */
    FBC_PREC_F32C = 1          /* FP32 pairs with two-sum / two-product compensation */
};

typedef struct fbc_ctx fbc_ctx;
typedef struct fbc_graph fbc_graph;
typedef struct fbc_clusters fbc_clusters;

/* Context on HIP device 'device' (must be gfx950).  Replaces nothing in the
 * reference (its computers hold no device state). */
int fbc_create(int device, fbc_ctx **out);
void fbc_destroy(fbc_ctx *ctx);
const char *fbc_last_error(const fbc_ctx *ctx);

/* Tuning knobs (0 = automatic).  bucket_delta is the delta-stepping window;
 * slots bounds the sources in flight (one workgroup each). */
int fbc_set_tuning(fbc_ctx *ctx, uint32_t bucket_delta, uint32_t slots);

/* Named tuning option (0 = automatic / off): "delta", "slots" (as above),
 * "frontier_sort" (workgroup teams, i.e. views of more than 65,536
 * vertices, group each round's near queue by vertex-id range before
 * expanding it, in chunks of this many entries, up to 4096; 0, the default:
 * off), "level_sort" (the same for the shortest-path DAG's levels, which
 * the dependency phase then reads in that order too), "relabel"
 * (1, the default: the batched pivot phase and fbc_exact_bc run on a copy of
 * the graph renumbered into BFS-grown blocks of 16 vertices, for cache-line
 * locality; 0: on the graph as given), "scratch_mb" (device memory for the
 * sources in flight, in MiB; automatic: up to 8 whole-graph sources per CU
 * within half the device, less what is not free and a sixteenth kept back),
 * "class_key_bits" (bits of the row hashes fbc_cluster_classes
 * keeps; 0: all 64; a few bits force collisions, a test hook),
 * "half_team_max" (views of at most this many vertices run one source per
 * half wavefront, two sources sharing its instruction stream; default 2048,
 * 0: off), "autotune" (1, the default: a large batched run of the pivot
 * phase, exact BC or the cluster evaluation on a graph with no tuned
 * configuration yet tunes on its first tasks -- one launch of one task per
 * slot with mixed delta candidates (powers of two of the mean arc weight, 8
 * down to 1/16) locates the region (the pivot phase skips it when the
 * cluster evaluation has tuned on the same graph, and walks from that
 * choice), launches at one delta each walk to the best neighbour, and for
 * views small enough for half-wavefront teams one launch of whole
 * wavefronts decides the team form -- and keeps the choice
 * for the rest of the run and for later runs on the graph; no task runs
 * twice; an explicit "delta" turns it off; 0: off), "ballot_append" (1:
 * queue and level appends take one
 * atomic per wavefront -- a ballot of the appending lanes, the first adds
 * their count -- instead of one per lane; same results; 0, the default,
 * because it measured 0.5-1% slower on the 1M-vertex grid), "lds_view"
 * (cluster evaluation with each cluster's short rows staged in LDS, shared by
 * the 16 wavefront -- or 32 half-wavefront -- teams of a 1024-thread
 * workgroup, one workgroup per CU; clusters too large for the LDS read their
 * rows from HBM: 0 never, 1 always, 2, the default: when the clusters whose
 * rows fit carry at least half of the run's work, n_c x m_c summed, and every
 * team gets 32 tasks or more), "tune_log" (1: the autotuner prints a line per
 * trial launch on stderr -- its delta, tasks, time and edges per ms, and each
 * region candidate's edges per busy tick; 0, the default: silent),
 * "kmeans_budget_mb" (device memory for the k-means Gram matrices, in MiB;
 * sets share a Gram pass while theirs fit; 0, the default: 64 GiB or a
 * quarter of the device, within what is free, or the slot memory of the
 * Brandes launches, idle during the k-means, when that is larger; a test
 * hook).  No counterpart in the reference. */
int fbc_set_option(fbc_ctx *ctx, const char *name, int64_t value);

/* Upload a graph: row_ptr[n+1] (uint64, row_ptr[n] = m), col[m], w[m].
 * Replaces the device-side use of DirectedWeightedGraph's adjacency maps
 * (DirectedWeightedGraph.h:57-58, forwardStar :111-115). */
int fbc_graph_upload(fbc_ctx *ctx, uint32_t n, uint64_t m, const uint64_t *row_ptr,
                     const uint32_t *col, const uint32_t *w, fbc_graph **out);
void fbc_graph_free(fbc_graph *g);

/* Test hook: distances and sigma from nsrc sources, row-major [nsrc x n].
 * Unreached vertices get dist = UINT64_MAX and sigma = 0.
 * Follows the forward pass of DijkstraSSBrandesBC::_dijkstra_SSSP
 * (brandes/DijkstraSSBrandesBC.h:83-147). */
int fbc_sssp_sigma(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *src, uint32_t nsrc,
                   uint64_t *dist, uint64_t *sigma);

/* ISSBrandesBC::singleSourceBrandes (brandes/DijkstraSSBrandesBC.h:45-81):
 * dep[n] = dependency of every vertex on paths from src (0 for src and for
 * unreached vertices).  One source per call: the plain form.  With
 * FBC_PREC_F64 the source is spread over one cooperative launch (16
 * workgroups unless fbc_set_tuning's 'slots' says otherwise; grid-wide
 * barriers) and the sums follow the reference's successor
 * order, so dep equals the reference's bit for bit; FBC_PREC_F32C runs the
 * batched kernel with a single source. */
int fbc_single_source_dep(fbc_ctx *ctx, const fbc_graph *g, uint32_t src, int prec,
                          double *dep);

/* The same for nsrc sources in one launch; dep is [nsrc x n] row-major. */
int fbc_dep_batch(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *src, uint32_t nsrc,
                  int prec, double *dep);

/* fbc_dep_batch for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  d_src (nsrc source ids) and d_dep
 * (nsrc x n doubles, row-major, every entry written) are device pointers the
 * caller owns; the graph is one fbc_graph_upload made earlier.  The call
 * allocates nothing: its per-source state lives in a 64 MiB array that each
 * GPU running the library holds until the process ends, so a graph whose
 * one source (a wavefront team's four, or a half-wavefront team's eight, for
 * graphs of at most 65,536 or 2,048 vertices) needs more than that --
 * about 40 bytes per vertex, beyond roughly 1.6 million vertices -- is
 * refused with FBC_E_UNSUPPORTED.  Values equal fbc_dep_batch's bit for bit.
 * Calls that overlap on one GPU (several contexts or threads) share the
 * array and race. */
int fbc_dep_batch_noalloc(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *d_src, uint32_t nsrc,
                          int prec, double *d_dep);

/* ExactBrandesBC::computeBC (brandes/ExactBrandesBC.h:46-95): bc[n] = sum over
 * all sources s of dep_s(v), v != s.  All sources in one persistent launch. */
int fbc_exact_bc(fbc_ctx *ctx, const fbc_graph *g, int prec, double *bc);

/* Clusters: ncl vertex sets given as cluster_ptr[ncl+1] / cluster_verts (each
 * set ascending, sets disjoint).  Builds SubGraph views (SubGraph.h:59-148):
 * border = vertex with an arc leaving or entering its set; the intra-cluster
 * forward stars.  border_ptr[ncl+1] receives the border counts as offsets. */
int fbc_clusters_create(fbc_ctx *ctx, const fbc_graph *g, uint32_t ncl,
                        const uint32_t *cluster_ptr, const uint32_t *cluster_verts,
                        fbc_clusters **out, uint32_t *border_ptr);
/* Ascending border ids of every cluster, concatenated per border_ptr. */
int fbc_clusters_borders(const fbc_clusters *cl, uint32_t *borders);
/* Internal arc count of cluster c (SubGraph::edges, SubGraph.h:106). */
int64_t fbc_clusters_edges(const fbc_clusters *cl, uint32_t c);
void fbc_clusters_free(fbc_clusters *cl);

/* DijkstraClusterEvaluator::evaluateCluster (DijkstraClusterEvaluator.h:50-100)
 * for every cluster with at least one border, all sources in one launch.
 *   intra_bc[n]   += intra-cluster betweenness (FP64, summed over the sources in
 *                    vertex order, as the reference's one-thread team does);
 *   vi_len / vi_cnt  rows of VertexInfo: for cluster c, source row r (r-th
 *                    vertex of the set) and border j the entry sits at
 *                    vi_off[c] + r * B_c + j, vi_off[c] = sum over c' < c of
 *                    n_c' * B_c'.  len = path length to the border (0 when
 *                    unreached), cnt = (int)sigma as the reference converts
 *                    its FP64 count (INT_MIN beyond the int range,
 *                    DijkstraClusterEvaluator.h:166-177).
 * Borderless clusters are skipped (FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED). */
int fbc_cluster_eval(fbc_ctx *ctx, const fbc_clusters *cl, double *intra_bc,
                     uint32_t *vi_len, int32_t *vi_cnt);

/* fbc_cluster_eval with the VertexInfo rows kept in device memory with the
 * clusters (for fbc_cluster_classes) instead of copied out; fbc_cluster_rows
 * copies them out later.  Same rows, same intra_bc. */
int fbc_cluster_eval_resident(fbc_ctx *ctx, fbc_clusters *cl, double *intra_bc);
int fbc_cluster_rows(fbc_ctx *ctx, const fbc_clusters *cl, uint32_t *vi_len, int32_t *vi_cnt);

/* VertexInfoPivotSelector::selectPivots (VertexInfoPivotSelector.h:28-127)
 * for every bordered cluster, on the resident rows: the exact classes of
 * vertices with equal normalized rows (lengths minus the row's minimum,
 * counts), in order of first appearance; each class's pivot is its first
 * non-border member with the smallest intra_bc (its first member when all
 * are borders, as FASTBC_BRANDES_ENABLE_PIVOT_BORDER, main.cpp:5, allows) and
 * its weight the class size.  Outputs: cls_ptr[ncl + 1] (classes of cluster
 * c are [cls_ptr[c], cls_ptr[c+1])), piv_row (row of the pivot within its
 * cluster's vertex list) and weight, up to one entry per cluster vertex.
 * Classes are found through a 64-bit hash of each row, verified entry by
 * entry; a hash collision returns FBC_E_UNSUPPORTED, and the caller then
 * selects on the host from fbc_cluster_rows.  The classes stay on the device
 * for fbc_cluster_kmeans. */
int fbc_cluster_classes(fbc_ctx *ctx, fbc_clusters *cl, const double *intra_bc, uint32_t *cls_ptr,
                        uint32_t *piv_row, uint32_t *weight);

/* KMeansPivotSelector's k-means (KMeansPivotSelector.h:79-132 with
 * PlusPlusKMeans, as fbc_kmeans_batch) over each cluster's classes from the
 * last fbc_cluster_classes: k_c = max((int)(P_c * kfrac), 1) for the P_c
 * classes of cluster c, points = the classes' normalized rows, weights = the
 * class sizes.  Outputs kptr[ncl + 1] and, in [kptr[c], kptr[c+1]), the
 * centroids (class indices within the cluster) and their weights.  A cluster
 * with classes but an empty range needs the host k-means (its features are
 * too large for exact FP64 distances or k_c > FBC_KMEANS_MAX_K); its points
 * come from fbc_cluster_class_features: P_c rows of 2 * B_c int32 features
 * (length minus the row minimum, count, per border). */
int fbc_cluster_kmeans(fbc_ctx *ctx, const fbc_clusters *cl, double kfrac, double min_variance,
                       uint32_t max_iteration, uint32_t *kptr, uint32_t *centroids, int32_t *cweights);
int fbc_cluster_class_features(fbc_ctx *ctx, const fbc_clusters *cl, uint32_t c, int32_t *features);

/* The pivot loop of ClusteredBrandeBC::computeBC (ClusteredBrandesBC.h:141-166):
 *   bc[v] += w_p * dep_p(v)                for every v,
 *   bc[v] -= w_p * intra_bc[v]             for v in the pivot's cluster,
 * over npiv pivots (pivot id, weight, cluster index into cl) in one launch.
 * The dependency sums use 'prec'; the per-vertex reduction is FP64. */
int fbc_pivot_phase(fbc_ctx *ctx, const fbc_clusters *cl, const uint32_t *pivot,
                    const uint32_t *weight, const uint32_t *pivot_cluster, uint32_t npiv,
                    const double *intra_bc, int prec, double *bc);

/* The same pivot loop in the plain form, as the reference runs it: one
 * singleSourceBrandes per pivot (ClusteredBrandesBC.h:143-148 calling
 * DijkstraSSBrandesBC.h:45-81), here one cooperative launch per
 * pivot with FP64 sums in the reference's successor order.  Same result
 * contract as fbc_pivot_phase with FBC_PREC_F64; the baseline the batched
 * form is measured against.  fbc_set_tuning's 'slots' sets the workgroups of
 * the grid (0: 16, a few more than a thin frontier keeps busy; the grid
 * barrier of every round and level costs more with a larger grid).  If the
 * runtime refuses the first cooperative launch of a call for want of device
 * resources (another process on the device has made cooperative launches:
 * HIP status 401, hipErrorIllegalState), this and fbc_single_source_dep
 * finish on the batched kernel, with the same values, and fbc_last_counters
 * reports it; any other launch failure is returned as FBC_E_HIP. */
int fbc_pivot_phase_plain(fbc_ctx *ctx, const fbc_clusters *cl, const uint32_t *pivot,
                          const uint32_t *weight, const uint32_t *pivot_cluster, uint32_t npiv,
                          const double *intra_bc, double *bc);

/* PlusPlusKMeans::computeCentroids (kmeans/PlusPlusKMeans.h:70-229) for nsets
 * independent point sets in one batch (one set per cluster's exact pivots, as
 * KMeansPivotSelector calls it, brandes/KMeansPivotSelector.h:99-101).
 * Set s has npoints[s] rows of nfeat[s] features (row-major; rows of all sets
 * concatenated in 'features'), npoints[s] weights and k[s] centroids.
 * Features are VertexInfo rows (length then count per border), integers
 * with 2 * nfeat * max|feature|^2 < 2^53, which makes every point-to-point
 * distance exact, and k[s] <= FBC_KMEANS_MAX_K (FBC_E_UNSUPPORTED
 * otherwise).  Outputs, k[s] entries per set,
 * concatenated: centroid point indices and their weights.  Each set's Gram
 * matrix (npoints^2 doubles) is held on the device while the set runs; sets
 * are processed in groups that fit a fraction of device memory. */
#define FBC_KMEANS_MAX_K 12288
int fbc_kmeans_batch(fbc_ctx *ctx, uint32_t nsets, const uint32_t *npoints, const uint32_t *nfeat,
                     const uint32_t *k, const int32_t *features, const int32_t *weights,
                     double min_variance, uint32_t max_iteration, uint32_t *centroids,
                     int32_t *cweights);

/* Kernel time of the last launch of the entry points above, in milliseconds,
 * and the number of forward arcs it examined (traversed edges). */
int fbc_last_stats(const fbc_ctx *ctx, double *kernel_ms, double *traversed_edges);

/* Work counters of the last launch, summed over its sources: distance-phase
 * rounds, windows, expanded vertices, relaxations, improvements, far
 * insertions, DAG levels, then the nanoseconds spent in the distance, sigma and
 * dependency phases (per workgroup, so the sum exceeds the kernel time by the
 * number of workgroups), the sources whose sigma counts outgrew the packed
 * 48-bit form and were counted again in 64 bits; then, for the plain form
 * (fbc_single_source_dep with FBC_PREC_F64, fbc_pivot_phase_plain), the
 * sources run by cooperative launches, the sources finished on the
 * batched kernel after a refused first launch, and the HIP status of that
 * refusal (0: none); then the delta and half-team limit of the last run's
 * (main) launch and how they were chosen (0: fixed or set by option, 1:
 * measured by the autotuner in this run, 2: the graph's tuned choice from an
 * earlier run); then 1 when the last run met a shortest-path count of 2^53
 * or more (exact here, where the reference's FP64 count is rounded; counts
 * of 2^64 or more return FBC_E_SIGMA_OVERFLOW); then 1 when the last
 * batched cluster evaluation ran on the LDS-view kernel.  Fills at most 'count'
 * entries, returns how many it wrote.  Instrumentation only; no counterpart
 * in the reference. */
int fbc_last_counters(const fbc_ctx *ctx, uint64_t *counters, int count);

#ifdef __cplusplus
}
#endif

#endif
