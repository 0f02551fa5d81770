/*
 * fbc_host.h -- C ABI of the host side of the clustered betweenness path.
 *
 * Louvain partitioning, pivot selection and k-means stay on the host (the
 * partition must be bit-identical to the reference's); the end-to-end entry
 * point drives the GPU through fbc_gpu.h.  Same conventions as fbc_gpu.h:
 * 0 or a negative FBC_E_* status, caller-owned buffers, blocking calls.
 * Graphs are arc lists in file order with the loader's semantics
 * (DirectedWeightedGraph.h:67-96,135-170): repeated arcs are summed, the
 * vertex count is 1 + the largest source id.
 */
#ifndef FBC_HOST_C_H
#define FBC_HOST_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LouvainGraphPartition::partitionGraph (louvain/LouvainGraphPartition.h:84-130):
 * comm_of[n] receives each vertex's community; returns the community count.
 * Seeds are deduplicated and used in ascending order, one instance each. */
int fbc_louvain(uint64_t m, const int32_t *src, const int32_t *dst, const double *w,
                const uint64_t *seeds, int nseeds, double precision, int32_t *comm_of);

/* VertexInfoPivotSelector::selectPivots (brandes/VertexInfoPivotSelector.h:29-127)
 * on the nv x nb VertexInfo rows of one cluster (row r = r-th vertex of the
 * ascending vertex list).  vi_len is normalized in place.  Fills piv_rows /
 * weights (capacity nv); returns the pivot count. */
int fbc_select_pivots(int nv, int nb, double *vi_len, const int32_t *vi_cnt,
                      const double *bc_row, const uint8_t *is_border, int32_t *piv_rows,
                      int32_t *weights);

/* PlusPlusKMeans::computeCentroids (kmeans/PlusPlusKMeans.h:70-170) over P
 * rows; fills centroids / cweights (capacity k), returns k. */
int fbc_kmeans(int k, int P, int nb, const double *vi_len, const int32_t *vi_cnt,
               const int32_t *weights, double min_variance, int max_iteration,
               int32_t *centroids, int32_t *cweights);

/* How fbc_clustered_bc runs its pivot phase. */
enum {
    FBC_FORM_BATCHED = 0,      /* every pivot in one persistent launch (fbc_pivot_phase) */
    FBC_FORM_PLAIN = 1,        /* one cooperative launch per pivot, FP64 (fbc_pivot_phase_plain) */
    FBC_FORM_SELECT_ONLY = 2   /* stop after pivot selection: bc = intra-cluster BC */
};

/* ClusteredBrandeBC::computeBC (brandes/ClusteredBrandesBC.h:63-169) as main.cpp
 * composes it (main.cpp:223-260): kfrac <= 0 selects the exact pivot selector.
 * bc[n] (pre-clamp), optional comm_of[n], intra_bc[n], pivots / weights /
 * pivot_cluster (capacity n), phase_seconds[8] = {louvain, clusters, pivot
 * selection, pivot phase, total, cluster traversed edges, pivot traversed
 * edges, pivot-phase kernel ms}.  'form' is one of FBC_FORM_*; prec applies
 * to the batched form.  Returns the pivot count. */
int fbc_clustered_bc(int device, uint64_t m, const int32_t *src, const int32_t *dst,
                     const double *w, const uint64_t *seeds, int nseeds, double precision,
                     double kfrac, int prec, int form, double *bc, int32_t *comm_of,
                     double *intra_bc, uint32_t *pivots, uint32_t *weights,
                     uint32_t *pivot_cluster, double *phase_seconds);

/* fbc_clustered_bc with the caller's partition in place of Louvain, as
 * ClusteredBrandeBC runs with any IGraphPartition strategy
 * (brandes/ClusteredBrandesBC.h:80-82 calling IGraphPartition::partitionGraph,
 * IGraphPartition.h:19): comm_in[n] gives each vertex's community (negative:
 * in none, as a strategy need not cover every vertex); community c holds the
 * vertices with index c in ascending order, the communities follow
 * ascending c (empty ones dropped), and pivot_cluster counts the non-empty
 * ones.  Other arguments as fbc_clustered_bc. */
int fbc_clustered_bc_partition(int device, uint64_t m, const int32_t *src, const int32_t *dst,
                               const double *w, const int32_t *comm_in, double kfrac, int prec,
                               int form, double *bc, double *intra_bc, uint32_t *pivots,
                               uint32_t *weights, uint32_t *pivot_cluster, double *phase_seconds);

/* ExactBrandesBC::computeBC (brandes/ExactBrandesBC.h:46-95) from arcs. */
int fbc_exact_bc_arcs(int device, uint64_t m, const int32_t *src, const int32_t *dst,
                      const double *w, int prec, double *bc);

/* The autotuner's choice in the last fbc_clustered_bc / _partition call:
 * out[6] = {delta, half-team limit, how chosen} for the cluster launch, then
 * the same for the pivot launch ("how": 0 fixed or set by option, 1
 * measured in that call, 2 the graph's choice from an earlier call; see
 * fbc_set_option "autotune").  No counterpart in the reference. */
int fbc_host_last_tuning(uint64_t *out);

/* The same with the cluster launch's kernel form: out[0..3] = {delta,
 * half-team limit, how chosen, 1 when the cluster evaluation ran on the
 * LDS-view kernel} for the cluster launch, out[4..7] the same for the pivot
 * launch (its LDS entry is always 0).  Writes at most 'count' entries and
 * returns how many.  No counterpart in the reference. */
int fbc_host_last_tuning_ex(uint64_t *out, int count);

/* Message of the calling thread's last failed call above (empty when none
 * failed yet).  No counterpart in the reference, whose errors are
 * exceptions. */
const char *fbc_host_last_error(void);

/* fbc_clustered_bc and fbc_exact_bc_arcs keep one device context per device
 * (configured from FBC_OPTIONS, re-created when it changes) and the graph of
 * the last arcs they were given, so that a repeated call on the same arcs
 * finds the graph and its device copies resident.  This releases them (and
 * the device memory they hold).  No counterpart in the reference. */
void fbc_host_release(void);

#ifdef __cplusplus
}
#endif

#endif
