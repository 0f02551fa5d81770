/*
 * fbc oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain CPU restatement of the reference's clustered-betweenness path
 * (15831944/fast-bc, libfastbc/include), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (fast-bc_amd/) includes, links or calls this code.
 *
 * Every function follows the reference file:line cited next to it and keeps
 * the reference's floating-point operation order (build with
 * -ffp-contract=off; the reference host build has no FMA, x86-64 baseline).
 *
 * Where the reference's result depends on the OpenMP runtime (array-section
 * reductions combined in thread-arrival order: ExactBrandesBC.h:59,
 * ClusteredBrandesBC.h:143) the oracle uses libgomp's static schedule for
 * 'threads' threads and combines the per-thread partial sums in thread order.
 *
 * Pinning (which check fixes each function) is listed in oracle/README.md.
 */
#ifndef FBC_ORACLE_H
#define FBC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph orc_graph;

/* DirectedWeightedGraph(std::istream&), DirectedWeightedGraph.h:67-96.
 * Returns NULL on error and writes a message into err. */
orc_graph *orc_graph_from_file(const char *path, char *err, int errlen);
/* Same merge semantics (DirectedWeightedGraph.h:135-170) from arc arrays. */
orc_graph *orc_graph_from_arcs(uint64_t m, const uint32_t *src, const uint32_t *dst,
                               const double *w, char *err, int errlen);
void orc_graph_free(orc_graph *g);
int orc_graph_n(const orc_graph *g);
long long orc_graph_edges(const orc_graph *g);
/* Forward CSR sorted by target id: rowptr[n+1], col[edges], w[edges]. */
int orc_graph_csr(const orc_graph *g, uint64_t *rowptr, uint32_t *col, double *w);
/* Backward star of one vertex (DirectedWeightedGraph.h:117-121); returns size,
 * fills at most cap entries. */
int orc_graph_bstar(const orc_graph *g, int v, int cap, int *src, double *w);

/* Forward Dijkstra of DijkstraSSBrandesBC::_dijkstra_SSSP
 * (DijkstraSSBrandesBC.h:83-147): dist (+inf when unreached), sigma (FP64 as
 * in the reference) and the settle order (-1 padded).  Returns #settled. */
int orc_sssp(const orc_graph *g, int src, double *dist, double *sigma, int *order);

/* DijkstraSSBrandesBC::singleSourceBrandes (DijkstraSSBrandesBC.h:45-81). */
int orc_single_source_dep(const orc_graph *g, int src, double *dep);

/* ExactBrandesBC::computeBC (ExactBrandesBC.h:46-95). */
int orc_exact_bc(const orc_graph *g, int threads, double *bc);
/* Exact BC over the sources [s0, s1) only, in source order (sampled timing). */
int orc_exact_bc_range(const orc_graph *g, int s0, int s1, int threads, double *bc);

/* SubGraph(vertices, graph) (SubGraph.h:59-148): ascending borders and the
 * internal arc count.  borders must hold nv entries. */
int orc_subgraph(const orc_graph *g, int nv, const int *verts, int *nborders,
                 int *borders, long long *edges);

/* DijkstraClusterEvaluator::evaluateCluster (DijkstraClusterEvaluator.h:50-100):
 * adds intra-cluster BC into bc[n]; vi_len/vi_cnt are nv x nborders row-major
 * (row = cluster vertex in the given order, column = ascending border). */
int orc_cluster_eval(const orc_graph *g, int nv, const int *verts, double *bc,
                     int *nborders, int *borders, double *vi_len, int *vi_cnt);

/* VertexInfoPivotSelector::selectPivots (VertexInfoPivotSelector.h:29-127) on
 * explicit rows.  Normalizes vi_len in place.  Returns #pivots. */
int orc_select_pivots(int nv, const int *verts, int nb, double *vi_len,
                      const int *vi_cnt, const double *bc_of_vert, int nborders,
                      const int *borders, int *pivots, int *weights);

/* PlusPlusKMeans::computeCentroids (kmeans/PlusPlusKMeans.h:70-170) on P points
 * (indexed 0..P-1, rows of vi_len / vi_cnt of width nb).  Returns k and fills
 * centroids (point indices) and weights. */
int orc_kmeans(int k, int P, int nb, const double *vi_len, const int *vi_cnt,
               const int *weights, double min_variance, int max_iteration,
               int *centroids, int *out_weights);

/* VertexInfo::squaredDistance (brandes/VertexInfo.h:293-307) of two rows:
 * the FP64 instance the path uses, and an FP32 instance (lengths of the
 * second row given as int, as VertexInfo<short,int> in the reference's test). */
double orc_vi_sqdist(int nb, const double *len_a, const int *cnt_a, const double *len_b,
                     const int *cnt_b);
float orc_vi_sqdist_f32(int nb, const float *len_a, const int *cnt_a, const int *len_b,
                        const int *cnt_b);

/* LouvainGraphPartition::partitionGraph (louvain/LouvainGraphPartition.h:84-130).
 * seeds are deduplicated and sorted as a std::set; comm_of[n] receives each
 * vertex's community index.  Returns the community count. */
int orc_louvain(const orc_graph *g, int nseeds, const uint64_t *seeds,
                double precision, int *comm_of);

/* ClusteredBrandeBC::computeBC (brandes/ClusteredBrandesBC.h:63-169) with the
 * strategies main.cpp composes (main.cpp:223-260).  kfrac <= 0 selects the
 * plain VertexInfoPivotSelector, otherwise KMeansPivotSelector(kfrac).
 * Optional outputs (may be NULL): comm_of[n], intra_bc[n], pivots / weights /
 * pivot_cluster (capacity n), te (traversed-edge count of the run).
 * Returns the pivot count or a negative error. */
int orc_clustered_bc(const orc_graph *g, int nseeds, const uint64_t *seeds,
                     double precision, double kfrac, int threads, double *bc,
                     int *comm_of, double *intra_bc, int *pivots, int *weights,
                     int *pivot_cluster, double *te);

/* ClusteredBrandeBC::computeBC with a caller's partition in place of Louvain
 * (the reference takes any IGraphPartition, ClusteredBrandesBC.h:80-82):
 * community c = the vertices with comm_in == c, ascending, communities in
 * ascending c (empty ones dropped); comm_in < 0: in no community.  Outputs as
 * orc_clustered_bc (pivot_cluster indexes the non-empty communities).
 * Returns the pivot count or a negative error. */
int orc_clustered_bc_partition(const orc_graph *g, const int *comm_in, double kfrac, int threads,
                               double *bc, double *intra_bc, int *pivots, int *weights, int *pivot_cluster);

/* KMeansPivotSelector::selectPivots after the exact selector
 * (KMeansPivotSelector.h:91-129): P exact pivots (rows 0..P-1 of vi_len /
 * vi_cnt, class sizes in weights), k = max((int)(P * kfrac), 1),
 * PlusPlusKMeans, then repeated centroids dropped keeping the first
 * occurrence and its weight.  Fills pivots (row indices) and out_weights
 * (capacity P); returns their count. */
int orc_kmeans_pivots(int P, int nb, const double *vi_len, const int *vi_cnt, const int *weights, double kfrac,
                      int *pivots, int *out_weights);

/* Phases 1-3 of the clustered run only (partition, intra-cluster BC, pivots);
 * all outputs have capacity n.  Returns the pivot count. */
int orc_clustered_pivots(const orc_graph *g, int nseeds, const uint64_t *seeds, double precision,
                         double kfrac, int threads, int *comm_of, double *intra_bc, int *pivots,
                         int *weights, int *pivot_cluster);

/* Sampled pivot phase: only the first max_pivots pivots (in cluster order) are
 * run.  Used for bounded CPU timing. */
int orc_clustered_bc_sampled(const orc_graph *g, int nseeds, const uint64_t *seeds,
                             double precision, double kfrac, int threads,
                             int max_pivots, double *bc, double *phase_seconds);

#ifdef __cplusplus
}
#endif

#endif
