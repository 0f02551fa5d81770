/*
 * Host engine of the clustered betweenness path (non-template, V = int,
 * W = double as the reference's CLI instantiates it, main.cpp:25-31).
 *
 * The public C++ API (include/ under this directory, reference class names)
 * and the fbc command line are thin layers over these functions.  Device work
 * goes through the C ABI of include/fbc_gpu.h only.
 */
#ifndef FBC_HOST_H
#define FBC_HOST_H

#include "fbc_gpu.h"
#include "../../include/fbc_host.h"

#include <cstdint>
#include <functional>
#include <istream>
#include <mutex>
#include <string>
#include <vector>

namespace fbchost {

/* A loaded graph with DirectedWeightedGraph semantics
 * (DirectedWeightedGraph.h:67-96,135-182): n = 1 + largest source id, repeated
 * arcs summed and counted once, stars sorted by neighbour id. */
struct Graph {
    int n = 0;
    long long edges = 0;
    std::vector<uint64_t> rp;     /* forward CSR, rows sorted by target */
    std::vector<int> col;
    std::vector<double> w;
    std::vector<uint64_t> rrp;    /* backward CSR, rows sorted by source */
    std::vector<int> rcol;
    std::vector<double> rw;
    std::vector<double> in_wdeg, out_wdeg;
    double total_weight = 0;      /* sum of all arc weights in input order */
    uint64_t id = 0;              /* unique per built graph (device upload cache key) */
};

/* Edge-list text loader.  Throws std::invalid_argument with the reference's
 * message for a weight <= 0 (DirectedWeightedGraph.h:78-81), and for inputs the
 * reference cannot represent (arc into an id >= n, malformed line). */
Graph load_edge_list(std::istream &in);
Graph load_edge_list_file(const std::string &path);
/* load_edge_list_file through a binary CSR cache at cache_path (no
 * counterpart in the reference, which parses the text on every run): a cache
 * written for the same text file -- same size, modification time and a hash
 * of its first and last MiB -- is read instead of the text; otherwise the
 * text is parsed and the cache (re)written.  *hit says which happened. */
Graph load_edge_list_cached(const std::string &path, const std::string &cache_path, bool *hit = nullptr);
/* Same semantics from arc arrays (file order). */
Graph from_arcs(size_t m, const int *src, const int *dst, const double *w);
/* The same with an explicit vertex count n (ids must be < n): the vertex set
 * of an IGraph view, e.g. a SubGraph whose last vertex has no internal arc. */
Graph from_arcs_n(int n, size_t m, const int *src, const int *dst, const double *w);
void finish_graph(Graph &g);

/* LouvainGraphPartition::partitionGraph (louvain/LouvainGraphPartition.h:84-130)
 * with Partition::one_level / partition2graph (louvain/Partition.h), bit for bit:
 * same libstdc++ std::shuffle over std::mt19937, same FP64 operation order.
 * seeds: one instance per distinct seed, ascending (a std::set in main.cpp:133). */
std::vector<std::vector<int>> louvain(const Graph &g, const std::vector<uint64_t> &seeds,
                                      double precision);

/* VertexInfoPivotSelector::selectPivots (VertexInfoPivotSelector.h:29-127) on
 * the rows of one cluster: len is normalized in place (VertexInfo.h:268-278);
 * rows with equal (count, length) vectors form a class in first-appearance
 * order; each class' pivot is its non-border member of smallest bc (first wins),
 * or its first member when all are borders; weight = class size. */
void select_pivots(int nv, int nb, double *len, const int *cnt, const double *bc_row,
                   const char *is_border, std::vector<int> &piv_rows, std::vector<int> &weights);

/* PlusPlusKMeans::computeCentroids (kmeans/PlusPlusKMeans.h:70-170) on P rows
 * (FP64, the reference's order of operations; ties go to the first index). */
void kmeans(int k, int P, int nb, const double *L, const int *C, const int *weights,
            double min_variance, size_t max_iteration, std::vector<int> &centroids,
            std::vector<int> &cweights);

/* KMeansPivotSelector::selectPivots (KMeansPivotSelector.h:79-132): k-means over
 * the exact pivots, k = max((int)(P * kfrac), 1), repeated centroids dropped
 * keeping the first occurrence. */
void kmeans_pivots(const std::vector<int> &rows, const std::vector<int> &w, int nb,
                   const double *len, const int *cnt, double kfrac, std::vector<int> &out_rows,
                   std::vector<int> &out_w);

/* A device context plus the uploaded graph, shared by the C++ classes. */
class Device {
public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    fbc_ctx *ctx() const { return ctx_; }
    /* Upload g (integer weights required); cached by address. */
    fbc_graph *graph(const Graph &g);
    std::mutex &lock() { return mu_; }
    void check(int rc) const;   /* throws std::runtime_error with the library message */
    static Device &shared();    /* device 0, created on first use */

private:
    fbc_ctx *ctx_ = nullptr;
    uint64_t cached_ = 0;
    fbc_graph *graph_ = nullptr;
    std::mutex mu_;
};

struct PhaseTimes {
    double louvain = 0, cluster = 0, pivots = 0, pivot_phase = 0, total = 0, kmeans = 0;
    double kmeans_prep = 0, kmeans_gpu_ms = 0;
    size_t kmeans_gpu_sets = 0, kmeans_host_sets = 0;
    double cluster_te = 0, pivot_te = 0, cluster_kernel_ms = 0, pivot_kernel_ms = 0;
    size_t clusters = 0, bordered = 0, pivot_count = 0;
    /* the autotuner's choice for the cluster and pivot launches: delta,
     * half-team limit, and how chosen (fbc_last_counters' last three) */
    /* delta, half-team limit, how chosen, LDS-view kernel (cluster launch) */
    uint64_t cluster_tuning[4] = {0, 0, 0, 0}, pivot_tuning[4] = {0, 0, 0, 0};
};

struct ClusteredOptions {
    /* the caller's partition, one community index per vertex (the reference
     * takes any IGraphPartition, ClusteredBrandesBC.h:80-82); empty: Louvain
     * over 'seeds' and 'precision' */
    std::vector<int> partition;
    std::vector<uint64_t> seeds;
    double precision = 0.01;
    double kfrac = 0;            /* <= 0: plain VertexInfoPivotSelector */
    int prec = FBC_PREC_F64;     /* dependency arithmetic of the pivot phase */
    int form = FBC_FORM_BATCHED; /* FBC_FORM_*: batched, plain, or stop after selection */
    /* progress messages as the phases go (the reference's log lines at
     * level 0 = info; per-chunk lines at level 1 = debug); may be empty */
    std::function<void(int, const std::string &)> progress;
};

struct ClusteredRun {
    std::vector<double> bc, intra;
    std::vector<int> comm;                       /* cluster of each vertex */
    std::vector<uint32_t> pivots, weights, pivot_cluster;
    PhaseTimes t;
};

/* ClusteredBrandeBC::computeBC (brandes/ClusteredBrandesBC.h:63-169), batched:
 * Louvain on the host, every (cluster, source) in one persistent launch, pivot
 * selection on the host, every pivot in one persistent launch. */
ClusteredRun clustered_bc(Device &dev, const Graph &g, const ClusteredOptions &opt);

/* ExactBrandesBC::computeBC (brandes/ExactBrandesBC.h:46-95). */
std::vector<double> exact_bc(Device &dev, const Graph &g, int prec);

/* DijkstraSSBrandesBC::singleSourceBrandes (brandes/DijkstraSSBrandesBC.h:45-81). */
std::vector<double> single_source_dep(Device &dev, const Graph &g, int src, int prec);

} // namespace fbchost

#endif
