/* Shared between the HIP kernels and the C-ABI layer (not a public header). */
#ifndef FBC_INTERNAL_H
#define FBC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbcdev {

constexpr uint32_t DINF = 0xFFFFFFFFu;     /* unreached */
/* Short rows, copied beside the CSR: a vertex v's arcs in one 16-byte row of
 * ELL_W words, each (weight << 16) | (uint16)(target - v), padded with
 * ELL_NONE; a row that does not fit (more than ELL_W arcs, a target more than
 * 32767 ids away, a weight of 0xFFFF or more) is marked ELL_WIDE in its
 * first word and read from the CSR.  The distance phase reads a vertex's arcs
 * in one row instead of a row-start line and an arc line, and the rows of a
 * whole graph take a quarter of its CSR arcs, so more of them stay in L2. */
constexpr uint32_t ELL_W = 4;
constexpr uint32_t ELL_NONE = 0xFFFFFFFFu;
constexpr uint32_t ELL_WIDE = 0xFFFFFFFEu;
constexpr uint32_t ELL_WMAX = 0xFFFEu;   /* largest weight a short row holds */
constexpr int WG = 256;                    /* threads per source (4 wavefronts) */

enum Mode : int {
    MODE_SSSP = 0,      /* write dist / sigma rows (test hook) */
    MODE_DEP = 1,       /* write dependency rows */
    MODE_EXACT = 2,     /* bc[v] += dep(v) */
    MODE_PIVOT = 3,     /* bc[v] += w * dep(v) */
    MODE_CLUSTER = 4    /* dependency matrix rows + VertexInfo rows */
};

enum ErrBits : uint32_t {
    ERR_DIST = 1u,          /* path length beyond 2^32-2 */
    ERR_SIGMA = 2u,         /* sigma beyond 2^64-1 */
    ERR_SIGMA_INEXACT = 4u, /* sigma >= 2^53: the reference's FP64 sigma is inexact */
    ERR_SOURCE = 8u         /* a task's source id is not a vertex of its view */
};

/* delta candidates of one autotuner trial launch at most */
constexpr uint32_t TRIAL_MAX = 8;

/* Everything a launch needs, passed by value. */
struct BrandesParams {
    /* graph view used when cl_n == nullptr */
    const uint2 *arcs;        /* {dst, w}, rows sorted by dst */
    const uint32_t *rp;       /* row starts, n+1 */
    const uint2 *rarcs;       /* reverse arcs {src, w} */
    const uint32_t *rrp;
    const uint4 *ell;         /* short rows (nullptr: the CSR only); cluster views at cl_voff */
    uint32_t n;

    /* cluster views: concatenated local CSRs (row starts are absolute) */
    const uint32_t *cl_n;     /* vertices per cluster */
    const uint64_t *cl_rpoff; /* offset of the cluster's row-start block (n_c + 1 entries) */
    const uint64_t *cl_voff;  /* offset of the cluster's vertex block */
    const uint32_t *cl_gid;   /* local -> global id, by cl_voff */
    const uint32_t *cl_bptr;  /* border offsets (ncl + 1) */
    const uint32_t *cl_blocal;/* local ids of the borders, ascending */
    const uint64_t *cl_vioff; /* VertexInfo block offset per cluster */
    const uint64_t *cl_doff;  /* dependency-matrix offset per cluster (this chunk) */

    /* tasks: source (local id in cluster mode), aux = cluster index or weight;
     * the batched kernels' queue index q runs task task_base + q (the
     * autotuner's launches each take the next run of tasks) */
    uint32_t ntasks;
    uint32_t task_base;
    /* the autotuner's trial launch (trial_sets > 0): queue index q runs with
     * delta trial_delta[q % trial_sets] and adds its traversed edges and busy
     * wall-clock ticks to trial_stats[2 (q % trial_sets)] and [... + 1] */
    uint32_t trial_sets;
    uint32_t trial_delta[TRIAL_MAX];
    unsigned long long *trial_stats;
    const uint32_t *task_src;
    const uint32_t *task_aux;
    uint32_t *queue_head;

    /* per-workgroup scratch */
    uint8_t *scratch;
    uint64_t slot_bytes;
    uint32_t nmax;
    uint32_t delta;           /* near-far window */

    /* outputs */
    double *bc;
    double *out_dep;
    uint64_t *out_dist;
    uint64_t *out_sigma;
    double *dmat;
    uint32_t *vi_len;
    int32_t *vi_cnt;
    uint32_t *err;
    unsigned long long *te;   /* traversed-edge counter */
    unsigned long long *stats;/* STAT_COUNT work counters, summed over tasks */
    uint32_t *grid_ctl;       /* plain kernel: traversal counters shared by the grid (16 words) */
    uint32_t sort_cap;        /* workgroup teams: group near queues by id, in chunks of this many entries (0: off) */
    uint32_t level_sort;      /* workgroup teams: the same for DAG levels (0: off) */
    uint32_t nslots;          /* slots the scratch holds (the wavefront kernels check their slot against it) */
    uint32_t half_max;        /* views up to this many vertices: half-wavefront teams */
    uint32_t ballot_append;   /* 1: queue and level appends take one atomic per wavefront */
    /* cluster evaluation on the LDS-view kernel (lds_rows_max > 0): views of
     * up to lds_rows_max vertices can have their short rows staged in LDS */
    uint32_t lds_rows_max;
};

/* Workgroup teams group a round's near queue and each DAG level by vertex
 * id before expanding them, with a one-pass counting sort in LDS: ids go to
 * SORT_BUCKETS buckets by v >> shift (the smallest shift that maps the view
 * into them), at most SORT_CHUNK_MAX entries at a time (longer lists are
 * grouped chunk by chunk). */
constexpr uint32_t SORT_CHUNK_MAX = 4096;
constexpr uint32_t SORT_BUCKETS = 2048;

/* Work counters of the persistent kernel (fbc_last_stats_ex). */
enum Stat : int {
    STAT_ROUNDS = 0,      /* near rounds (workgroup barriers) of the distance phase */
    STAT_WINDOWS = 1,     /* near-far windows */
    STAT_VISITS = 2,      /* near-queue vertices expanded */
    STAT_RELAX = 3,       /* arcs relaxed in the distance phase */
    STAT_IMPROVE = 4,     /* relaxations that lowered a distance */
    STAT_FARPUSH = 5,     /* far-list insertions */
    STAT_LEVELS = 6,      /* levels of the shortest-path DAG */
    STAT_T_DIST = 7,      /* time in the distance phase, summed over sources */
    STAT_T_SIGMA = 8,     /* time in the sigma phase */
    STAT_T_DEP = 9,       /* time in the dependency phase */
    STAT_SIGMA_WIDE = 10, /* sources whose sigma sweep ran again in the wide form */
    STAT_COUNT = 11
};

/* Views of at most this many vertices run one source per wavefront (their
 * frontiers are a few dozen to a few hundred vertices wide); larger ones one
 * per workgroup.  On the 4M-vertex grid's clusters (13K-24K vertices) the
 * wavefront form runs the cluster launch 32% faster than workgroups, on a
 * 32K-vertex R-MAT graph exact BC 7% faster. */
constexpr uint32_t WAVE_TEAM_MAX_N = 65536;
constexpr int WAVES = WG / 64;             /* slots of a wavefront-per-source workgroup */
inline bool wave_team(uint32_t nmax) { return nmax <= WAVE_TEAM_MAX_N; }
/* Views of at most half_max vertices (an option; 0 = never) run one source
 * per half wavefront: two sources share a wavefront's instruction stream.
 * Measured: clusters of a few hundred vertices (65K-vertex grid) run their
 * launch 16-30% faster that way; at ~5K vertices the wavefront form is 1%
 * faster, and a whole 64K-vertex graph runs 2x slower. */
constexpr uint32_t HALF_TEAM_MAX_N = 2048;
constexpr int HALVES = WG / 32;            /* slots of a half-wavefront-per-source workgroup */
enum TeamKind { TEAM_WG = 0, TEAM_WAVE = 1, TEAM_HALF = 2 };
/* The LDS-view cluster kernel: one workgroup of LWG threads per CU (its LDS
 * holds one cluster's short rows), as many wavefront or half-wavefront teams
 * as the other kernels keep resident per CU. */
constexpr int LWG = 1024;
constexpr uint32_t LDS_TEAMS_WAVE = LWG / 64, LDS_TEAMS_HALF = LWG / 32;
inline int team_kind(uint32_t nmax, uint32_t half_max)
{
    return nmax <= half_max ? TEAM_HALF : wave_team(nmax) ? TEAM_WAVE : TEAM_WG;
}

/* dst[v] += src[perm[v]] for v < n (sums of a relabelled run back to the
 * original numbering). */
hipError_t launch_gather_add(uint32_t n, const uint32_t *perm, const double *src, double *dst, hipStream_t stream);

/* p[0, nwords) = 0 by a kernel (safe inside captured graphs, unlike a
 * memset node). */
hipError_t launch_clear_words(uint32_t *p, uint32_t nwords, hipStream_t stream);

/* Bytes of one slot for views of at most nmax vertices. */
uint64_t slot_bytes(uint32_t nmax);

/* Plain form: one source (task index 'task') per cooperative launch of
 * 'blocks' workgroups; modes MODE_DEP and MODE_PIVOT, FP64 in the reference's
 * successor order.  plain_max_blocks gives the co-resident limit. */
hipError_t plain_max_blocks(int mode, int cu_count, uint32_t *blocks);
hipError_t launch_plain(int mode, const BrandesParams &p, uint32_t task, uint32_t blocks,
                        hipStream_t stream);

/* Dynamic LDS an LDS-view workgroup can hold (bytes; rows of 16 bytes). */
hipError_t lds_view_capacity(bool half, uint32_t *bytes);

/* Launch the persistent kernel (grid = slots workgroups) on 'stream'. */
hipError_t launch_brandes(int mode, int prec, bool ordered, const BrandesParams &p,
                          uint32_t slots, hipStream_t stream);

/* intra_bc[gid(v)] += sum over s (ascending) of dmat[doff_c + s*n_c + v] for
 * the concatenated cluster vertices [vbegin, vbegin + vcount) whose cluster
 * has a matrix in this chunk (doff_c != ~0). */
hipError_t launch_intra_reduce(const BrandesParams &p, uint64_t vbegin, uint64_t vcount,
                               const uint32_t *v_cluster, double *intra, hipStream_t stream);

/* bc[v] -= intra[v] * w_p for every pivot p of v's cluster, pivots in order. */
hipError_t launch_pivot_subtract(uint64_t nv, const uint32_t *v_gid, const uint32_t *v_cluster,
                                 const uint32_t *cpiv_ptr, const uint32_t *piv_w,
                                 const double *intra, double *bc, hipStream_t stream);

} // namespace fbcdev

#endif
