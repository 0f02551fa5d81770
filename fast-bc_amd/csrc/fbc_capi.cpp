/*
 * C ABI of the gfx950 betweenness kernels (declared in include/fbc_gpu.h).
 *
 * Owns device memory, builds the device CSR layouts (packed {dst, w} arcs,
 * uint32 row starts, reverse CSR, per-cluster local CSRs) and drives the
 * persistent kernel in brandes.hip.
 */
#include "fbc_gpu.h"
#include "fbc_internal.h"
#include "fbc_kmeans.h"
#include "fbc_classes.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

using fbcdev::BrandesParams;

/* Device memory of fbc_dep_batch_noalloc, the entry point for callers that
 * cannot allocate: NOALLOC_CTL_BYTES of control words (work-queue head, error
 * bits, traversed edges, work counters), then the slots of the sources in
 * flight.  Every GPU the library runs on holds this array from the
 * library's load until the process ends.  Its name is a host-side stand-in:
 * host code uses only the per-device address that fbc_create obtains with
 * hipGetSymbolAddress. */
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr size_t NOALLOC_CTL_BYTES = 256;
__device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex noalloc_mu;
std::vector<uint8_t *> noalloc_ptr;   /* by device: the array's address there */
}

struct fbc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    uint32_t delta_override = 0, slots_override = 0;
    uint32_t half_max = fbcdev::HALF_TEAM_MAX_N;   /* views up to this many vertices: half-wavefront teams */
    uint32_t sort_cap = 0;     /* workgroup teams: near-queue sort chunk (0: off, the default) */
    uint32_t level_sort = 0;   /* workgroup teams: DAG-level sort chunk (0: off, the default) */
    uint32_t scratch_mb = 0;          /* slot memory limit in MiB (0: automatic) */
    uint64_t class_key_mask = ~0ull;  /* class hashes kept (fewer bits: collisions, a test hook) */
    uint32_t relabel = 1;             /* pivot phase / exact on the locality-relabelled graph */
    /* plain form, last call: sources run by cooperative launches, sources
     * finished on the batched kernel after a refused first launch, and the
     * HIP status of that refusal (0: none) */
    uint64_t plain_sources = 0, plain_fallback_sources = 0, plain_refusal = 0;
    int cu_count = 0;
    size_t total_mem = 0;
    uint8_t *scratch = nullptr;
    size_t scratch_bytes = 0;
    uint32_t *d_queue = nullptr, *d_err = nullptr;
    unsigned long long *d_te = nullptr, *d_stats = nullptr;
    unsigned long long *d_trial = nullptr;   /* autotuner: edges and busy ticks per trial set */
    uint32_t *d_grid_ctl = nullptr;   /* plain kernel: grid-shared traversal counters */
    double last_ms = 0, last_te = 0;
    unsigned long long last_counters[fbcdev::STAT_COUNT] = {};
    double wall_khz = 100000.0;   /* device wall clock (wall_clock64) rate */
    uint8_t *noalloc = nullptr;   /* this device's address of 'scratch' */
    uint32_t autotune = 1;        /* tune delta / team form on large batched runs */
    uint32_t ballot_append = 0;   /* wavefront-aggregated list appends */
    uint32_t tune_log = 0;        /* autotuner: a line per trial launch on stderr */
    uint32_t km_budget_mb = 0;    /* k-means Gram budget in MiB (0: automatic; a test hook) */
    /* cluster evaluation on the LDS-view kernel: 0 never, 1 always, 2 when
     * the clusters whose rows fit in LDS carry at least half of the run's
     * work (lds_fit_share, set by the cluster evaluation for each launch) and
     * every team gets 32 tasks or more */
    uint32_t lds_view = 2;
    double lds_fit_share = 0;
    /* the last batched run: its delta and half-team limit, and whether they
     * were measured by that run (1), taken from the graph's cache (2) or
     * fixed (0) */
    uint32_t last_delta = 0, last_half_max = 0, last_tuned = 0;
    uint32_t last_lds = 0;   /* the last batched run used the LDS-view kernel */
    /* the last run met a sigma of 2^53 or more (exact here; the reference's
     * FP64 count would be rounded) */
    uint32_t last_sigma_inexact = 0;
};

/* An autotuned launch configuration. */
struct TuneChoice {
    bool set = false;
    uint32_t delta = 0, half_max = 0;
    bool lds = false;   /* cluster evaluation on the LDS-view kernel */
};

struct fbc_graph {
    fbc_ctx *ctx = nullptr;
    uint32_t n = 0;
    uint64_t m = 0;
    uint32_t maxw = 0;
    double meanw = 0;
    std::vector<uint32_t> rp, col, w, rrp, rsrc, rw;   /* host copies */
    uint2 *d_arcs = nullptr, *d_rarcs = nullptr;
    uint32_t *d_rp = nullptr, *d_rrp = nullptr;
    /* Locality-relabelled copy for the modes whose sums are order-free
     * (pivot phase, exact): vertex ids regrouped so that each 16-id block is a
     * BFS-grown blob of neighbouring vertices; built on first use. */
    bool blocked = false;
    std::vector<uint32_t> perm;                 /* original id -> blocked id */
    uint32_t *d_perm = nullptr;
    uint2 *d_barcs = nullptr, *d_brarcs = nullptr;
    uint4 *d_bell = nullptr;
    uint32_t *d_brp = nullptr, *d_brrp = nullptr;
    /* the autotuner's choices for this graph: pivot phase, exact, clusters */
    TuneChoice tuned[3];
};

struct fbc_clusters {
    const fbc_graph *g = nullptr;
    uint32_t ncl = 0;
    uint32_t nmax = 0;
    std::vector<uint32_t> n, bptr, blocal, gid, vcl;
    std::vector<uint64_t> rpoff, voff, vioff, edges;
    uint64_t vi_total = 0;
    uint32_t *d_n = nullptr, *d_gid = nullptr, *d_bptr = nullptr, *d_blocal = nullptr,
             *d_vcl = nullptr, *d_rp = nullptr, *d_rrp = nullptr;
    uint64_t *d_rpoff = nullptr, *d_voff = nullptr, *d_vioff = nullptr, *d_doff = nullptr;
    uint2 *d_arcs = nullptr, *d_rarcs = nullptr;
    uint4 *d_ell = nullptr;
    /* VertexInfo rows left on the device by fbc_cluster_eval_resident */
    uint32_t *d_vil = nullptr;
    int32_t *d_vic = nullptr;
    /* classes of the last fbc_cluster_classes (device, in class order) */
    uint64_t ncls = 0;
    std::vector<uint32_t> cls_ptr, cls_size;
    uint32_t *d_cls_cluster = nullptr, *d_cls_grow = nullptr, *d_cls_ptr = nullptr, *d_rmin = nullptr;
};

namespace {

int fail(fbc_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(fbc_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, e == hipErrorOutOfMemory ? FBC_E_OOM : FBC_E_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
    } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (!count) count = 1;
    return hipMalloc((void **)p, count * sizeof(T));
}

template <typename T>
hipError_t upload(T **p, const std::vector<T> &h, hipStream_t s)
{
    hipError_t e = dalloc(p, h.size());
    if (e != hipSuccess || h.empty()) return e;
    e = hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

template <typename T>
hipError_t upload_n(T **p, const T *h, size_t n, hipStream_t s)
{
    hipError_t e = dalloc(p, n);
    if (e != hipSuccess || !n) return e;
    e = hipMemcpyAsync(*p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

/* Near-far window: a few mean arc weights. */
uint32_t auto_delta(const fbc_ctx *ctx, const fbc_graph *g)
{
    if (ctx->delta_override) return ctx->delta_override;
    double d = g->meanw * 4.0;
    if (d < 1) d = 1;
    if (d > 1e9) d = 1e9;
    return (uint32_t)d;
}

/* Make room for 'slots' slots of views up to nmax vertices; returns the slot
 * count actually available. */
int ensure_scratch(fbc_ctx *ctx, uint32_t nmax, uint32_t want_slots, uint32_t *slots_out, uint32_t half_max,
                   bool lds = false)
{
    const uint64_t sb = fbcdev::slot_bytes(nmax);
    /* whole-graph slots: 8 per CU (on the 1M-vertex grid's pivot phase 4, 6, 8
     * and 12 per CU took 3277, 3206, 3159 and 3162 ms); a wavefront-per-source
     * workgroup holds 4 slots, 8 such workgroups per CU */
    const int kind = fbcdev::team_kind(nmax, half_max);
    /* the LDS-view kernel: one workgroup of LDS_TEAMS_* teams per CU */
    const uint32_t lds_teams = kind == fbcdev::TEAM_HALF ? fbcdev::LDS_TEAMS_HALF : fbcdev::LDS_TEAMS_WAVE;
    const uint32_t per_cu =
        lds ? lds_teams : kind == fbcdev::TEAM_HALF ? 64u : kind == fbcdev::TEAM_WAVE ? 32u : 8u;
    uint32_t slots = ctx->slots_override ? ctx->slots_override : (uint32_t)ctx->cu_count * per_cu;
    if (slots > want_slots) slots = want_slots;
    if (slots < 1) slots = 1;
    /* slots (40 B per vertex each) are also bounded by memory: half the
     * device, less what is not free now (the current scratch aside) and a
     * sixteenth kept for the caller (8 per CU of a 1M-vertex graph take 86 GB;
     * a 4M-vertex graph gets ~3 per CU, whose pivot phase runs 12% faster than
     * with a third of the device) */
    size_t budget = ctx->scratch_mb ? (size_t)ctx->scratch_mb << 20 : ctx->total_mem / 2;
    if (!ctx->scratch_mb) {
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
            const size_t avail = freeb + ctx->scratch_bytes, reserve = ctx->total_mem / 16;
            budget = std::min(budget, avail > 2 * reserve ? avail - reserve : avail / 2);
        }
    }
    if ((uint64_t)slots * sb > budget) slots = (uint32_t)std::max<uint64_t>(1, budget / sb);
    /* the wavefront kernels' workgroups address WAVES (HALVES) slots each:
     * hold whole workgroups' worth */
    const uint32_t unit =
        lds ? lds_teams : kind == fbcdev::TEAM_HALF ? fbcdev::HALVES : kind == fbcdev::TEAM_WAVE ? fbcdev::WAVES : 1;
    slots = (slots + unit - 1) / unit * unit;
    if ((size_t)slots * sb > ctx->scratch_bytes) {
        dfree(ctx->scratch);
        ctx->scratch_bytes = 0;
        /* fewer slots when the device cannot hold them all */
        for (;;) {
            const size_t need = (size_t)slots * sb;
            hipError_t e = hipMalloc((void **)&ctx->scratch, need);
            if (e == hipSuccess) {
                ctx->scratch_bytes = need;
                break;
            }
            (void)hipGetLastError();
            ctx->scratch = nullptr;
            if (e != hipErrorOutOfMemory || slots <= unit) return hip_fail(ctx, e, "scratch allocation");
            slots = std::max(unit, (slots / 2 + unit - 1) / unit * unit);
        }
    }
    *slots_out = slots;
    return 0;
}

void base_params(BrandesParams &p, const fbc_graph *g)
{
    std::memset(&p, 0, sizeof(p));
    p.arcs = g->d_arcs;
    p.rp = g->d_rp;
    p.rarcs = g->d_rarcs;
    p.rrp = g->d_rrp;
    p.n = g->n;
    p.nmax = g->n;
}

/* Statuses with which the runtime refuses a cooperative launch for want of
 * device resources rather than for a bad launch.  hipErrorIllegalState is
 * what a process gets for its first cooperative launch while another process
 * on the device has made cooperative launches of its own (whatever the grid
 * size: 16 and 256 workgroups alike). */
bool refused_cooperative(hipError_t e)
{
    return e == hipErrorCooperativeLaunchTooLarge || e == hipErrorNotSupported ||
           e == hipErrorLaunchOutOfResources || e == hipErrorOutOfMemory || e == hipErrorIllegalState;
}

/* One launch of the batched kernel over q's tasks with q's delta and half-team
 * limit, waited for.  Its time, traversed edges and work counters add to the
 * context's figures of the current run; its error bits are or-ed into *err. */
int timed_launch(fbc_ctx *ctx, int mode, int prec, bool ordered, BrandesParams q, uint32_t *err, double *ms_out,
                 double *te_out)
{
    uint32_t slots = 0;
    int rc = ensure_scratch(ctx, q.nmax, q.ntasks, &slots, q.half_max, q.lds_rows_max > 0);
    if (rc) return rc;
    q.scratch = ctx->scratch;
    q.slot_bytes = fbcdev::slot_bytes(q.nmax);
    q.nslots = slots;
    q.queue_head = ctx->d_queue;
    q.err = ctx->d_err;
    q.te = ctx->d_te;
    q.stats = ctx->d_stats;
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_stats, 0, sizeof(ctx->last_counters), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_queue, 0, sizeof(uint32_t), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), ctx->stream));
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_te, 0, sizeof(unsigned long long), ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, fbcdev::launch_brandes(mode, prec, ordered, q, slots, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    uint32_t e = 0;
    unsigned long long te = 0, c[fbcdev::STAT_COUNT] = {};
    HIP_TRY(ctx, hipMemcpy(&e, ctx->d_err, sizeof(e), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(&te, ctx->d_te, sizeof(te), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(c, ctx->d_stats, sizeof(c), hipMemcpyDeviceToHost));
    for (int i = fbcdev::STAT_T_DIST; i <= fbcdev::STAT_T_DEP; ++i)   /* wall-clock ticks -> ns */
        c[i] = (unsigned long long)((double)c[i] * 1e6 / ctx->wall_khz);
    for (int i = 0; i < fbcdev::STAT_COUNT; ++i) ctx->last_counters[i] += c[i];
    ctx->last_ms += ms;
    ctx->last_te += (double)te;
    *err |= e;
    if (ms_out) *ms_out = ms;
    if (te_out) *te_out = (double)te;
    return 0;
}

/* Put configuration u on launch parameters r: delta, half-team limit and,
 * for cluster evaluation, the LDS-view kernel with as many rows as a
 * workgroup's LDS holds. */
int apply_choice(fbc_ctx *ctx, BrandesParams &r, const TuneChoice &u)
{
    r.delta = u.delta;
    r.half_max = u.half_max;
    r.lds_rows_max = 0;
    if (u.lds) {
        const bool half = fbcdev::team_kind(r.nmax, r.half_max) == fbcdev::TEAM_HALF;
        uint32_t bytes = 0;
        HIP_TRY(ctx, fbcdev::lds_view_capacity(half, &bytes));
        r.lds_rows_max = std::min<uint32_t>(bytes / (uint32_t)sizeof(uint4), r.nmax);
    }
    return 0;
}

/* The autotuner's slot for a mode (-1: the mode is not tuned). */
int tune_slot(int mode)
{
    return mode == fbcdev::MODE_PIVOT ? 0 : mode == fbcdev::MODE_EXACT ? 1 : mode == fbcdev::MODE_CLUSTER ? 2 : -1;
}

/* The batched form of a run.  Large runs of the additive modes (pivot phase,
 * exact BC, cluster evaluation: every task's output is its own) on a graph
 * with no tuned configuration yet tune one on their own first tasks, then
 * keep it for the graph.  No task runs twice.
 *
 * 1. Region: one launch of one task per slot in which task q runs with delta
 *    candidate q mod C (powers of two of the mean arc weight, 8 down to 1/16,
 *    distinct as integers); the kernel adds each task's traversed edges and
 *    busy time to its candidate, and the most edges per busy tick wins.
 *    Candidates apart by large factors separate reliably this way (a
 *    32K-vertex R-MAT graph whose weights average 128: exact BC 1916 / 1218 /
 *    1197 / 1357 ms at delta 511 / 8 / 12 / 32), but a greedier candidate
 *    takes memory bandwidth from the others in a mixed launch, so close ones
 *    are ranked by:
 *    The pivot phase of a clustered run skips this launch when the cluster
 *    evaluation has tuned on the same graph (its views are pieces of it, and
 *    its region launch is short): the walk starts at that choice instead.
 *    A whole-graph region launch pays for its slowest candidates: on the
 *    1M-vertex grid 850 ms for work that the chosen delta does in ~460.
 * 2. Walk: launches of one task per slot, all at one delta, by edges per
 *    millisecond: the region's choice, then its smaller neighbours while
 *    they win, else its larger ones while they win (the 1M-vertex grid's
 *    pivot phase: 3,223 ms at 2 mean weights, 3,312 at 4).
 * 3. Team form, for views small enough for half-wavefront teams: one launch
 *    of whole wavefronts at the chosen delta against the walk's figure.
 * Each step runs only while a full wave of tasks would remain for the final
 * launch; runs of fewer than two tasks per slot are not tuned. */
int run_batched(fbc_ctx *ctx, fbc_graph *g, int mode, int prec, bool ordered, BrandesParams &p, uint32_t *err)
{
    TuneChoice def;
    def.delta = auto_delta(ctx, g);
    def.half_max = ctx->half_max;
    const bool cluster = mode == fbcdev::MODE_CLUSTER;
    /* the LDS-view kernel: measured on the 1M-vertex grid's clusters (81% of
     * the work in views of up to ~9,100 vertices, which fit) 1,116 -> 1,091
     * ms; on the 65K-vertex grid's (8 tasks per team) 8.4 ms either way */
    const uint32_t lds_teams = fbcdev::team_kind(p.nmax, ctx->half_max) == fbcdev::TEAM_HALF
                                   ? fbcdev::LDS_TEAMS_HALF
                                   : fbcdev::LDS_TEAMS_WAVE;
    def.lds = cluster && (ctx->lds_view == 1 ||
                          (ctx->lds_view == 2 && ctx->lds_fit_share >= 0.5 &&
                           (uint64_t)p.ntasks >= 32ull * lds_teams * (uint64_t)ctx->cu_count));
    const int ts = tune_slot(mode);
    const bool tunable = ctx->autotune && !ctx->delta_override && ts >= 0;
    ctx->last_tuned = 0;
    TuneChoice use = def;
    if (tunable && g->tuned[ts].set) {
        use = g->tuned[ts];
        use.lds = def.lds;   /* a property of this run's views, not of the graph */
        ctx->last_tuned = 2;
    } else if (tunable) {
        auto tlog = [&](const char *step, uint32_t delta, uint32_t half, uint64_t n, double ms, double te) {
            if (ctx->tune_log)
                std::fprintf(stderr, "[fbc tune] mode %d %s: delta %u, half_team_max %u, %llu tasks, %.1f ms, %.4g edges/ms\n",
                             mode, step, delta, half, (unsigned long long)n, ms, te / std::max(ms, 1e-3));
        };
        uint32_t slots = 0;
        int rc = ensure_scratch(ctx, p.nmax, p.ntasks, &slots, def.half_max, def.lds);
        if (rc) return rc;
        if ((uint64_t)p.ntasks >= 2ull * slots) {
            std::vector<uint32_t> cand;   /* descending */
            for (double m = 8.0; m >= 1.0 / 16 && cand.size() < fbcdev::TRIAL_MAX; m *= 0.5) {
                const uint32_t d = (uint32_t)std::min(std::max(g->meanw * m, 1.0), 1e9);
                if (cand.empty() || d < cand.back()) cand.push_back(d);
            }
            const uint32_t C = (uint32_t)cand.size();
            uint64_t done = 0;
            uint32_t bi = 0;
            const TuneChoice &cl = g->tuned[tune_slot(fbcdev::MODE_CLUSTER)];
            if (mode == fbcdev::MODE_PIVOT && cl.set) {
                /* the cluster evaluation's choice on this graph: the walk
                 * starts at the candidate nearest to it */
                double bd = 1e300;
                for (uint32_t k = 0; k < C; ++k) {
                    const double dd = std::fabs(std::log((double)cand[k] / (double)std::max(cl.delta, 1u)));
                    if (dd < bd) {
                        bd = dd;
                        bi = k;
                    }
                }
                if (ctx->tune_log)
                    std::fprintf(stderr, "[fbc tune] mode %d: no region launch, the walk starts at delta %u (cluster "
                                         "evaluation's choice %u)\n",
                                 mode, cand[bi], cl.delta);
            } else {
                /* 1. region */
                {
                    BrandesParams q = p;
                    q.ntasks = slots;
                    if ((rc = apply_choice(ctx, q, def))) return rc;
                    q.trial_sets = C;
                    for (uint32_t k = 0; k < C; ++k) q.trial_delta[k] = cand[k];
                    q.trial_stats = ctx->d_trial;
                    HIP_TRY(ctx, hipMemsetAsync(ctx->d_trial, 0, 2 * fbcdev::TRIAL_MAX * sizeof(unsigned long long),
                                                ctx->stream));
                    double ms = 0, te = 0;
                    if ((rc = timed_launch(ctx, mode, prec, ordered, q, err, &ms, &te))) return rc;
                    done += slots;
                    tlog("region", 0, def.half_max, slots, ms, te);
                }
                unsigned long long st[2 * fbcdev::TRIAL_MAX] = {};
                HIP_TRY(ctx, hipMemcpy(st, ctx->d_trial, sizeof(st), hipMemcpyDeviceToHost));
                double br = -1;
                for (uint32_t k = 0; k < C; ++k) {
                    const double rate = st[2 * k + 1] ? (double)st[2 * k] / (double)st[2 * k + 1] : 0.0;
                    if (ctx->tune_log)
                        std::fprintf(stderr, "[fbc tune]   candidate delta %u: %llu edges in %llu busy ticks, %.4g per tick\n",
                                     cand[k], st[2 * k], st[2 * k + 1], rate);
                    if (rate > br) {
                        br = rate;
                        bi = k;
                    }
                }
            }
            /* one launch of n tasks at one configuration; its edges per ms */
            auto pure = [&](const TuneChoice &u, uint32_t n, double *rate) -> int {
                BrandesParams r = p;
                r.task_base = (uint32_t)done;
                r.ntasks = n;
                int e = apply_choice(ctx, r, u);
                if (e) return e;
                double ms = 0, te = 0;
                e = timed_launch(ctx, mode, prec, ordered, r, err, &ms, &te);
                if (e) return e;
                done += n;
                *rate = te / std::max(ms, 1e-3);
                tlog("trial", u.delta, u.half_max, n, ms, te);
                return 0;
            };
            auto room = [&](uint64_t n) { return p.ntasks - done >= n + slots; };
            /* 2. walk */
            use.delta = cand[bi];
            double best = -1;
            if (room(slots)) {
                if ((rc = pure(use, slots, &best))) return rc;
                uint32_t at = bi;
                for (int dir : {+1, -1}) {   /* +1: smaller deltas */
                    bool moved = false;
                    for (int k = (int)at + dir; k >= 0 && k < (int)C && room(slots); k += dir) {
                        TuneChoice u = use;
                        u.delta = cand[(uint32_t)k];
                        double r = 0;
                        if ((rc = pure(u, slots, &r))) return rc;
                        if (r <= best) break;
                        best = r;
                        at = (uint32_t)k;
                        moved = true;
                    }
                    if (moved) break;
                }
                use.delta = cand[at];
            }
            /* 3. team form */
            if (best > 0 && fbcdev::team_kind(p.nmax, def.half_max) == fbcdev::TEAM_HALF) {
                uint32_t wave_slots = 0;
                if ((rc = ensure_scratch(ctx, p.nmax, p.ntasks, &wave_slots, 0, use.lds))) return rc;
                if (room(wave_slots)) {
                    TuneChoice u = use;
                    u.half_max = 0;
                    double r = 0;
                    if ((rc = pure(u, wave_slots, &r))) return rc;
                    if (r > best) {
                        use.half_max = 0;
                        best = r;
                    }
                }
            }
            use.set = true;
            g->tuned[ts] = use;
            ctx->last_tuned = 1;
            ctx->last_delta = use.delta;
            ctx->last_half_max = use.half_max;
            ctx->last_lds = use.lds;
            BrandesParams r = p;
            r.task_base = (uint32_t)done;
            r.ntasks = p.ntasks - (uint32_t)done;
            if ((rc = apply_choice(ctx, r, use))) return rc;
            if (!r.ntasks) return 0;
            double ms = 0, te = 0;
            if ((rc = timed_launch(ctx, mode, prec, ordered, r, err, &ms, &te))) return rc;
            tlog("rest", use.delta, use.half_max, r.ntasks, ms, te);
            return 0;
        }
    }
    ctx->last_delta = use.delta;
    ctx->last_half_max = use.half_max;
    ctx->last_lds = use.lds;
    BrandesParams q = p;
    if (int rc = apply_choice(ctx, q, use)) return rc;
    return timed_launch(ctx, mode, prec, ordered, q, err, nullptr, nullptr);
}

/* Launch one run over p.ntasks tasks and wait for it: the persistent batched
 * kernel (autotuned, run_batched), or (plain) one cooperative grid launch per
 * task. */
int run(fbc_ctx *ctx, const fbc_graph *cg, int mode, int prec, bool ordered, BrandesParams &p,
        bool plain = false)
{
    fbc_graph *g = const_cast<fbc_graph *>(cg);   /* the autotuner's cache */
    ctx->last_ms = 0;
    ctx->last_te = 0;
    std::memset(ctx->last_counters, 0, sizeof(ctx->last_counters));
    ctx->plain_sources = ctx->plain_fallback_sources = ctx->plain_refusal = 0;
    ctx->last_sigma_inexact = 0;
    if (p.ntasks == 0) return 0;
    p.sort_cap = ctx->sort_cap;
    p.level_sort = ctx->level_sort;
    p.ballot_append = ctx->ballot_append;
    uint32_t err = 0;
    int rc = 0;
    if (!plain) {
        rc = run_batched(ctx, g, mode, prec, ordered, p, &err);
        if (rc) return rc;
    } else {
        uint32_t slots = 0;
        rc = ensure_scratch(ctx, p.nmax, 1, &slots, ctx->half_max);
        if (rc) return rc;
        p.scratch = ctx->scratch;
        p.slot_bytes = fbcdev::slot_bytes(p.nmax);
        p.nslots = slots;
        p.half_max = ctx->half_max;
        p.delta = auto_delta(ctx, g);
        p.queue_head = ctx->d_queue;
        p.err = ctx->d_err;
        p.te = ctx->d_te;
        p.stats = ctx->d_stats;
        ctx->last_delta = p.delta;
        ctx->last_half_max = p.half_max;
        ctx->last_tuned = 0;
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_stats, 0, sizeof(ctx->last_counters), ctx->stream));
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_err, 0, sizeof(uint32_t), ctx->stream));
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_te, 0, sizeof(unsigned long long), ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        uint32_t blocks = 0;
        HIP_TRY(ctx, fbcdev::plain_max_blocks(mode, ctx->cu_count, &blocks));
        /* 16 workgroups unless set: every near round and DAG level ends in a
         * grid barrier, whose cost grows with the grid while a thin frontier
         * leaves most workgroups idle (on the 1M-vertex grid's pivots 82 ms per
         * source at 16 workgroups, 228 ms at one per CU) */
        const uint32_t want = ctx->slots_override ? ctx->slots_override : std::min<uint32_t>(16u, ctx->cu_count);
        if (blocks > want) blocks = want;
        p.grid_ctl = ctx->d_grid_ctl;
        uint32_t t = 0;
        for (; t < p.ntasks; ++t) {
            const hipError_t e = fbcdev::launch_plain(mode, p, t, blocks, ctx->stream);
            if (e == hipSuccess) continue;
            (void)hipGetLastError();
            /* Only a refusal of the very first cooperative launch, with a
             * status that says the device cannot take one now (another
             * process holds its cooperative-launch resources), is finished on
             * the batched kernel; any other failure is returned. */
            if (t != 0 || !refused_cooperative(e)) {
                (void)hipStreamSynchronize(ctx->stream);
                const std::string what = "cooperative launch of the plain form (source " + std::to_string(t) +
                                         " of " + std::to_string(p.ntasks) + ", " + std::to_string(blocks) +
                                         " workgroups, HIP status " + std::to_string((int)e) + ")";
                return hip_fail(ctx, e, what.c_str());
            }
            ctx->plain_refusal = (uint64_t)e;
            break;
        }
        ctx->plain_sources = t;
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        unsigned long long te = 0;
        HIP_TRY(ctx, hipMemcpy(&err, ctx->d_err, sizeof(err), hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(&te, ctx->d_te, sizeof(te), hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(ctx->last_counters, ctx->d_stats, sizeof(ctx->last_counters), hipMemcpyDeviceToHost));
        for (int i = fbcdev::STAT_T_DIST; i <= fbcdev::STAT_T_DEP; ++i)   /* wall-clock ticks -> ns */
            ctx->last_counters[i] = (unsigned long long)((double)ctx->last_counters[i] * 1e6 / ctx->wall_khz);
        ctx->last_ms = ms;
        ctx->last_te = (double)te;
        if (t < p.ntasks) {
            /* The remaining sources run on the batched kernel, which computes
             * the same values (FP64, reference successor order); the count is
             * reported by fbc_last_counters. */
            ctx->plain_fallback_sources = p.ntasks - t;
            BrandesParams q = p;
            rc = timed_launch(ctx, mode, 0, true, q, &err, nullptr, nullptr);
            if (rc) return rc;
        }
    }
    ctx->last_sigma_inexact = (err & fbcdev::ERR_SIGMA_INEXACT) ? 1 : 0;
    if (err & fbcdev::ERR_SOURCE) return fail(ctx, FBC_E_ARG, "source out of range");
    if (err & fbcdev::ERR_SIGMA)
        return fail(ctx, FBC_E_SIGMA_OVERFLOW, "shortest-path count exceeds 2^64-1");
    if (err & fbcdev::ERR_DIST) return fail(ctx, FBC_E_DIST_OVERFLOW, "path length exceeds 2^32-2");
    return 0;
}

bool valid_prec(int prec) { return prec == FBC_PREC_F64 || prec == FBC_PREC_F32C; }

} // namespace

extern "C" {

int fbc_create(int device, fbc_ctx **out)
{
    if (!out) return FBC_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return FBC_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FBC_E_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FBC_E_DEVICE;
    fbc_ctx *ctx = new (std::nothrow) fbc_ctx;
    if (!ctx) return FBC_E_OOM;
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->total_mem = prop.totalGlobalMem;
    int wall_khz = 0;
    if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wall_khz > 0)
        ctx->wall_khz = wall_khz;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        dalloc(&ctx->d_queue, 1) != hipSuccess || dalloc(&ctx->d_err, 1) != hipSuccess ||
        dalloc(&ctx->d_te, 1) != hipSuccess || dalloc(&ctx->d_stats, fbcdev::STAT_COUNT) != hipSuccess ||
        dalloc(&ctx->d_trial, 2 * fbcdev::TRIAL_MAX) != hipSuccess ||
        dalloc(&ctx->d_grid_ctl, 16) != hipSuccess || hipMemset(ctx->d_grid_ctl, 0, 16 * sizeof(uint32_t)) != hipSuccess) {
        fbc_destroy(ctx);
        return FBC_E_HIP;
    }
    {
        /* once per device, with the device current: the address of the
         * allocation-free entry point's array on this device */
        std::lock_guard<std::mutex> lk(noalloc_mu);
        if (noalloc_ptr.size() < (size_t)ndev) noalloc_ptr.resize((size_t)ndev, nullptr);
        if (!noalloc_ptr[(size_t)device]) {
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                fbc_destroy(ctx);
                return FBC_E_HIP;
            }
            noalloc_ptr[(size_t)device] = static_cast<uint8_t *>(ptr);
        }
        ctx->noalloc = noalloc_ptr[(size_t)device];
    }
    *out = ctx;
    return 0;
}

void fbc_destroy(fbc_ctx *ctx)
{
    if (!ctx) return;
    dfree(ctx->scratch);
    dfree(ctx->d_queue);
    dfree(ctx->d_err);
    dfree(ctx->d_te);
    dfree(ctx->d_stats);
    dfree(ctx->d_trial);
    dfree(ctx->d_grid_ctl);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *fbc_last_error(const fbc_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int fbc_set_tuning(fbc_ctx *ctx, uint32_t bucket_delta, uint32_t slots)
{
    if (!ctx) return FBC_E_ARG;
    ctx->delta_override = bucket_delta;
    ctx->slots_override = slots;
    return 0;
}

int fbc_set_option(fbc_ctx *ctx, const char *name, int64_t value)
{
    if (!ctx || !name) return FBC_E_ARG;
    if (value < 0 || value > 0xFFFFFFFFll) return fail(ctx, FBC_E_ARG, "option value out of range");
    const uint32_t v = (uint32_t)value;
    if (!std::strcmp(name, "delta")) ctx->delta_override = v;
    else if (!std::strcmp(name, "slots")) ctx->slots_override = v;
    else if (!std::strcmp(name, "frontier_sort")) ctx->sort_cap = std::min(v, fbcdev::SORT_CHUNK_MAX);
    else if (!std::strcmp(name, "level_sort")) ctx->level_sort = std::min(v, fbcdev::SORT_CHUNK_MAX);
    else if (!std::strcmp(name, "relabel")) ctx->relabel = v ? 1 : 0;
    else if (!std::strcmp(name, "scratch_mb")) ctx->scratch_mb = v;
    else if (!std::strcmp(name, "half_team_max")) ctx->half_max = std::min<uint32_t>(v, fbcdev::WAVE_TEAM_MAX_N);
    else if (!std::strcmp(name, "autotune")) ctx->autotune = v ? 1 : 0;
    else if (!std::strcmp(name, "ballot_append")) ctx->ballot_append = v ? 1 : 0;
    else if (!std::strcmp(name, "lds_view")) ctx->lds_view = std::min<uint32_t>(v, 2);
    else if (!std::strcmp(name, "tune_log")) ctx->tune_log = v ? 1 : 0;
    else if (!std::strcmp(name, "kmeans_budget_mb")) ctx->km_budget_mb = v;
    else if (!std::strcmp(name, "class_key_bits"))
        ctx->class_key_mask = (v == 0 || v >= 64) ? ~0ull : ((1ull << v) - 1);
    else return fail(ctx, FBC_E_ARG, std::string("unknown option: ") + name);
    return 0;
}

int fbc_graph_upload(fbc_ctx *ctx, uint32_t n, uint64_t m, const uint64_t *row_ptr,
                     const uint32_t *col, const uint32_t *w, fbc_graph **out)
{
    if (!ctx || !out || !row_ptr || (m && (!col || !w))) return fail(ctx, FBC_E_ARG, "null argument");
    *out = nullptr;
    if (n == 0 || n >= 0xFFFFFFF0u) return fail(ctx, FBC_E_ARG, "vertex count out of range");
    if (m >= 0xFFFFFFF0ull) return fail(ctx, FBC_E_UNSUPPORTED, "more than 2^32-16 arcs");
    if (row_ptr[0] != 0 || row_ptr[n] != m) return fail(ctx, FBC_E_ARG, "row_ptr must run from 0 to m");
    fbc_graph *g = new (std::nothrow) fbc_graph;
    if (!g) return fail(ctx, FBC_E_OOM, "host allocation");
    g->ctx = ctx;
    g->n = n;
    g->m = m;
    try {
        g->rp.resize((size_t)n + 1);
        g->col.assign(col, col + m);
        g->w.assign(w, w + m);
        double wsum = 0;
        for (uint32_t v = 0; v <= n; ++v) {
            g->rp[v] = (uint32_t)row_ptr[v];
            if (v < n && row_ptr[v + 1] < row_ptr[v]) {
                delete g;
                return fail(ctx, FBC_E_ARG, "row_ptr is not monotone");
            }
        }
        for (uint64_t e = 0; e < m; ++e) {
            if (col[e] >= n || w[e] == 0) {
                delete g;
                return fail(ctx, FBC_E_ARG, "arc target out of range or zero weight");
            }
            g->maxw = std::max(g->maxw, w[e]);
            wsum += w[e];
        }
        g->meanw = m ? wsum / (double)m : 1.0;
        /* reverse CSR, rows in ascending source order */
        g->rrp.assign((size_t)n + 1, 0);
        for (uint64_t e = 0; e < m; ++e) g->rrp[col[e] + 1]++;
        for (uint32_t v = 0; v < n; ++v) g->rrp[v + 1] += g->rrp[v];
        g->rsrc.resize(m);
        g->rw.resize(m);
        std::vector<uint32_t> fill(g->rrp.begin(), g->rrp.end() - 1);
        for (uint32_t u = 0; u < n; ++u)
            for (uint32_t e = g->rp[u]; e < g->rp[u + 1]; ++e) {
                uint32_t pos = fill[col[e]]++;
                g->rsrc[pos] = u;
                g->rw[pos] = w[e];
            }
        std::vector<uint2> arcs(m), rarcs(m);
        for (uint64_t e = 0; e < m; ++e) {
            arcs[e] = make_uint2(col[e], w[e]);
            rarcs[e] = make_uint2(g->rsrc[e], g->rw[e]);
        }
        hipError_t e;
        if ((e = upload(&g->d_arcs, arcs, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_rarcs, rarcs, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_rp, g->rp, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_rrp, g->rrp, ctx->stream)) != hipSuccess) {
            fbc_graph_free(g);
            return hip_fail(ctx, e, "graph upload");
        }
    } catch (const std::bad_alloc &) {
        delete g;
        return fail(ctx, FBC_E_OOM, "host allocation");
    }
    *out = g;
    return 0;
}

void fbc_graph_free(fbc_graph *g)
{
    if (!g) return;
    dfree(g->d_arcs);
    dfree(g->d_rarcs);
    dfree(g->d_rp);
    dfree(g->d_rrp);
    dfree(g->d_perm);
    dfree(g->d_barcs);
    dfree(g->d_bell);
    dfree(g->d_brarcs);
    dfree(g->d_brp);
    dfree(g->d_brrp);
    delete g;
}

/* Blocked vertex order, hierarchical: the graph is cut into BFS-grown blobs
 * of 4096 vertices (seeds in ascending id), each blob into blobs of 256, each
 * of those into blobs of 16, and vertices are numbered blob by blob.  A
 * 16-vertex blob's per-vertex state then shares cache lines, and blobs that
 * are close in the graph are close in memory at every scale (the way a
 * space-filling curve orders a grid). */
/* Short row of vertex v (arcs row[0, deg)): (weight << 16) | (uint16)(target
 * - v) per arc, padded with ELL_NONE, or ELL_WIDE when the row does not fit. */
static uint4 ell_row(uint32_t v, const uint2 *row, uint32_t deg)
{
    uint32_t e[fbcdev::ELL_W];
    bool fits = deg <= fbcdev::ELL_W;
    for (uint32_t j = 0; j < fbcdev::ELL_W; ++j) {
        e[j] = fbcdev::ELL_NONE;
        if (!fits || j >= deg) continue;
        const int64_t d = (int64_t)row[j].x - (int64_t)v;
        if (d < -32768 || d > 32767 || row[j].y > fbcdev::ELL_WMAX) {
            fits = false;
            continue;
        }
        e[j] = (row[j].y << 16) | (uint32_t)(uint16_t)(int16_t)d;
    }
    if (!fits) return make_uint4(fbcdev::ELL_WIDE, fbcdev::ELL_NONE, fbcdev::ELL_NONE, fbcdev::ELL_NONE);
    return make_uint4(e[0], e[1], e[2], e[3]);
}

static void blocked_order(const fbc_graph *g, std::vector<uint32_t> &perm)
{
    static const uint32_t SIZES[] = {4096, 256, 16};
    const uint32_t n = g->n;
    std::vector<uint32_t> region(n, 0), stamp(n, 0);   /* region id of each vertex at this level */
    std::vector<uint32_t> order(n), next_order(n), new_region(n);
    for (uint32_t v = 0; v < n; ++v) order[v] = v;
    /* regions are runs of 'order'; rstart[r] = first position of region r */
    std::vector<uint64_t> rstart = {0, n};
    uint32_t tick = 0;
    for (uint32_t size : SIZES) {
        /* split every region into BFS blobs of 'size' (seeds in the region's
         * order); regions are disjoint, so they split in parallel, each
         * writing its own run of next_order and local blob numbers */
        ++tick;   /* stamp == tick: already placed at this level */
        const uint32_t nreg = (uint32_t)rstart.size() - 1;
        std::vector<uint32_t> nblob(nreg, 0);
#pragma omp parallel
        {
            std::vector<uint32_t> q;
#pragma omp for schedule(dynamic, 16)
            for (uint32_t r = 0; r < nreg; ++r) {
                uint64_t out = rstart[r];
                uint32_t b = 0;
                for (uint64_t i = rstart[r]; i < rstart[r + 1]; ++i) {
                    const uint32_t s = order[i];
                    if (stamp[s] == tick) continue;
                    q.clear();
                    q.push_back(s);
                    stamp[s] = tick;
                    for (size_t h = 0; h < q.size() && q.size() < size; ++h) {
                        const uint32_t u = q[h];
                        for (int dir = 0; dir < 2 && q.size() < size; ++dir) {
                            const std::vector<uint32_t> &ptr = dir ? g->rrp : g->rp;
                            const std::vector<uint32_t> &nb = dir ? g->rsrc : g->col;
                            for (uint32_t e = ptr[u]; e < ptr[u + 1] && q.size() < size; ++e) {
                                const uint32_t w = nb[e];
                                if (region[w] != r || stamp[w] == tick) continue;
                                stamp[w] = tick;
                                q.push_back(w);
                            }
                        }
                    }
                    for (uint32_t w : q) {
                        new_region[w] = b;
                        next_order[out++] = w;
                    }
                    ++b;
                }
                nblob[r] = b;
            }
        }
        /* global blob numbers: each region's blobs after those of the
         * regions before it */
        std::vector<uint64_t> boff((size_t)nreg + 1, 0);
        for (uint32_t r = 0; r < nreg; ++r) boff[r + 1] = boff[r] + nblob[r];
        std::vector<uint64_t> next_start((size_t)boff[nreg] + 1, 0);
#pragma omp parallel for schedule(static, 4096)
        for (uint32_t i = 0; i < n; ++i) new_region[next_order[i]] += (uint32_t)boff[region[next_order[i]]];
        for (uint32_t i = 0; i < n; ++i) next_start[new_region[next_order[i]] + 1]++;
        for (uint64_t b = 0; b < boff[nreg]; ++b) next_start[b + 1] += next_start[b];
        order.swap(next_order);
        region.swap(new_region);
        rstart.swap(next_start);
    }
    perm.resize(n);
#pragma omp parallel for schedule(static, 4096)
    for (uint32_t i = 0; i < n; ++i) perm[order[i]] = i;
}

/* Build and upload the relabelled CSRs (forward and reverse, rows in the new
 * order, arcs by ascending new target / source). */
static int ensure_blocked(fbc_ctx *ctx, fbc_graph *g)
{
    if (g->blocked) return 0;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    try {
        blocked_order(g, g->perm);
        std::vector<uint32_t> inv(n);
#pragma omp parallel for schedule(static, 4096)
        for (uint32_t v = 0; v < n; ++v) inv[g->perm[v]] = v;
        auto build = [&](const std::vector<uint32_t> &ptr, const std::vector<uint32_t> &nb,
                         const std::vector<uint32_t> &wt, std::vector<uint32_t> &bp, std::vector<uint2> &arcs) {
            bp.assign((size_t)n + 1, 0);
            arcs.resize(m);
            for (uint32_t nu = 0; nu < n; ++nu) bp[nu + 1] = bp[nu] + (ptr[inv[nu] + 1] - ptr[inv[nu]]);
#pragma omp parallel for schedule(static, 1024)
            for (uint32_t nu = 0; nu < n; ++nu) {
                const uint32_t u = inv[nu];
                uint2 *row = &arcs[bp[nu]];
                for (uint32_t e = ptr[u]; e < ptr[u + 1]; ++e) row[e - ptr[u]] = make_uint2(g->perm[nb[e]], wt[e]);
                std::sort(row, row + (ptr[u + 1] - ptr[u]), [](const uint2 &a, const uint2 &b) { return a.x < b.x; });
            }
        };
        std::vector<uint32_t> brp, brrp;
        std::vector<uint2> barcs, brarcs;
        build(g->rp, g->col, g->w, brp, barcs);
        build(g->rrp, g->rsrc, g->rw, brrp, brarcs);
        std::vector<uint4> bell(n);
#pragma omp parallel for schedule(static, 4096)
        for (uint32_t nu = 0; nu < n; ++nu) bell[nu] = ell_row(nu, &barcs[brp[nu]], brp[nu + 1] - brp[nu]);
        hipError_t e;
        if ((e = upload(&g->d_bell, bell, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_barcs, barcs, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_brarcs, brarcs, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_brp, brp, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_brrp, brrp, ctx->stream)) != hipSuccess ||
            (e = upload(&g->d_perm, g->perm, ctx->stream)) != hipSuccess)
            return hip_fail(ctx, e, "relabelled graph upload");
    } catch (const std::bad_alloc &) {
        return fail(ctx, FBC_E_OOM, "host allocation");
    }
    g->blocked = true;
    return 0;
}

/* Point p at the relabelled graph. */
static void use_blocked(BrandesParams &p, const fbc_graph *g)
{
    p.arcs = g->d_barcs;
    p.rp = g->d_brp;
    p.rarcs = g->d_brarcs;
    p.rrp = g->d_brrp;
    p.ell = g->d_bell;
}

/* Output-row modes (sssp / dep): sources in chunks that fit a device budget. */
static int row_mode(fbc_ctx *ctx, const fbc_graph *g, int mode, int prec, const uint32_t *src,
                    uint32_t nsrc, uint64_t *dist, uint64_t *sigma, double *dep, bool plain = false)
{
    for (uint32_t i = 0; i < nsrc; ++i)
        if (src[i] >= g->n) return fail(ctx, FBC_E_ARG, "source out of range");
    const size_t row_bytes = (size_t)g->n * (mode == fbcdev::MODE_SSSP ? 16 : 8);
    const size_t budget = (size_t)4 << 30;
    uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(nsrc, budget / row_bytes));
    uint64_t *d_dist = nullptr, *d_sigma = nullptr;
    double *d_dep = nullptr;
    uint32_t *d_src = nullptr;
    int rc = 0;
    double ms = 0, te = 0;
    hipError_t e = dalloc(&d_src, chunk);
    if (e == hipSuccess) {
        if (mode == fbcdev::MODE_SSSP) {
            e = dalloc(&d_dist, (size_t)chunk * g->n);
            if (e == hipSuccess) e = dalloc(&d_sigma, (size_t)chunk * g->n);
        } else {
            e = dalloc(&d_dep, (size_t)chunk * g->n);
        }
    }
    if (e != hipSuccess) rc = hip_fail(ctx, e, "output allocation");
    for (uint32_t s0 = 0; !rc && s0 < nsrc; s0 += chunk) {
        const uint32_t k = std::min(chunk, nsrc - s0);
        BrandesParams p;
        base_params(p, g);
        p.ntasks = k;
        p.task_src = d_src;
        p.out_dist = d_dist;
        p.out_sigma = d_sigma;
        p.out_dep = d_dep;
        if ((e = hipMemcpyAsync(d_src, src + s0, k * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
            (d_dep && (e = hipMemsetAsync(d_dep, 0, (size_t)k * g->n * sizeof(double), ctx->stream)) != hipSuccess)) {
            rc = hip_fail(ctx, e, "chunk setup");
            break;
        }
        rc = run(ctx, g, mode, prec, true, p, plain);
        if (rc) break;
        ms += ctx->last_ms;
        te += ctx->last_te;
        if (mode == fbcdev::MODE_SSSP) {
            e = hipMemcpy(dist + (size_t)s0 * g->n, d_dist, (size_t)k * g->n * 8, hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(sigma + (size_t)s0 * g->n, d_sigma, (size_t)k * g->n * 8, hipMemcpyDeviceToHost);
        } else {
            e = hipMemcpy(dep + (size_t)s0 * g->n, d_dep, (size_t)k * g->n * 8, hipMemcpyDeviceToHost);
        }
        if (e != hipSuccess) rc = hip_fail(ctx, e, "result download");
    }
    dfree(d_src);
    dfree(d_dist);
    dfree(d_sigma);
    dfree(d_dep);
    ctx->last_ms = ms;
    ctx->last_te = te;
    return rc;
}

int fbc_sssp_sigma(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *src, uint32_t nsrc,
                   uint64_t *dist, uint64_t *sigma)
{
    if (!ctx || !g || (nsrc && (!src || !dist || !sigma))) return fail(ctx, FBC_E_ARG, "null argument");
    return row_mode(ctx, g, fbcdev::MODE_SSSP, 0, src, nsrc, dist, sigma, nullptr);
}

int fbc_dep_batch(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *src, uint32_t nsrc, int prec,
                  double *dep)
{
    if (!ctx || !g || (nsrc && (!src || !dep))) return fail(ctx, FBC_E_ARG, "null argument");
    if (!valid_prec(prec)) return fail(ctx, FBC_E_ARG, "unknown precision");
    return row_mode(ctx, g, fbcdev::MODE_DEP, prec, src, nsrc, nullptr, nullptr, dep);
}

int fbc_dep_batch_noalloc(fbc_ctx *ctx, const fbc_graph *g, const uint32_t *d_src, uint32_t nsrc, int prec,
                          double *d_dep)
{
    if (!ctx || !g || (nsrc && (!d_src || !d_dep))) return fail(ctx, FBC_E_ARG, "null argument");
    if (!valid_prec(prec)) return fail(ctx, FBC_E_ARG, "unknown precision");
    if (!ctx->noalloc) return fail(ctx, FBC_E_HIP, "static scratch not resolved on this device");
    ctx->last_ms = 0;
    ctx->last_te = 0;
    ctx->plain_sources = ctx->plain_fallback_sources = ctx->plain_refusal = 0;
    if (nsrc == 0) return 0;
    /* slots: the batched kernel's team for this graph size, as many whole
     * workgroups' worth as the array holds (and the device keeps busy) */
    const uint64_t sb = fbcdev::slot_bytes(g->n);
    const int kind = fbcdev::team_kind(g->n, ctx->half_max);
    const uint64_t unit = kind == fbcdev::TEAM_HALF ? fbcdev::HALVES : kind == fbcdev::TEAM_WAVE ? fbcdev::WAVES : 1;
    const uint64_t per_cu = kind == fbcdev::TEAM_HALF ? 64 : kind == fbcdev::TEAM_WAVE ? 32 : 4;
    const uint64_t fit = (NOALLOC_SCRATCH_BYTES - NOALLOC_CTL_BYTES) / (sb * unit) * unit;
    if (fit == 0) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "graph of %u vertices: one source needs %llu bytes of scratch (%llu with its team), beyond the "
                      "%zu MiB static array of the allocation-free entry point",
                      g->n, (unsigned long long)sb, (unsigned long long)(sb * unit), NOALLOC_SCRATCH_BYTES >> 20);
        return fail(ctx, FBC_E_UNSUPPORTED, msg);
    }
    uint64_t slots = std::min<uint64_t>(nsrc, (uint64_t)ctx->cu_count * per_cu);
    slots = std::min(fit, (slots + unit - 1) / unit * unit);
    uint8_t *const base = ctx->noalloc;
    BrandesParams p;
    base_params(p, g);
    p.ntasks = nsrc;
    p.task_src = d_src;
    p.out_dep = d_dep;
    p.scratch = base + NOALLOC_CTL_BYTES;
    p.slot_bytes = sb;
    p.nslots = (uint32_t)slots;
    p.half_max = ctx->half_max;
    p.delta = auto_delta(ctx, g);
    p.sort_cap = ctx->sort_cap;
    p.level_sort = ctx->level_sort;
    p.ballot_append = ctx->ballot_append;
    p.queue_head = reinterpret_cast<uint32_t *>(base);
    p.err = reinterpret_cast<uint32_t *>(base + 4);
    p.te = reinterpret_cast<unsigned long long *>(base + 8);
    p.stats = reinterpret_cast<unsigned long long *>(base + 16);
    static_assert(16 + 8 * fbcdev::STAT_COUNT <= NOALLOC_CTL_BYTES, "control words");
    /* the last call's control words are still there: clear them */
    HIP_TRY(ctx, fbcdev::launch_clear_words(reinterpret_cast<uint32_t *>(base), NOALLOC_CTL_BYTES / 4, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, fbcdev::launch_brandes(fbcdev::MODE_DEP, prec, true, p, (uint32_t)slots, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    unsigned long long ctl[2 + fbcdev::STAT_COUNT] = {};
    HIP_TRY(ctx, hipMemcpy(ctl, base, sizeof(ctl), hipMemcpyDeviceToHost));
    const uint32_t err = (uint32_t)(ctl[0] >> 32);
    for (int i = 0; i < fbcdev::STAT_COUNT; ++i) ctx->last_counters[i] = ctl[2 + i];
    for (int i = fbcdev::STAT_T_DIST; i <= fbcdev::STAT_T_DEP; ++i)
        ctx->last_counters[i] = (unsigned long long)((double)ctx->last_counters[i] * 1e6 / ctx->wall_khz);
    ctx->last_ms = ms;
    ctx->last_te = (double)ctl[1];
    if (err & fbcdev::ERR_SOURCE) return fail(ctx, FBC_E_ARG, "source out of range");
    if (err & fbcdev::ERR_SIGMA) return fail(ctx, FBC_E_SIGMA_OVERFLOW, "shortest-path count exceeds 2^64-1");
    if (err & fbcdev::ERR_DIST) return fail(ctx, FBC_E_DIST_OVERFLOW, "path length exceeds 2^32-2");
    return 0;
}

int fbc_single_source_dep(fbc_ctx *ctx, const fbc_graph *g, uint32_t src, int prec, double *dep)
{
    if (!ctx || !g || !dep) return fail(ctx, FBC_E_ARG, "null argument");
    if (!valid_prec(prec)) return fail(ctx, FBC_E_ARG, "unknown precision");
    if (prec == FBC_PREC_F32C) return fbc_dep_batch(ctx, g, &src, 1, prec, dep);
    return row_mode(ctx, g, fbcdev::MODE_DEP, prec, &src, 1, nullptr, nullptr, dep, true);
}

int fbc_exact_bc(fbc_ctx *ctx, const fbc_graph *g, int prec, double *bc)
{
    if (!ctx || !g || !bc) return fail(ctx, FBC_E_ARG, "null argument");
    if (!valid_prec(prec)) return fail(ctx, FBC_E_ARG, "unknown precision");
    const bool blocked = ctx->relabel != 0;
    if (blocked) {
        const int rc = ensure_blocked(ctx, const_cast<fbc_graph *>(g));
        if (rc) return rc;
    }
    std::vector<uint32_t> src(g->n);
    std::iota(src.begin(), src.end(), 0u);   /* every vertex (in the graph's own numbering) */
    uint32_t *d_src = nullptr;
    double *d_bc = nullptr;
    hipError_t e;
    if ((e = upload(&d_src, src, ctx->stream)) != hipSuccess || (e = dalloc(&d_bc, g->n)) != hipSuccess ||
        (e = hipMemsetAsync(d_bc, 0, (size_t)g->n * sizeof(double), ctx->stream)) != hipSuccess) {
        dfree(d_src);
        dfree(d_bc);
        return hip_fail(ctx, e, "exact setup");
    }
    BrandesParams p;
    base_params(p, g);
    if (blocked) use_blocked(p, g);
    p.ntasks = g->n;
    p.task_src = d_src;
    p.bc = d_bc;
    int rc = run(ctx, g, fbcdev::MODE_EXACT, prec, false, p);
    if (!rc && blocked) {
        std::vector<double> nb(g->n);
        if ((e = hipMemcpy(nb.data(), d_bc, (size_t)g->n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_fail(ctx, e, "result download");
        else
            for (uint32_t v = 0; v < g->n; ++v) bc[v] = nb[g->perm[v]];
    } else if (!rc && (e = hipMemcpy(bc, d_bc, (size_t)g->n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess) {
        rc = hip_fail(ctx, e, "result download");
    }
    dfree(d_src);
    dfree(d_bc);
    return rc;
}

int fbc_clusters_create(fbc_ctx *ctx, const fbc_graph *g, uint32_t ncl, const uint32_t *cluster_ptr,
                        const uint32_t *cluster_verts, fbc_clusters **out, uint32_t *border_ptr)
{
    if (!ctx || !g || !out || !cluster_ptr || (ncl && !cluster_verts)) return fail(ctx, FBC_E_ARG, "null argument");
    *out = nullptr;
    const uint32_t n = g->n;
    const uint64_t total = cluster_ptr[ncl];
    if (cluster_ptr[0] != 0 || total > n) return fail(ctx, FBC_E_ARG, "bad cluster_ptr");
    fbc_clusters *cl = new (std::nothrow) fbc_clusters;
    if (!cl) return fail(ctx, FBC_E_OOM, "host allocation");
    cl->g = g;
    cl->ncl = ncl;
    try {
        std::vector<int32_t> owner(n, -1);
        std::vector<uint32_t> lid(n, 0);
        for (uint32_t c = 0; c < ncl; ++c) {
            if (cluster_ptr[c + 1] < cluster_ptr[c]) throw std::invalid_argument("bad cluster_ptr");
            for (uint32_t i = cluster_ptr[c]; i < cluster_ptr[c + 1]; ++i) {
                const uint32_t v = cluster_verts[i];
                if (v >= n || owner[v] != -1) throw std::invalid_argument("cluster vertex out of range or repeated");
                if (i > cluster_ptr[c] && v <= cluster_verts[i - 1])
                    throw std::invalid_argument("cluster vertex lists must be ascending");
                owner[v] = (int32_t)c;
                lid[v] = i - cluster_ptr[c];
            }
        }
        cl->n.resize(ncl);
        cl->rpoff.resize(ncl);
        cl->voff.resize(ncl);
        cl->vioff.resize(ncl);
        cl->edges.resize(ncl);
        cl->bptr.assign((size_t)ncl + 1, 0);
        cl->gid.assign(cluster_verts, cluster_verts + total);
        cl->vcl.resize(total);
        /* per position (cluster-major, ascending ids): internal out- and
         * in-degrees and the border flag, then offsets by prefix sums, then the
         * rows filled in parallel; rp/rrp blocks are n_c + 1 entries per
         * cluster, row starts absolute */
        std::vector<uint32_t> odeg(total), ideg(total);
        std::vector<uint8_t> isb(total);
#pragma omp parallel for schedule(dynamic, 4)
        for (uint32_t c = 0; c < ncl; ++c)
            for (uint32_t i = cluster_ptr[c]; i < cluster_ptr[c + 1]; ++i) {
                const uint32_t v = cluster_verts[i];
                uint32_t od = 0, id = 0;
                bool border = false;
                for (uint32_t e = g->rp[v]; e < g->rp[v + 1]; ++e) {
                    if (owner[g->col[e]] == (int32_t)c) ++od;
                    else border = true;
                }
                for (uint32_t e = g->rrp[v]; e < g->rrp[v + 1]; ++e) {
                    if (owner[g->rsrc[e]] == (int32_t)c) ++id;
                    else border = true;
                }
                odeg[i] = od;
                ideg[i] = id;
                isb[i] = border;
                cl->vcl[i] = c;
            }
        std::vector<uint64_t> aoff((size_t)total + 1, 0), roff((size_t)total + 1, 0);
        for (uint64_t i = 0; i < total; ++i) {
            aoff[i + 1] = aoff[i] + odeg[i];
            roff[i + 1] = roff[i] + ideg[i];
        }
        if (aoff[total] > 0xFFFFFFFFull || roff[total] > 0xFFFFFFFFull) throw std::invalid_argument("too many cluster arcs");
        std::vector<uint32_t> rp((size_t)total + ncl), rrp((size_t)total + ncl);
        std::vector<uint2> arcs(aoff[total]), rarcs(roff[total]);
        std::vector<uint4> ell(total);
        std::vector<uint32_t> nb(ncl);
        uint64_t vi = 0;
        for (uint32_t c = 0; c < ncl; ++c) {
            const uint32_t v0 = cluster_ptr[c], nc = cluster_ptr[c + 1] - v0;
            uint32_t cnt = 0;
            for (uint32_t i = v0; i < v0 + nc; ++i) cnt += isb[i];
            nb[c] = cnt;
            cl->n[c] = nc;
            cl->nmax = std::max(cl->nmax, nc);
            cl->voff[c] = v0;
            cl->rpoff[c] = (uint64_t)v0 + c;
            cl->vioff[c] = vi;
            cl->edges[c] = aoff[v0 + nc] - aoff[v0];
            cl->bptr[c + 1] = cl->bptr[c] + cnt;
            vi += (uint64_t)nc * cnt;
        }
        cl->blocal.resize(cl->bptr[ncl]);
#pragma omp parallel for schedule(dynamic, 4)
        for (uint32_t c = 0; c < ncl; ++c) {
            const uint32_t v0 = cluster_ptr[c], nc = cluster_ptr[c + 1] - v0;
            const uint64_t r0 = cl->rpoff[c];
            uint32_t b = cl->bptr[c];
            for (uint32_t i = 0; i < nc; ++i) {
                const uint64_t pos = (uint64_t)v0 + i;
                const uint32_t v = cluster_verts[pos];
                rp[r0 + i] = (uint32_t)aoff[pos];
                rrp[r0 + i] = (uint32_t)roff[pos];
                uint64_t a = aoff[pos], r = roff[pos];
                for (uint32_t e = g->rp[v]; e < g->rp[v + 1]; ++e) {
                    const uint32_t t = g->col[e];
                    if (owner[t] == (int32_t)c) arcs[a++] = make_uint2(lid[t], g->w[e]);
                }
                for (uint32_t e = g->rrp[v]; e < g->rrp[v + 1]; ++e) {
                    const uint32_t s2 = g->rsrc[e];
                    if (owner[s2] == (int32_t)c) rarcs[r++] = make_uint2(lid[s2], g->rw[e]);
                }
                ell[pos] = ell_row(i, arcs.data() + aoff[pos], odeg[pos]);
                if (isb[pos]) cl->blocal[b++] = i;
            }
            rp[r0 + nc] = (uint32_t)aoff[(uint64_t)v0 + nc];
            rrp[r0 + nc] = (uint32_t)roff[(uint64_t)v0 + nc];
        }
        cl->vi_total = vi;
        hipError_t e;
        if ((e = upload(&cl->d_n, cl->n, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_gid, cl->gid, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_bptr, cl->bptr, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_blocal, cl->blocal, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_vcl, cl->vcl, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_rp, rp, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_rrp, rrp, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_arcs, arcs, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_rarcs, rarcs, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_ell, ell, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_rpoff, cl->rpoff, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_voff, cl->voff, ctx->stream)) != hipSuccess ||
            (e = upload(&cl->d_vioff, cl->vioff, ctx->stream)) != hipSuccess ||
            (e = dalloc(&cl->d_doff, ncl)) != hipSuccess) {
            fbc_clusters_free(cl);
            return hip_fail(ctx, e, "cluster upload");
        }
    } catch (const std::bad_alloc &) {
        fbc_clusters_free(cl);
        return fail(ctx, FBC_E_OOM, "host allocation");
    } catch (const std::invalid_argument &ex) {
        fbc_clusters_free(cl);
        return fail(ctx, FBC_E_ARG, ex.what());
    }
    if (border_ptr) std::copy(cl->bptr.begin(), cl->bptr.end(), border_ptr);
    *out = cl;
    return 0;
}

int fbc_clusters_borders(const fbc_clusters *cl, uint32_t *borders)
{
    if (!cl || (cl->bptr[cl->ncl] && !borders)) return FBC_E_ARG;
    for (uint32_t c = 0; c < cl->ncl; ++c)
        for (uint32_t j = cl->bptr[c]; j < cl->bptr[c + 1]; ++j)
            borders[j] = cl->gid[cl->voff[c] + cl->blocal[j]];
    return 0;
}

int64_t fbc_clusters_edges(const fbc_clusters *cl, uint32_t c)
{
    if (!cl || c >= cl->ncl) return FBC_E_ARG;
    return (int64_t)cl->edges[c];
}

void fbc_clusters_free(fbc_clusters *cl)
{
    if (!cl) return;
    dfree(cl->d_n);
    dfree(cl->d_gid);
    dfree(cl->d_bptr);
    dfree(cl->d_blocal);
    dfree(cl->d_vcl);
    dfree(cl->d_rp);
    dfree(cl->d_rrp);
    dfree(cl->d_arcs);
    dfree(cl->d_ell);
    dfree(cl->d_rarcs);
    dfree(cl->d_rpoff);
    dfree(cl->d_voff);
    dfree(cl->d_vioff);
    dfree(cl->d_doff);
    dfree(cl->d_vil);
    dfree(cl->d_vic);
    dfree(cl->d_cls_cluster);
    dfree(cl->d_cls_grow);
    dfree(cl->d_cls_ptr);
    dfree(cl->d_rmin);
    delete cl;
}

static int cluster_eval(fbc_ctx *ctx, fbc_clusters *cl, double *intra_bc, uint32_t *vi_len, int32_t *vi_cnt,
                        bool resident)
{
    const fbc_graph *g = cl->g;
    /* a new evaluation replaces the rows and classes of an earlier one */
    dfree(cl->d_vil);
    dfree(cl->d_vic);
    cl->ncls = 0;
    cl->cls_ptr.clear();
    cl->cls_size.clear();
    double *d_intra = nullptr, *d_dmat = nullptr;
    uint32_t *d_vil = nullptr;
    int32_t *d_vic = nullptr;
    uint32_t *d_tsrc = nullptr, *d_taux = nullptr;
    hipError_t e;
    int rc = 0;
    /* dependency matrices (n_c x n_c doubles): all bordered clusters at once
     * when they fit in a quarter of device memory (at most 64 GiB), else
     * chunks; never less than the largest one */
    const size_t dbudget = std::min<size_t>(ctx->total_mem / 4, (size_t)64 << 30) / sizeof(double);
    size_t dcap = 0, dall = 0;
    for (uint32_t c = 0; c < cl->ncl; ++c)
        if (cl->bptr[c + 1] > cl->bptr[c]) {
            dcap = std::max(dcap, (size_t)cl->n[c] * cl->n[c]);
            dall += (size_t)cl->n[c] * cl->n[c];
        }
    dcap = std::max(dcap, std::min(dall, dbudget));
    if ((e = dalloc(&d_intra, g->n)) != hipSuccess || (e = dalloc(&d_dmat, dcap)) != hipSuccess ||
        (e = dalloc(&d_vil, cl->vi_total)) != hipSuccess || (e = dalloc(&d_vic, cl->vi_total)) != hipSuccess ||
        (e = hipMemcpy(d_intra, intra_bc, (size_t)g->n * 8, hipMemcpyHostToDevice)) != hipSuccess) {
        rc = hip_fail(ctx, e, "cluster eval setup");
    }
    double ms = 0, te = 0;
    uint32_t c0 = 0;
    while (!rc && c0 < cl->ncl) {
        /* a chunk of consecutive clusters whose dependency matrices fit */
        std::vector<uint64_t> doff(cl->ncl, ~uint64_t(0));
        std::vector<uint32_t> order;
        size_t used = 0;
        uint32_t c1 = c0;
        while (c1 < cl->ncl) {
            const size_t sz = (size_t)cl->n[c1] * cl->n[c1];
            const bool bordered = cl->bptr[c1 + 1] > cl->bptr[c1];
            if (bordered && used + sz > dcap && !order.empty()) break;
            if (bordered) {
                doff[c1] = used;
                used += sz;
                order.push_back(c1);
            }
            ++c1;
        }
        /* largest clusters first */
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return (double)cl->n[a] * cl->edges[a] > (double)cl->n[b] * cl->edges[b];
        });
        std::vector<uint32_t> tsrc, taux;
        for (uint32_t c : order)
            for (uint32_t s = 0; s < cl->n[c]; ++s) {
                tsrc.push_back(s);
                taux.push_back(c);
            }
        if (!tsrc.empty()) {
            dfree(d_tsrc);
            dfree(d_taux);
            if ((e = upload(&d_tsrc, tsrc, ctx->stream)) != hipSuccess ||
                (e = upload(&d_taux, taux, ctx->stream)) != hipSuccess ||
                (e = hipMemcpy(cl->d_doff, doff.data(), doff.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) {
                rc = hip_fail(ctx, e, "cluster chunk setup");
                break;
            }
            BrandesParams p;
            base_params(p, g);
            p.arcs = cl->d_arcs;
            p.rarcs = cl->d_rarcs;
            p.rp = cl->d_rp;
            p.rrp = cl->d_rrp;
            p.ell = cl->d_ell;
            p.nmax = cl->nmax;
            p.cl_n = cl->d_n;
            p.cl_rpoff = cl->d_rpoff;
            p.cl_voff = cl->d_voff;
            p.cl_gid = cl->d_gid;
            p.cl_bptr = cl->d_bptr;
            p.cl_blocal = cl->d_blocal;
            p.cl_vioff = cl->d_vioff;
            p.cl_doff = cl->d_doff;
            p.ntasks = (uint32_t)tsrc.size();
            p.task_src = d_tsrc;
            p.task_aux = d_taux;
            p.dmat = d_dmat;
            p.vi_len = d_vil;
            p.vi_cnt = d_vic;
            /* the share of the chunk's work (n_c * m_c) in views whose short
             * rows fit in an LDS-view workgroup */
            {
                const bool half = fbcdev::team_kind(cl->nmax, ctx->half_max) == fbcdev::TEAM_HALF;
                uint32_t bytes = 0;
                if ((e = fbcdev::lds_view_capacity(half, &bytes)) != hipSuccess) {
                    rc = hip_fail(ctx, e, "LDS capacity");
                    break;
                }
                double fit = 0, all = 0;
                for (uint32_t c : order) {
                    const double wk = (double)cl->n[c] * (double)cl->edges[c];
                    all += wk;
                    if ((uint64_t)cl->n[c] * sizeof(uint4) <= bytes) fit += wk;
                }
                ctx->lds_fit_share = all > 0 ? fit / all : 0.0;
            }
            rc = run(ctx, g, fbcdev::MODE_CLUSTER, 0, true, p);
            ctx->lds_fit_share = 0;
            if (rc) break;
            ms += ctx->last_ms;
            te += ctx->last_te;
            /* ordered per-vertex sums over the chunk's vertex range (borderless
             * clusters carry an all-ones matrix offset and are skipped) */
            const uint64_t vb = cl->voff[c0];
            const uint64_t ve = cl->voff[c1 - 1] + cl->n[c1 - 1];
            e = fbcdev::launch_intra_reduce(p, vb, ve - vb, cl->d_vcl, d_intra, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "intra reduce");
        }
        c0 = c1;
    }
    if (!rc) {
        if ((e = hipMemcpy(intra_bc, d_intra, (size_t)g->n * 8, hipMemcpyDeviceToHost)) != hipSuccess ||
            (!resident && cl->vi_total &&
             (e = hipMemcpy(vi_len, d_vil, cl->vi_total * 4, hipMemcpyDeviceToHost)) != hipSuccess) ||
            (!resident && cl->vi_total &&
             (e = hipMemcpy(vi_cnt, d_vic, cl->vi_total * 4, hipMemcpyDeviceToHost)) != hipSuccess))
            rc = hip_fail(ctx, e, "result download");
    }
    if (!rc && resident) {
        cl->d_vil = d_vil;
        cl->d_vic = d_vic;
        d_vil = nullptr;
        d_vic = nullptr;
    }
    dfree(d_intra);
    dfree(d_dmat);
    dfree(d_vil);
    dfree(d_vic);
    dfree(d_tsrc);
    dfree(d_taux);
    ctx->last_ms = ms;
    ctx->last_te = te;
    return rc;
}

int fbc_cluster_eval(fbc_ctx *ctx, const fbc_clusters *cl, double *intra_bc, uint32_t *vi_len,
                     int32_t *vi_cnt)
{
    if (!ctx || !cl || !intra_bc || (cl->vi_total && (!vi_len || !vi_cnt)))
        return fail(ctx, FBC_E_ARG, "null argument");
    return cluster_eval(ctx, const_cast<fbc_clusters *>(cl), intra_bc, vi_len, vi_cnt, false);
}

int fbc_cluster_eval_resident(fbc_ctx *ctx, fbc_clusters *cl, double *intra_bc)
{
    if (!ctx || !cl || !intra_bc) return fail(ctx, FBC_E_ARG, "null argument");
    return cluster_eval(ctx, cl, intra_bc, nullptr, nullptr, true);
}

int fbc_cluster_rows(fbc_ctx *ctx, const fbc_clusters *cl, uint32_t *vi_len, int32_t *vi_cnt)
{
    if (!ctx || !cl || (cl->vi_total && (!vi_len || !vi_cnt))) return fail(ctx, FBC_E_ARG, "null argument");
    if (cl->vi_total && !cl->d_vil) return fail(ctx, FBC_E_ARG, "no resident VertexInfo rows");
    if (!cl->vi_total) return 0;
    hipError_t e = hipMemcpy(vi_len, cl->d_vil, cl->vi_total * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(vi_cnt, cl->d_vic, cl->vi_total * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "row download");
}

namespace {

fbcdev::RowSet row_set(const fbc_clusters *cl)
{
    fbcdev::RowSet rs;
    rs.nrows = cl->voff.empty() ? 0 : cl->voff.back() + cl->n.back();
    rs.ncl = cl->ncl;
    rs.vcl = cl->d_vcl;
    rs.voff = cl->d_voff;
    rs.bptr = cl->d_bptr;
    rs.blocal = cl->d_blocal;
    rs.vioff = cl->d_vioff;
    rs.vi_len = cl->d_vil;
    rs.vi_cnt = cl->d_vic;
    rs.gid = cl->d_gid;
    return rs;
}

} // namespace

int fbc_cluster_classes(fbc_ctx *ctx, fbc_clusters *cl, const double *intra_bc, uint32_t *cls_ptr,
                        uint32_t *piv_row, uint32_t *weight)
{
    if (!ctx || !cl || !intra_bc || !cls_ptr || !piv_row || !weight) return fail(ctx, FBC_E_ARG, "null argument");
    if (cl->vi_total && !cl->d_vil) return fail(ctx, FBC_E_ARG, "no resident VertexInfo rows");
    const fbcdev::RowSet rs = row_set(cl);
    const uint64_t n = rs.nrows;
    if (n >= 0x7FFFFFFFull) return fail(ctx, FBC_E_UNSUPPORTED, "too many cluster vertices for the class sort");
    dfree(cl->d_cls_cluster);
    dfree(cl->d_cls_grow);
    dfree(cl->d_cls_ptr);
    dfree(cl->d_rmin);
    cl->ncls = 0;
    cl->cls_ptr.clear();
    cl->cls_size.clear();
    uint64_t *d_key = nullptr, *d_skey = nullptr;
    uint32_t *d_val = nullptr, *d_sval = nullptr, *d_head = nullptr, *d_size = nullptr, *d_piv = nullptr,
             *d_pos = nullptr, *d_crow = nullptr, *d_csize = nullptr, *d_flag = nullptr;
    uint8_t *d_isb = nullptr;
    double *d_intra = nullptr;
    void *d_tmp = nullptr;
    size_t sort_bytes = 0, scan_bytes = 0;
    hipError_t e;
    int rc = 0;
    if ((e = fbcdev::classes_sort_bytes(n, &sort_bytes)) != hipSuccess ||
        (e = fbcdev::classes_scan_bytes(n, &scan_bytes)) != hipSuccess ||
        (e = dalloc(&d_key, n)) != hipSuccess || (e = dalloc(&d_skey, n)) != hipSuccess ||
        (e = dalloc(&d_val, n)) != hipSuccess || (e = dalloc(&d_sval, n)) != hipSuccess ||
        (e = dalloc(&cl->d_rmin, n)) != hipSuccess || (e = dalloc(&d_head, n)) != hipSuccess ||
        (e = dalloc(&d_size, n)) != hipSuccess || (e = dalloc(&d_piv, n)) != hipSuccess ||
        (e = dalloc(&d_pos, n)) != hipSuccess || (e = dalloc(&d_isb, n)) != hipSuccess ||
        (e = dalloc(&d_flag, 1)) != hipSuccess ||
        (e = dalloc((uint8_t **)&d_tmp, std::max(sort_bytes, scan_bytes))) != hipSuccess ||
        (e = upload_n(&d_intra, intra_bc, cl->g->n, ctx->stream)) != hipSuccess)
        rc = hip_fail(ctx, e, "class setup");
    uint32_t collision = 0;
    if (!rc) {
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        if ((e = fbcdev::classes_launch_hash(rs, ctx->class_key_mask, d_key, d_val, cl->d_rmin, ctx->stream)) !=
                hipSuccess ||
            (e = fbcdev::classes_launch_borders(rs, d_isb, ctx->stream)) != hipSuccess ||
            (e = fbcdev::classes_sort(d_tmp, sort_bytes, d_key, d_skey, d_val, d_sval, n, ctx->stream)) !=
                hipSuccess ||
            (e = fbcdev::classes_launch_runs(rs, d_skey, d_sval, cl->d_rmin, d_isb, d_intra, d_head, d_size, d_piv,
                                             d_flag, ctx->stream)) != hipSuccess ||
            (e = fbcdev::classes_scan(d_tmp, scan_bytes, d_head, d_pos, n, ctx->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(&collision, d_flag, 4, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
            rc = hip_fail(ctx, e, "class grouping");
    }
    if (!rc && collision) rc = fail(ctx, FBC_E_UNSUPPORTED, "hash collision in the class grouping (use the host selector)");
    uint32_t total = 0;
    if (!rc) {
        uint32_t last[2] = {0, 0};
        if (n && ((e = hipMemcpy(&last[0], d_pos + n - 1, 4, hipMemcpyDeviceToHost)) != hipSuccess ||
                  (e = hipMemcpy(&last[1], d_head + n - 1, 4, hipMemcpyDeviceToHost)) != hipSuccess))
            rc = hip_fail(ctx, e, "class count");
        total = last[0] + last[1];
    }
    if (!rc) {
        if ((e = dalloc(&d_crow, total)) != hipSuccess || (e = dalloc(&d_csize, total)) != hipSuccess ||
            (e = dalloc(&cl->d_cls_cluster, total)) != hipSuccess || (e = dalloc(&cl->d_cls_grow, total)) != hipSuccess ||
            (e = dalloc(&cl->d_cls_ptr, (size_t)cl->ncl + 1)) != hipSuccess ||
            (e = fbcdev::classes_launch_compact(rs, d_head, d_pos, d_size, d_piv, d_crow, d_csize, cl->d_cls_cluster,
                                                cl->d_cls_grow, cl->d_cls_ptr, ctx->stream)) != hipSuccess ||
            (e = hipEventRecord(ctx->ev1, ctx->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess ||
            (e = hipMemcpy(cls_ptr, cl->d_cls_ptr, ((size_t)cl->ncl + 1) * 4, hipMemcpyDeviceToHost)) != hipSuccess ||
            (total && (e = hipMemcpy(piv_row, d_crow, (size_t)total * 4, hipMemcpyDeviceToHost)) != hipSuccess) ||
            (total && (e = hipMemcpy(weight, d_csize, (size_t)total * 4, hipMemcpyDeviceToHost)) != hipSuccess))
            rc = hip_fail(ctx, e, "class list");
    }
    if (!rc) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        ctx->last_ms = ms;
        ctx->last_te = 0;
        cl->ncls = total;
        cl->cls_ptr.assign(cls_ptr, cls_ptr + cl->ncl + 1);
        cl->cls_size.assign(weight, weight + total);
    } else {
        dfree(cl->d_cls_cluster);
        dfree(cl->d_cls_grow);
        dfree(cl->d_cls_ptr);
        dfree(cl->d_rmin);
    }
    for (auto *q : {d_key, d_skey}) dfree(q);
    for (auto *q : {d_val, d_sval, d_head, d_size, d_piv, d_pos, d_crow, d_csize, d_flag}) dfree(q);
    dfree(d_isb);
    dfree(d_intra);
    if (d_tmp) (void)hipFree(d_tmp);
    return rc ? rc : (int)0;
}

int fbc_cluster_class_features(fbc_ctx *ctx, const fbc_clusters *cl, uint32_t c, int32_t *features)
{
    if (!ctx || !cl || !features) return fail(ctx, FBC_E_ARG, "null argument");
    if (c >= cl->ncl || cl->cls_ptr.empty()) return fail(ctx, FBC_E_ARG, "no classes for this cluster");
    const uint64_t P = cl->cls_ptr[c + 1] - cl->cls_ptr[c], D = 2ull * (cl->bptr[c + 1] - cl->bptr[c]);
    if (!P || !D) return 0;
    std::vector<uint64_t> xoff(cl->ncl, ~uint64_t(0));
    xoff[c] = 0;
    uint64_t *d_xoff = nullptr;
    int32_t *d_X = nullptr;
    hipError_t e;
    int rc = 0;
    if ((e = upload(&d_xoff, xoff, ctx->stream)) != hipSuccess || (e = dalloc(&d_X, P * D)) != hipSuccess ||
        (e = fbcdev::classes_launch_gather(row_set(cl), cl->ncls, cl->d_cls_cluster, cl->d_cls_grow, cl->d_rmin,
                                           cl->d_cls_ptr, d_xoff, d_X, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess ||
        (e = hipMemcpy(features, d_X, P * D * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        rc = hip_fail(ctx, e, "class features");
    dfree(d_xoff);
    dfree(d_X);
    return rc;
}

static int kmeans_device(fbc_ctx *ctx, uint32_t nsets, const uint32_t *npoints, const uint32_t *nfeat,
                         const uint32_t *k, const int32_t *dX, const std::vector<uint64_t> &xoff,
                         const int32_t *dW, double min_variance, uint32_t max_iteration, uint32_t *centroids,
                         int32_t *cweights);

int fbc_cluster_kmeans(fbc_ctx *ctx, const fbc_clusters *cl, double kfrac, double min_variance,
                       uint32_t max_iteration, uint32_t *kptr, uint32_t *centroids, int32_t *cweights)
{
    if (!ctx || !cl || !kptr || !centroids || !cweights) return fail(ctx, FBC_E_ARG, "null argument");
    if (cl->cls_ptr.empty()) return fail(ctx, FBC_E_ARG, "no classes: call fbc_cluster_classes first");
    const uint32_t ncl = cl->ncl;
    /* the largest |feature| of every cluster decides where its k-means runs */
    unsigned long long *d_fmax = nullptr;
    std::vector<unsigned long long> fmax(ncl ? ncl : 1);
    hipError_t e;
    int rc = 0;
    if ((e = dalloc(&d_fmax, ncl)) != hipSuccess ||
        (e = fbcdev::classes_launch_feature_max(row_set(cl), cl->ncls, cl->d_cls_cluster, cl->d_cls_grow,
                                                cl->d_rmin, d_fmax, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess ||
        (ncl && (e = hipMemcpy(fmax.data(), d_fmax, ncl * sizeof(unsigned long long), hipMemcpyDeviceToHost)) !=
                    hipSuccess))
        rc = hip_fail(ctx, e, "k-means feature range");
    dfree(d_fmax);
    if (rc) return rc;
    std::vector<uint32_t> sets, P, D, K;
    std::vector<uint64_t> xoff, xoff_cl(ncl, ~uint64_t(0));
    std::vector<uint32_t> setk(ncl, 0);
    uint64_t xs = 0;
    for (uint32_t c = 0; c < ncl; ++c) {
        const uint32_t pc = cl->cls_ptr[c + 1] - cl->cls_ptr[c], nb = cl->bptr[c + 1] - cl->bptr[c];
        if (!pc || !nb) continue;
        const uint32_t kc = (uint32_t)std::max((int)(pc * kfrac), 1);
        const double mx = (double)fmax[c];
        if (2.0 * 2 * nb * mx * mx >= 9007199254740992.0 || kc > FBC_KMEANS_MAX_K || kc > pc) continue; /* host */
        sets.push_back(c);
        P.push_back(pc);
        D.push_back(2 * nb);
        K.push_back(kc);
        xoff.push_back(xs);
        xoff_cl[c] = xs;
        setk[c] = kc;
        xs += (uint64_t)pc * 2 * nb;
    }
    kptr[0] = 0;
    for (uint32_t c = 0; c < ncl; ++c) kptr[c + 1] = kptr[c] + setk[c];
    if (sets.empty()) return 0;
    /* weights: class sizes, in set order */
    std::vector<int32_t> W;
    uint64_t *d_xoff = nullptr;
    int32_t *d_X = nullptr, *d_W = nullptr;
    std::vector<uint32_t> oc(kptr[ncl]);
    std::vector<int32_t> ow(kptr[ncl]);
    rc = 0;
    if ((e = upload(&d_xoff, xoff_cl, ctx->stream)) != hipSuccess || (e = dalloc(&d_X, xs)) != hipSuccess ||
        (e = fbcdev::classes_launch_gather(row_set(cl), cl->ncls, cl->d_cls_cluster, cl->d_cls_grow, cl->d_rmin,
                                           cl->d_cls_ptr, d_xoff, d_X, ctx->stream)) != hipSuccess)
        rc = hip_fail(ctx, e, "k-means points");
    if (!rc) {
        for (uint32_t c : sets)
            W.insert(W.end(), cl->cls_size.begin() + cl->cls_ptr[c], cl->cls_size.begin() + cl->cls_ptr[c + 1]);
        if ((e = upload(&d_W, W, ctx->stream)) != hipSuccess) rc = hip_fail(ctx, e, "k-means weights");
    }
    if (!rc)
        rc = kmeans_device(ctx, (uint32_t)sets.size(), P.data(), D.data(), K.data(), d_X, xoff, d_W, min_variance,
                           max_iteration, oc.data(), ow.data());
    if (!rc) {
        size_t o = 0;
        for (size_t i = 0; i < sets.size(); ++i) {
            std::copy(oc.begin() + o, oc.begin() + o + K[i], centroids + kptr[sets[i]]);
            std::copy(ow.begin() + o, ow.begin() + o + K[i], cweights + kptr[sets[i]]);
            o += K[i];
        }
    }
    dfree(d_xoff);
    dfree(d_X);
    dfree(d_W);
    return rc;
}

static int pivot_phase(fbc_ctx *ctx, const fbc_clusters *cl, const uint32_t *pivot, const uint32_t *weight,
                       const uint32_t *pivot_cluster, uint32_t npiv, const double *intra_bc, int prec,
                       double *bc, bool plain)
{
    if (!ctx || !cl || !bc || !intra_bc || (npiv && (!pivot || !weight || !pivot_cluster)))
        return fail(ctx, FBC_E_ARG, "null argument");
    if (!valid_prec(prec)) return fail(ctx, FBC_E_ARG, "unknown precision");
    const fbc_graph *g = cl->g;
    for (uint32_t i = 0; i < npiv; ++i)
        if (pivot[i] >= g->n || pivot_cluster[i] >= cl->ncl) return fail(ctx, FBC_E_ARG, "pivot out of range");
    /* pivot weights grouped by cluster, in the given order */
    std::vector<uint32_t> cptr((size_t)cl->ncl + 1, 0), pw(npiv);
    for (uint32_t i = 0; i < npiv; ++i) cptr[pivot_cluster[i] + 1]++;
    for (uint32_t c = 0; c < cl->ncl; ++c) cptr[c + 1] += cptr[c];
    {
        std::vector<uint32_t> fill(cptr.begin(), cptr.end() - 1);
        for (uint32_t i = 0; i < npiv; ++i) pw[fill[pivot_cluster[i]]++] = weight[i];
    }
    /* the batched form runs on the relabelled graph and gathers its sums back */
    const bool blocked = !plain && ctx->relabel != 0;
    if (blocked) {
        const int rc = ensure_blocked(ctx, const_cast<fbc_graph *>(g));
        if (rc) return rc;
    }
    std::vector<uint32_t> tsrc(pivot, pivot + npiv), taux(weight, weight + npiv);
    if (blocked)
        for (uint32_t &x : tsrc) x = g->perm[x];
    double *d_bc = nullptr, *d_intra = nullptr, *d_bcb = nullptr;
    uint32_t *d_tsrc = nullptr, *d_taux = nullptr, *d_cptr = nullptr, *d_pw = nullptr;
    hipError_t e;
    int rc = 0;
    if ((e = dalloc(&d_bc, g->n)) != hipSuccess || (e = dalloc(&d_intra, g->n)) != hipSuccess ||
        (e = upload(&d_tsrc, tsrc, ctx->stream)) != hipSuccess || (e = upload(&d_taux, taux, ctx->stream)) != hipSuccess ||
        (e = upload(&d_cptr, cptr, ctx->stream)) != hipSuccess || (e = upload(&d_pw, pw, ctx->stream)) != hipSuccess ||
        (e = hipMemcpy(d_bc, bc, (size_t)g->n * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_intra, intra_bc, (size_t)g->n * 8, hipMemcpyHostToDevice)) != hipSuccess)
        rc = hip_fail(ctx, e, "pivot setup");
    if (!rc && blocked &&
        ((e = dalloc(&d_bcb, g->n)) != hipSuccess ||
         (e = hipMemsetAsync(d_bcb, 0, (size_t)g->n * 8, ctx->stream)) != hipSuccess))
        rc = hip_fail(ctx, e, "pivot setup");
    if (!rc) {
        BrandesParams p;
        base_params(p, g);
        if (blocked) use_blocked(p, g);
        p.ntasks = npiv;
        p.task_src = d_tsrc;
        p.task_aux = d_taux;
        p.bc = blocked ? d_bcb : d_bc;
        rc = run(ctx, g, fbcdev::MODE_PIVOT, prec, false, p, plain);
    }
    if (!rc && blocked && (e = fbcdev::launch_gather_add(g->n, g->d_perm, d_bcb, d_bc, ctx->stream)) != hipSuccess)
        rc = hip_fail(ctx, e, "pivot gather");
    if (!rc) {
        e = fbcdev::launch_pivot_subtract(cl->gid.size(), cl->d_gid, cl->d_vcl, d_cptr, d_pw, d_intra, d_bc,
                                          ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipMemcpy(bc, d_bc, (size_t)g->n * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "pivot finish");
    }
    dfree(d_bc);
    dfree(d_bcb);
    dfree(d_intra);
    dfree(d_tsrc);
    dfree(d_taux);
    dfree(d_cptr);
    dfree(d_pw);
    return rc;
}

int fbc_pivot_phase(fbc_ctx *ctx, const fbc_clusters *cl, const uint32_t *pivot, const uint32_t *weight,
                    const uint32_t *pivot_cluster, uint32_t npiv, const double *intra_bc, int prec,
                    double *bc)
{
    return pivot_phase(ctx, cl, pivot, weight, pivot_cluster, npiv, intra_bc, prec, bc, false);
}

int fbc_pivot_phase_plain(fbc_ctx *ctx, const fbc_clusters *cl, const uint32_t *pivot, const uint32_t *weight,
                          const uint32_t *pivot_cluster, uint32_t npiv, const double *intra_bc, double *bc)
{
    return pivot_phase(ctx, cl, pivot, weight, pivot_cluster, npiv, intra_bc, FBC_PREC_F64, bc, true);
}

/* k-means of sets whose features (dX, set s at xoff[s]) and weights (dW,
 * concatenated in set order) are already in device memory; sizes checked by
 * the caller. */
static int kmeans_device(fbc_ctx *ctx, uint32_t nsets, const uint32_t *npoints, const uint32_t *nfeat,
                         const uint32_t *k, const int32_t *dX, const std::vector<uint64_t> &xoff,
                         const int32_t *dW, double min_variance, uint32_t max_iteration, uint32_t *centroids,
                         int32_t *cweights)
{
    std::vector<uint64_t> poff(nsets), koff(nsets), moff(nsets), goff(nsets, 0);
    uint64_t ps = 0, ks = 0, ms = 0;
    for (uint32_t s = 0; s < nsets; ++s) {
        poff[s] = ps;
        koff[s] = ks;
        moff[s] = ms;
        ps += npoints[s];
        ks += k[s];
        ms += (uint64_t)k[s] * nfeat[s];
    }
    /* groups of sets whose Gram matrices fit the budget together, largest
     * sets first: a group's initialisation lasts as long as its largest set's
     * sequential choice of centroids, so similar sizes share a group */
    uint64_t budget = std::min<uint64_t>(ctx->total_mem / 4, (uint64_t)64 << 30);
    if (ctx->km_budget_mb) {
        budget = (uint64_t)ctx->km_budget_mb << 20;
    } else {
        /* and what is free now, less a sixteenth of the device */
        size_t freeb = 0, totalb = 0;
        if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
            const uint64_t reserve = ctx->total_mem / 16;
            budget = std::min<uint64_t>(budget, freeb > 2 * reserve ? freeb - reserve : freeb / 2);
        }
        budget = std::max<uint64_t>(budget, 1);
    }
    /* the slot memory of the Brandes launches is idle between them (every
     * task initialises its own slot), so the Gram matrices use it when it is
     * the larger: after a first computeBC it holds the pivot phase's slots
     * (86 GB on the 1M-vertex grid, 129 GB on the 4M-vertex one, whose Gram
     * matrices then need fewer groups) */
    bool in_scratch = ctx->scratch && ctx->scratch_bytes > budget;
    if (in_scratch) budget = ctx->scratch_bytes;
    std::vector<std::vector<uint32_t>> groups;
    uint64_t gmax = 0;
    {
        std::vector<uint32_t> by_size(nsets);
        for (uint32_t s = 0; s < nsets; ++s) by_size[s] = s;
        std::stable_sort(by_size.begin(), by_size.end(),
                         [&](uint32_t a, uint32_t b) { return npoints[a] > npoints[b]; });
        uint64_t used = 0;
        for (uint32_t s : by_size) {
            const uint64_t g = (uint64_t)npoints[s] * npoints[s];
            if (groups.empty() || (used + g) * 8 > budget) {
                groups.emplace_back();
                used = 0;
            }
            goff[s] = used;
            used += g;
            groups.back().push_back(s);
            gmax = std::max(gmax, used);
        }
    }
    /* a set larger than the budget has a group of its own: then the largest
     * group may not fit the slot memory, and gets its own allocation */
    if (in_scratch && gmax * sizeof(double) > ctx->scratch_bytes) in_scratch = false;
    std::vector<uint32_t> P(npoints, npoints + nsets), D(nfeat, nfeat + nsets), K(k, k + nsets), active(nsets, 0);
    fbcdev::KmParams p;
    std::memset(&p, 0, sizeof(p));
    p.ncl = nsets;
    p.min_variance = min_variance;
    uint32_t *dP = nullptr, *dD = nullptr, *dK = nullptr, *dtiles = nullptr;
    uint64_t *dxo = nullptr, *dpo = nullptr, *dko = nullptr, *dmo = nullptr, *dgo = nullptr;
    double *dxn = nullptr, *dcd = nullptr, *dmd = nullptr, *dmsum = nullptr, *dG = nullptr,
           *dkv = nullptr;
    int32_t *dcw = nullptr;
    uint32_t *dcent = nullptr, *dnext = nullptr, *dcnt = nullptr, *dstart = nullptr, *dasg = nullptr,
             *dmem = nullptr, *dact = nullptr, *duk = nullptr, *dnu = nullptr, *dfirst = nullptr;
    hipError_t e;
    int rc = 0;
    double ms_total = 0;
    if ((e = upload(&dP, P, ctx->stream)) != hipSuccess || (e = upload(&dD, D, ctx->stream)) != hipSuccess ||
        (e = upload(&dK, K, ctx->stream)) != hipSuccess || (e = upload(&dxo, xoff, ctx->stream)) != hipSuccess ||
        (e = upload(&dpo, poff, ctx->stream)) != hipSuccess || (e = upload(&dko, koff, ctx->stream)) != hipSuccess ||
        (e = upload(&dmo, moff, ctx->stream)) != hipSuccess || (e = upload(&dgo, goff, ctx->stream)) != hipSuccess ||
        (e = dalloc(&dxn, ps)) != hipSuccess || (e = upload(&dact, active, ctx->stream)) != hipSuccess ||
        (e = dalloc(&dcd, ps)) != hipSuccess || (e = dalloc(&dmd, ps)) != hipSuccess ||
        (e = dalloc(&dkv, ks)) != hipSuccess ||
        (e = dalloc(&dmsum, std::max<uint64_t>(ms, 1))) != hipSuccess || (e = dalloc(&dcw, ks)) != hipSuccess ||
        (e = dalloc(&dcent, ks)) != hipSuccess || (e = dalloc(&dnext, ks)) != hipSuccess ||
        (e = dalloc(&dcnt, ks)) != hipSuccess || (e = dalloc(&dstart, ks)) != hipSuccess ||
        (e = dalloc(&dasg, ps)) != hipSuccess || (e = dalloc(&dmem, ps)) != hipSuccess ||
        (e = dalloc(&duk, ks)) != hipSuccess || (e = dalloc(&dnu, nsets)) != hipSuccess ||
        (e = dalloc(&dfirst, ps)) != hipSuccess ||
        (e = hipMemsetAsync(dfirst, 0xFF, ps * sizeof(uint32_t), ctx->stream)) != hipSuccess ||
        (!in_scratch && (e = dalloc(&dG, gmax)) != hipSuccess))
        rc = hip_fail(ctx, e, "k-means setup");
    if (!rc && in_scratch) dG = reinterpret_cast<double *>(ctx->scratch);
    if (!rc) {
        p.P = dP;
        p.D = dD;
        p.K = dK;
        p.xoff = dxo;
        p.poff = dpo;
        p.koff = dko;
        p.moff = dmo;
        p.goff = dgo;
        p.X = dX;
        p.xnorm = dxn;
        p.G = dG;
        p.weights = dW;
        p.cdist = dcd;
        p.mdist = dmd;
        p.kvar = dkv;
        p.cent = dcent;
        p.next = dnext;
        p.cnt = dcnt;
        p.start = dstart;
        p.assign = dasg;
        p.members = dmem;
        p.msum = dmsum;
        p.cweights = dcw;
        p.active = dact;
        p.uk = duk;
        p.nu = dnu;
        p.firstk = dfirst;
        (void)hipEventRecord(ctx->ev0, ctx->stream);
        for (const std::vector<uint32_t> &grp : groups) {
            /* Gram tiles, then the iteration tiles, in one device array */
            std::vector<uint32_t> tiles, pc, pf, sum, mc, mf, kc, kf;
            std::fill(active.begin(), active.end(), 0u);
            for (uint32_t s : grp) {
                active[s] = 1;
                const uint32_t nt = (P[s] + 63) / 64;
                for (uint32_t ti = 0; ti < nt; ++ti)
                    for (uint32_t tj = ti; tj < nt; ++tj) {
                        tiles.push_back(s);
                        tiles.push_back(ti);
                        tiles.push_back(tj);
                    }
                for (uint32_t p0 = 0; p0 < P[s]; p0 += 256) {
                    pc.push_back(s);
                    pf.push_back(p0);
                }
                for (uint32_t m0 = 0; m0 < P[s]; m0 += 64) {
                    mc.push_back(s);
                    mf.push_back(m0);
                    for (uint32_t f0 = 0; f0 < D[s]; f0 += 256) {
                        sum.push_back(s);
                        sum.push_back(m0);
                        sum.push_back(f0);
                    }
                }
                for (uint32_t k0 = 0; k0 < K[s]; k0 += 256) {
                    kc.push_back(s);
                    kf.push_back(k0);
                }
            }
            std::vector<uint32_t> all;
            std::vector<size_t> at;
            for (const std::vector<uint32_t> *v : {&tiles, &pc, &pf, &sum, &mc, &mf, &kc, &kf}) {
                at.push_back(all.size());
                all.insert(all.end(), v->begin(), v->end());
            }
            dfree(dtiles);
            if ((e = upload(&dtiles, all, ctx->stream)) != hipSuccess ||
                (e = hipMemcpyAsync(dact, active.data(), nsets * sizeof(uint32_t), hipMemcpyHostToDevice,
                                    ctx->stream)) != hipSuccess ||
                (e = fbcdev::km_launch_gram(p, dG, (uint32_t)(tiles.size() / 3), dtiles + at[0], ctx->stream)) != hipSuccess ||
                (e = fbcdev::km_launch_init(p, ctx->stream)) != hipSuccess)
                break;
            fbcdev::KmTiles kt;
            kt.npt = (uint32_t)pc.size();
            kt.nsum = (uint32_t)(sum.size() / 3);
            kt.nmt = (uint32_t)mc.size();
            kt.nkt = (uint32_t)kc.size();
            kt.pt_cl = dtiles + at[1];
            kt.pt_first = dtiles + at[2];
            kt.sum = dtiles + at[3];
            kt.mt_cl = dtiles + at[4];
            kt.mt_first = dtiles + at[5];
            kt.kt_cl = dtiles + at[6];
            kt.kt_first = dtiles + at[7];
            for (uint32_t it = 1; it <= max_iteration + 1; ++it) {
                e = fbcdev::km_launch_iteration(p, kt, it, max_iteration, ctx->stream);
                if (e != hipSuccess) break;
                e = hipMemcpyAsync(active.data(), dact, nsets * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) break;
                bool any = false;
                for (uint32_t a : active) any = any || a;
                if (!any) break;
            }
            if (e != hipSuccess) break;
        }
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        float ms = 0;
        if (e == hipSuccess) (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        ms_total = ms;
        if (e == hipSuccess) e = hipMemcpy(centroids, dcent, ks * sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(cweights, dcw, ks * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "k-means run");
    }
    for (auto *q : {dP, dD, dK, dtiles, dcent, dnext, dcnt, dstart, dasg, dmem, dact, duk, dnu, dfirst}) dfree(q);
    for (auto *q : {dxo, dpo, dko, dmo, dgo}) dfree(q);
    for (auto *q : {dxn, dcd, dmd, dmsum, dkv}) dfree(q);
    if (!in_scratch) dfree(dG);
    dfree(dcw);
    ctx->last_ms = ms_total;
    ctx->last_te = 0;
    return rc;
}

int fbc_kmeans_batch(fbc_ctx *ctx, uint32_t nsets, const uint32_t *npoints, const uint32_t *nfeat,
                     const uint32_t *k, const int32_t *features, const int32_t *weights, double min_variance,
                     uint32_t max_iteration, uint32_t *centroids, int32_t *cweights)
{
    if (!ctx || (nsets && (!npoints || !nfeat || !k || !features || !weights || !centroids || !cweights)))
        return fail(ctx, FBC_E_ARG, "null argument");
    if (nsets == 0) return 0;
    std::vector<uint64_t> xoff(nsets);
    uint64_t xs = 0, ps = 0;
    for (uint32_t s = 0; s < nsets; ++s) {
        if (npoints[s] == 0 || k[s] == 0 || k[s] > npoints[s]) return fail(ctx, FBC_E_ARG, "bad k-means set sizes");
        if (k[s] > FBC_KMEANS_MAX_K) return fail(ctx, FBC_E_UNSUPPORTED, "k-means set with too many centroids");
        xoff[s] = xs;
        xs += (uint64_t)npoints[s] * nfeat[s];
        ps += npoints[s];
    }
    /* exactness precondition (the squared norms come from the Gram diagonal) */
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (uint32_t s = 0; s < nsets; ++s) {
        double mx = 0;
        const int32_t *X = features + xoff[s];
        for (uint64_t i = 0; i < (uint64_t)npoints[s] * nfeat[s]; ++i) mx = std::max(mx, std::fabs((double)X[i]));
        if (2.0 * nfeat[s] * mx * mx >= 9007199254740992.0) bad |= 2;
    }
    if (bad & 2) return fail(ctx, FBC_E_UNSUPPORTED, "k-means features too large for exact FP64 distances");
    int32_t *dX = nullptr, *dW = nullptr;
    hipError_t e;
    int rc = 0;
    if ((e = upload_n(&dX, features, xs, ctx->stream)) != hipSuccess ||
        (e = upload_n(&dW, weights, ps, ctx->stream)) != hipSuccess)
        rc = hip_fail(ctx, e, "k-means upload");
    if (!rc)
        rc = kmeans_device(ctx, nsets, npoints, nfeat, k, dX, xoff, dW, min_variance, max_iteration, centroids,
                           cweights);
    dfree(dX);
    dfree(dW);
    return rc;
}

int fbc_last_stats(const fbc_ctx *ctx, double *kernel_ms, double *traversed_edges)
{
    if (!ctx) return FBC_E_ARG;
    if (kernel_ms) *kernel_ms = ctx->last_ms;
    if (traversed_edges) *traversed_edges = ctx->last_te;
    return 0;
}

int fbc_last_counters(const fbc_ctx *ctx, uint64_t *counters, int count)
{
    if (!ctx || !counters || count < 0) return FBC_E_ARG;
    const uint64_t host[] = {ctx->plain_sources, ctx->plain_fallback_sources, ctx->plain_refusal,
                             ctx->last_delta,    ctx->last_half_max,           ctx->last_tuned,
                             ctx->last_sigma_inexact, ctx->last_lds};
    constexpr int NH = (int)(sizeof(host) / sizeof(host[0]));
    int i = 0;
    for (; i < count && i < fbcdev::STAT_COUNT; ++i) counters[i] = ctx->last_counters[i];
    for (; i < count && i < fbcdev::STAT_COUNT + NH; ++i) counters[i] = host[i - fbcdev::STAT_COUNT];
    return i;
}

} // extern "C"
