/*
 * Persistent single-source Brandes kernels for gfx950.
 *
 * A team -- one workgroup, one wavefront or one half wavefront, by the size
 * of the view -- owns one source at a time and pulls the next (view, source)
 * task from a device work queue, so every source of a run -- all vertices
 * (--exact), all cluster vertices, or all pivots -- goes through one launch.
 * Per source:
 *
 *   1. distances: near-far delta-stepping.  Each vertex has one 64-bit key,
 *      tentative distance (high 32 bits) and a tag (low 32): the near round
 *      it was last queued for, or the window whose far list holds it.  A
 *      relaxation is one 64-bit atomicMin on the key; the key it returns
 *      tells whether the vertex is queued for the next round (or on this
 *      window's far list) already, so each is appended at most once.
 *   2. sigma: a sweep in vertex order counts each vertex's shortest-path DAG
 *      in-degree over the reverse CSR and records its tight-arc mask; a Kahn
 *      sweep then pushes sigma along the masked arcs with exact 64-bit
 *      integer atomics (the in-degree packed above a 48-bit sigma, or, after
 *      an overflow, a wide form with an overflow check), appending every
 *      vertex to the level list when its last DAG predecessor is done.  The
 *      level list is a topological order of the DAG.
 *   3. dependencies: the levels in reverse; every vertex pulls
 *      sigma(v) / sigma(w) * (1 + delta(w)) from its masked DAG successors.
 *      In 'ordered' mode the successors are summed in descending (d(w), w),
 *      the order in which the reference's backward stack visits them
 *      (DijkstraSSBrandesBC.h:61-78), so the FP64 result equals the
 *      reference's bit for bit.  The mode's output follows in one pass in
 *      vertex order.
 *
 * Views are CSRs in HBM (the whole graph, or one cluster's local CSR whose
 * local ids follow ascending global ids), with 16-byte short rows beside them
 * for the distance phase.  Per-slot state (keys, queues, sigma/mask/delta
 * pairs) lives in a private scratch region of the slot: 40 bytes per vertex.
 */
#include "fbc_internal.h"

#include <hip/hip_cooperative_groups.h>
#include <type_traits>

namespace fbcdev {

constexpr int HEAVY_DEG = 96;   /* out-degree handled by the whole workgroup */
/* Between the in-degree sweep and v's dependency, the low 32 bits of
 * sd[2v+1] hold v's tight-arc mask: bit j says whether out-arc j of v lies on
 * a shortest path, for out-degrees up to MASK_ARCS; MASK_NONE marks wider
 * rows, whose arcs are tested against the distances.
 * v's count of unprocessed DAG predecessors sits beside sigma(v) in one of
 * two forms.  Packed: sd[2v] = count << 48 | sigma, so a Kahn step adds
 * sigma(u) and takes one predecessor off with a single 64-bit atomic, and a
 * count reaching zero leaves sd[2v] = sigma(v).  Wide: sd[2v] = sigma alone,
 * the count in the high half of sd[2v+1] (two atomics per step).  A source
 * runs packed; a sigma of 2^48 or more or an in-degree of 2^16 or more, which
 * the packed form cannot hold, sends the source's sigma sweep round again in
 * the wide form. */
constexpr uint32_t MASK_ARCS = 31;
constexpr uint32_t MASK_NONE = 0xFFFFFFFFu;
constexpr uint64_t PRED_ONE = 1ull << 32;      /* wide: one predecessor in sd[2v+1] */
constexpr uint64_t PK_ONE = 1ull << 48;        /* packed: one predecessor in sd[2v] */
constexpr uint64_t PK_SIGMA = PK_ONE - 1;      /* packed: the sigma bits */
constexpr uint32_t PK_CMAX = 0xFFFFu;          /* packed: largest count */
constexpr int HEAVY_CAP = 64;
/* Per team: RB arcs of one vertex have their loads and atomics issued
 * together; the ordered fold keeps up to SUCC successors in registers; the
 * masked Kahn step and fold take MB successors per block (a masked row has
 * one or two as a rule, and a block's unused slots still load, to keep its
 * loads free of branches: 2 for workgroup teams, where wider blocks cost the
 * pivot launch 1.3%, 4 for wavefront teams).  The wavefront team takes
 * smaller blocks to stay at 4 wavefronts per SIMD. */

/* ---- memory helpers ------------------------------------------------------ */

/* Atomics and atomic loads take the scope of the team that owns the state:
 * per-slot state of the batched kernel is touched by one workgroup only, so
 * workgroup scope keeps it in the XCD's L2; the plain kernel spreads one
 * source over the whole grid and needs device (agent) scope. */
template <int SC> __device__ __forceinline__ uint32_t ald32(const uint32_t *p)
{
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint64_t ald64(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ void ast32(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ void ast64(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint32_t aadd32(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint32_t amin32(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint32_t amax32(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint64_t aadd64(uint64_t *p, uint64_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint64_t asub64(uint64_t *p, uint64_t v)
{
    return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint64_t amin64(uint64_t *p, uint64_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, SC);
}
template <int SC> __device__ __forceinline__ uint64_t amax64(uint64_t *p, uint64_t v)
{
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, SC);
}

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

/* Per-source state.  key[v] packs the tentative distance (high 32 bits) with
 * a tag (low 32 bits): NEAR_TAG | r when v was last queued for near round r,
 * or w when it was last put on the far list of window w.  sd[2v] is
 * sigma(v) (with the predecessor count in the packed form); sd[2v+1] holds
 * v's tight-arc mask (and the count in the wide form) from the in-degree
 * sweep until the dependency phase replaces it by delta(v) (a double, or a
 * float pair, bit-cast), so a DAG successor's (sigma, delta) share one
 * 16-byte line segment. */
struct Slot {
    uint64_t *key;
    uint32_t *qa, *qb, *fa, *fb;
    uint64_t *sd;
    bool agg;   /* list appends aggregated per wavefront (team_append) */
};
constexpr uint32_t NEAR_TAG = 0x80000000u;

__host__ __device__ inline uint64_t slot_layout(uint32_t nmax, uint8_t *base, Slot *s)
{
    const uint64_t w32 = align256((uint64_t)(nmax + 1) * 4);
    const uint64_t w64 = align256((uint64_t)(nmax + 1) * 8);
    if (s) {
        s->key = (uint64_t *)(base);
        s->qa = (uint32_t *)(base + w64);
        s->qb = (uint32_t *)(base + w64 + w32);
        s->fa = (uint32_t *)(base + w64 + 2 * w32);
        s->fb = (uint32_t *)(base + w64 + 3 * w32);
        s->sd = (uint64_t *)(base + w64 + 4 * w32);
    }
    return 3 * w64 + 4 * w32;
}

/* distance half of key[v] */
template <int SC> __device__ __forceinline__ uint32_t dist_of(const Slot S, uint32_t v)
{
    return ald32<SC>(reinterpret_cast<const uint32_t *>(S.key) + 2 * (uint64_t)v + 1);
}
/* the same once the key is no longer written (dependency phase) */
__device__ __forceinline__ uint32_t dist_final(const Slot S, uint32_t v)
{
    return reinterpret_cast<const uint32_t *>(S.key)[2 * (uint64_t)v + 1];
}
template <int SC> __device__ __forceinline__ uint64_t sigma_of(const Slot S, uint32_t v)
{
    return ald64<SC>(&S.sd[2 * (uint64_t)v]);
}
/* (sigma, delta) of a DAG successor in one 16-byte load; valid once the
 * sigma sweep and the successor's level are complete */
__device__ __forceinline__ ulonglong2 sd_pair(const Slot S, uint32_t v)
{
    return *reinterpret_cast<const ulonglong2 *>(S.sd + 2 * (uint64_t)v);
}

uint64_t slot_bytes(uint32_t nmax) { return slot_layout(nmax, nullptr, nullptr); }

/* ---- FP32 pair arithmetic (two-sum / two-product, value = hi + lo) ------- */
struct F2 { float hi, lo; };

__device__ __forceinline__ F2 quick_two_sum(float a, float b)
{
    float s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ F2 two_sum(float a, float b)
{
    float s = a + b;
    float bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ F2 f2_add(F2 x, F2 y)
{
    F2 s = two_sum(x.hi, y.hi);
    F2 t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ F2 f2_mul(F2 x, F2 y)
{
    float p = x.hi * y.hi;
    float e = __builtin_fmaf(x.hi, y.hi, -p);
    e += x.hi * y.lo + x.lo * y.hi;
    return quick_two_sum(p, e);
}
__device__ __forceinline__ F2 f2_div(F2 x, F2 y)
{
    float q1 = x.hi / y.hi;
    F2 r = f2_add(x, f2_mul({-q1, 0.0f}, y));
    float q2 = r.hi / y.hi;
    return quick_two_sum(q1, q2);
}
/* v = x1 * 2^24 + x0 with x0, x1 < 2^24: both are exact in FP32 and their
 * two-sum is v exactly, for every v below 2^48 (the packed sigma form's
 * range); a larger v (the wide form) rounds its high part to 24 bits, the
 * pair's own precision. */
__device__ __forceinline__ F2 f2_from_u64(uint64_t v)
{
    const float x0 = (float)(uint32_t)(v & 0xFFFFFFu);
    const uint64_t h = v >> 24;
    const float x1 = (h >> 32 ? (float)(uint32_t)(h >> 16) * 65536.0f : (float)(uint32_t)h) * 16777216.0f;
    return two_sum(x1, x0);
}
__device__ __forceinline__ uint64_t f2_pack(F2 x)
{
    return ((uint64_t)__float_as_uint(x.lo) << 32) | __float_as_uint(x.hi);
}
__device__ __forceinline__ F2 f2_unpack(uint64_t b)
{
    return {__uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32))};
}

/* ---- one task ------------------------------------------------------------ */

struct View {
    const uint2 *arcs;
    const uint32_t *rp;
    const uint2 *rarcs;
    const uint32_t *rrp;
    const uint4 *ell;   /* short rows of the view, or nullptr */
    const uint4 *lrows; /* the same rows staged in LDS (the Kahn step and the
                           fold read masked arcs from them), or nullptr */
    uint32_t n;
};

/* Short row of v (ELL_WIDE in .x when the view's rows are not in LDS or v's
 * row does not fit), and arc j of it as (target, weight): a short row lists
 * v's arcs in CSR order. */
__device__ __forceinline__ uint4 lds_row(const View &V, uint32_t v)
{
    return V.lrows ? V.lrows[v] : make_uint4(ELL_WIDE, ELL_NONE, ELL_NONE, ELL_NONE);
}
__device__ __forceinline__ uint2 row_arc(const uint4 r, uint32_t v, uint32_t j)
{
    const uint32_t e = j == 0 ? r.x : j == 1 ? r.y : j == 2 ? r.z : r.w;
    return make_uint2(v + (uint32_t)(int32_t)(int16_t)(e & 0xFFFFu), e >> 16);
}

/* Per-thread work counters, flushed into BrandesParams::stats. */
struct Counters {
    unsigned long long c[STAT_COUNT] = {};
};

/* ---- execution teams ------------------------------------------------------ */

/* Counters of one source's traversal shared by all threads of its team: in
 * LDS for a workgroup or wavefront team, in device memory for the grid team. */
struct Ctl {
    uint32_t qcnt[3];     /* near-queue lengths by round mod 3 */
    uint32_t fcnt[2];     /* far-list lengths by window parity */
    uint32_t minv[3];     /* smallest far distance by window mod 3 */
    uint32_t lcnt[3];     /* level-list entries pushed, by level mod 3 */
    uint32_t wide;        /* the packed sigma form overflowed */
};

/* LDS of one local group (a workgroup, or one wavefront of a wavefront team):
 * the vertices of degree > HEAVY_DEG met in the current round or level, whose
 * arcs the group's threads split, and the traversed-edge count. */
struct Local {
    uint32_t nheavy[2];   /* by round (level) parity */
    uint32_t heavy[HEAVY_CAP];
    unsigned long long te;
};

/* Per-workgroup LDS of the workgroup-per-source and grid kernels. */
/* Near queues kept in LDS: the first LQ entries of a round's queue (two
 * queues, this round's and the next), the rest in the slot's global queue.
 * A round then reads its vertices from LDS instead of one more dependent
 * global load. */
constexpr uint32_t LQ_WG = 1024, LQ_WAVE = 256, LQ_HALF = 128;

struct Shared {
    Ctl ctl;
    Local loc;
    uint32_t task;
    /* the counting sort's words: the entries, the buckets, the waves' sums */
    uint32_t sortbuf[SORT_CHUNK_MAX + SORT_BUCKETS + WG / 64];
    uint32_t lq[2 * LQ_WG];
};

/* Per-workgroup LDS of the wavefront-per-source kernel. */
struct SharedWave {
    Ctl ctl[WAVES];
    Local loc[WAVES];
    uint32_t lq[WAVES][2 * LQ_WAVE];
};

/* Per-workgroup LDS of the half-wavefront kernel: HALVES teams. */
struct SharedHalf {
    Ctl ctl[HALVES];
    Local loc[HALVES];
    uint32_t lq[HALVES][2 * LQ_HALF];
};

/* Bucket shift of a view of n vertices: the smallest that maps every id
 * into SORT_BUCKETS buckets. */
__device__ __forceinline__ uint32_t sort_shift(uint32_t n)
{
    uint32_t sh = 0;
    while (sh < 32 && ((n - 1) >> sh) >= SORT_BUCKETS) ++sh;
    return sh;
}

/* Group q[0, n) by v >> shift (n <= SORT_CHUNK_MAX; ascending buckets, any
 * order within one) with one counting pass through LDS, so that a round's or
 * a level's vertices are taken in id order at bucket granularity:
 * neighbouring ids share cache lines of the per-vertex state and of the CSR,
 * and a wavefront's loads then fall on fewer lines.  buf holds
 * SORT_CHUNK_MAX entries, SORT_BUCKETS counters and WG/64 wave sums.  Whole
 * workgroup; ends with a barrier. */
__device__ __noinline__ void bucket_vertices(uint32_t *q, uint32_t n, uint32_t shift, uint32_t *buf)
{
    constexpr uint32_t PER = SORT_BUCKETS / WG;
    static_assert(SORT_BUCKETS % WG == 0 && WG % 64 == 0, "bucket scan layout");
    uint32_t *const hist = buf + SORT_CHUNK_MAX, *const wsum = hist + SORT_BUCKETS;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    for (uint32_t i = tid; i < SORT_BUCKETS; i += WG) hist[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += WG) {
        const uint32_t v = q[i];
        buf[i] = v;
        atomicAdd(&hist[min(v >> shift, SORT_BUCKETS - 1)], 1u);
    }
    __syncthreads();
    /* exclusive prefix sums of the buckets: PER consecutive ones per thread,
     * a wavefront scan of the threads' totals, then the waves' totals */
    uint32_t tot = 0;
    for (uint32_t k = 0; k < PER; ++k) tot += hist[tid * PER + k];
    uint32_t inc = tot;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[tid >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - tot;
    for (uint32_t w = 0; w < (tid >> 6); ++w) off += wsum[w];
    for (uint32_t k = 0; k < PER; ++k) {   /* each thread rewrites only its own counters */
        const uint32_t c = hist[tid * PER + k];
        hist[tid * PER + k] = off;
        off += c;
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += WG) {
        const uint32_t v = buf[i];
        q[atomicAdd(&hist[min(v >> shift, SORT_BUCKETS - 1)], 1u)] = v;
    }
    __syncthreads();
}


/* A team runs one source.  rank/size/leader/sync span the team; the l*
 * members span its local group, which shares one Local in LDS. */

/* One workgroup per source (batched kernel). */
struct TeamWG {
    static constexpr int SC = __HIP_MEMORY_SCOPE_WORKGROUP;
    static constexpr int RB = 8, SUCC = 8, SWEEP_U = 1, MB = 2;
    static constexpr uint32_t LQ = LQ_WG;
    Ctl *ctl;
    Local *loc;
    uint32_t *sortbuf;
    uint32_t *lq;
    __device__ uint32_t rank() const { return threadIdx.x; }
    __device__ uint32_t size() const { return WG; }
    __device__ bool leader() const { return threadIdx.x == 0; }
    __device__ void sync() const { __syncthreads(); }
    __device__ bool aborted() const { return false; }
    __device__ uint32_t lrank() const { return threadIdx.x; }
    __device__ uint32_t lsize() const { return WG; }
    __device__ bool lleader() const { return threadIdx.x == 0; }
    __device__ void lsync() const { __syncthreads(); }
};

/* Barrier of one wavefront: its lanes' memory operations before it are
 * visible to its lanes after it. */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

/* One wavefront per source (small views): WAVES sources per workgroup, no
 * workgroup barriers.  Slot state keeps workgroup scope (the wavefront is
 * part of one workgroup). */
template <uint32_t LQN>
struct TeamWaveT {
    static constexpr int SC = __HIP_MEMORY_SCOPE_WORKGROUP;
    static constexpr int RB = 4, SUCC = 4, SWEEP_U = 2, MB = 4;
    static constexpr uint32_t LQ = LQN, LANES = 64;
    Ctl *ctl;
    Local *loc;
    uint32_t *lq;
    __device__ uint32_t rank() const { return threadIdx.x & 63; }
    __device__ uint32_t size() const { return 64; }
    __device__ bool leader() const { return (threadIdx.x & 63) == 0; }
    __device__ void sync() const { wave_sync(); }
    __device__ bool aborted() const { return false; }
    __device__ uint32_t lrank() const { return threadIdx.x & 63; }
    __device__ uint32_t lsize() const { return 64; }
    __device__ bool lleader() const { return (threadIdx.x & 63) == 0; }
    __device__ void lsync() const { wave_sync(); }
};

/* One half wavefront (32 lanes) per source: the two sources of a wavefront
 * run in lockstep, each on its own slot, counters and LDS; a branch on which
 * they disagree runs each half in turn.  For small views whose rounds and
 * levels hold a few dozen vertices, two sources then share each memory
 * latency of the step. */
template <uint32_t LQN>
struct TeamHalfT {
    static constexpr int SC = __HIP_MEMORY_SCOPE_WORKGROUP;
    static constexpr int RB = 4, SUCC = 4, SWEEP_U = 2, MB = 4;
    static constexpr uint32_t LQ = LQN, LANES = 32;
    Ctl *ctl;
    Local *loc;
    uint32_t *lq;
    __device__ uint32_t rank() const { return threadIdx.x & 31; }
    __device__ uint32_t size() const { return 32; }
    __device__ bool leader() const { return (threadIdx.x & 31) == 0; }
    __device__ void sync() const { wave_sync(); }
    __device__ bool aborted() const { return false; }
    __device__ uint32_t lrank() const { return threadIdx.x & 31; }
    __device__ uint32_t lsize() const { return 32; }
    __device__ bool lleader() const { return (threadIdx.x & 31) == 0; }
    __device__ void lsync() const { wave_sync(); }
};

/* The whole co-resident grid per source (plain kernel, launched with
 * hipLaunchCooperativeKernel, which guarantees that every workgroup is
 * resident, so the grid barrier always completes). */
struct TeamGrid {
    static constexpr int SC = __HIP_MEMORY_SCOPE_AGENT;
    static constexpr int RB = 8, SUCC = 8, SWEEP_U = 1, MB = 2;
    static constexpr uint32_t LQ = 0;   /* the grid shares no LDS */
    Ctl *ctl;
    Local *loc;
    uint32_t *lq;
    __device__ uint32_t rank() const { return blockIdx.x * WG + threadIdx.x; }
    __device__ uint32_t size() const { return gridDim.x * WG; }
    __device__ bool leader() const { return blockIdx.x == 0 && threadIdx.x == 0; }
    __device__ void sync() const { cooperative_groups::this_grid().sync(); }
    __device__ bool aborted() const { return false; }
    __device__ uint32_t lrank() const { return threadIdx.x; }
    __device__ uint32_t lsize() const { return WG; }
    __device__ bool lleader() const { return threadIdx.x == 0; }
    __device__ void lsync() const { __syncthreads(); }
};

using TeamWave = TeamWaveT<LQ_WAVE>;
using TeamHalf = TeamHalfT<LQ_HALF>;

template <class T> struct IsWG { static constexpr bool value = false; };
template <> struct IsWG<TeamWG> { static constexpr bool value = true; };
template <class T> struct IsHalf { static constexpr bool value = false; };
template <uint32_t N> struct IsHalf<TeamHalfT<N>> { static constexpr bool value = true; };

/* A place in a list counted by *cnt for each calling lane, one atomic per
 * wavefront: the lanes that append -- exactly those executing this call,
 * the others being off in the branch that leads here -- take consecutive
 * places; the first of them adds their number and each offsets from it by
 * the appending lanes below it.  The two half-wavefront teams of a
 * wavefront count their own lists, so their lanes group by half. */
template <class T>
__device__ __forceinline__ uint32_t team_append(uint32_t *cnt)
{
    const uint32_t lane = __lane_id();
    const uint64_t all = __ballot(1);   /* the calling lanes (wavefront-uniform) */
    const uint64_t below = all & ((1ull << lane) - 1ull);
    if constexpr (IsHalf<T>::value) {
        constexpr uint64_t LO = 0x00000000FFFFFFFFull;
        const uint64_t lo = all & LO, hi = all & ~LO;
        const uint64_t mine = lane < 32 ? lo : hi;
        uint32_t base = 0;
        if ((below & (lane < 32 ? LO : ~LO)) == 0) base = aadd32<T::SC>(cnt, (uint32_t)__popcll(mine));
        /* each half's first lane holds its base; both leaders are uniform */
        const uint32_t b0 = lo ? (uint32_t)__builtin_amdgcn_readlane((int)base, __ffsll((unsigned long long)lo) - 1) : 0u;
        const uint32_t b1 = hi ? (uint32_t)__builtin_amdgcn_readlane((int)base, __ffsll((unsigned long long)hi) - 1) : 0u;
        return (lane < 32 ? b0 : b1) + (uint32_t)__popcll(below & (lane < 32 ? LO : ~LO));
    } else {
        uint32_t base = 0;
        if (below == 0) base = aadd32<T::SC>(cnt, (uint32_t)__popcll(all));
        /* the first calling lane is the first active one */
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)base) + (uint32_t)__popcll(below);
    }
}

/* Append w to a near queue whose first T::LQ entries live in LDS (lq) and
 * the rest in the slot's global queue (q). */
template <class T>
__device__ __forceinline__ void push_near(uint32_t *lq, uint32_t *q, uint32_t *cnt, uint32_t w, bool agg)
{
    const uint32_t i = agg ? team_append<T>(cnt) : aadd32<T::SC>(cnt, 1u);
    if (T::LQ && i < T::LQ)
        lq[i] = w;
    else
        q[i] = w;
}

/* Entry i of such a queue. */
template <class T>
__device__ __forceinline__ uint32_t near_entry(const uint32_t *lq, const uint32_t *q, uint32_t i)
{
    return T::LQ && i < T::LQ ? lq[i] : q[i];
}

/* Relax arc e of u (distance du); used for the arcs of heavy vertices, one
 * per thread. */
template <class T>
__device__ __forceinline__ void relax_arc(const View &V, const Slot S, uint32_t du, uint32_t e,
                                          uint32_t thr, uint32_t nxt, uint32_t win, uint32_t *lnext,
                                          uint32_t *qnext, uint32_t *qn_cnt, uint32_t *far, uint32_t *far_cnt,
                                          uint32_t *far_min, uint32_t *err, Counters &ct)
{
    constexpr int SC = T::SC;
    ++ct.c[STAT_RELAX];
    const uint2 a = V.arcs[e];
    const uint64_t nd64 = (uint64_t)du + a.y;
    if (nd64 >= DINF) {
        atomicOr(err, ERR_DIST);
        return;
    }
    const uint32_t nd = (uint32_t)nd64;
    const uint64_t nk = ((uint64_t)nd << 32) | (nd < thr ? NEAR_TAG | nxt : win);
    if (nk >= ald64<SC>(&S.key[a.x])) return;
    const uint64_t old = amin64<SC>(&S.key[a.x], nk);
    if ((old >> 32) <= nd) return;
    ++ct.c[STAT_IMPROVE];
    if (nd < thr) {
        /* queued already for the next round (with a larger distance): not again */
        if ((uint32_t)old != (NEAR_TAG | nxt)) push_near<T>(lnext, qnext, qn_cnt, a.x, S.agg);
    } else {
        amin32<SC>(far_min, nd);
        /* on this window's far list already: not again */
        if ((uint32_t)old != win) {
            far[S.agg ? team_append<T>(far_cnt) : aadd32<SC>(far_cnt, 1u)] = a.x;
            ++ct.c[STAT_FARPUSH];
        }
    }
}

/* Relax NB arcs of u (distance du; arc j taken when ok[j]): the key loads and
 * the atomicMins of the block are independent, so they are issued back to
 * back and their latencies overlap. */
template <class T, int NB>
__device__ __forceinline__ void relax_arcs(const Slot S, uint32_t du, const uint2 (&a)[NB], const bool (&ok)[NB],
                                           uint32_t thr, uint32_t nxt, uint32_t win, uint32_t *lnext,
                                           uint32_t *qnext, uint32_t *qn_cnt, uint32_t *far, uint32_t *far_cnt,
                                           uint32_t *far_min, uint32_t *err, Counters &ct)
{
    constexpr int SC = T::SC;
    uint32_t to[NB];
    uint64_t nk[NB], cur[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        to[j] = a[j].x;
        nk[j] = ~uint64_t(0);
        if (ok[j]) {
            const uint64_t nd64 = (uint64_t)du + a[j].y;
            if (nd64 >= DINF) {
                atomicOr(err, ERR_DIST);
            } else {
                const uint32_t nd = (uint32_t)nd64;
                nk[j] = ((uint64_t)nd << 32) | (nd < thr ? NEAR_TAG | nxt : win);
            }
        }
    }
    /* every load of the block is issued, an arc not taken reading key[0]
     * (always a vertex), so no branch separates them and they are in flight
     * together; its nk (all ones) then never passes the filter */
    uint32_t la[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) la[j] = nk[j] != ~uint64_t(0) ? to[j] : 0u;
#pragma unroll
    for (int j = 0; j < NB; ++j) cur[j] = ald64<SC>(&S.key[la[j]]);
#pragma unroll
    for (int j = 0; j < NB; ++j) cur[j] = nk[j] < cur[j] ? amin64<SC>(&S.key[to[j]], nk[j]) : 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        /* cur[j] is the key before our atomicMin (0 when none was issued) */
        const uint32_t nd = (uint32_t)(nk[j] >> 32);
        if ((cur[j] >> 32) <= nd) continue;
        ++ct.c[STAT_IMPROVE];
        if (nd < thr) {
            if ((uint32_t)cur[j] != (NEAR_TAG | nxt)) push_near<T>(lnext, qnext, qn_cnt, to[j], S.agg);
        } else {
            amin32<SC>(far_min, nd);
            if ((uint32_t)cur[j] != win) {
                far[S.agg ? team_append<T>(far_cnt) : aadd32<SC>(far_cnt, 1u)] = to[j];
                ++ct.c[STAT_FARPUSH];
            }
        }
    }
}

/* Relax the arcs [e0, e1) of u (distance du), RB at a time. */
template <class T>
__device__ __forceinline__ void relax_vertex(const View &V, const Slot S, uint32_t du, uint32_t e0,
                                             uint32_t e1, uint32_t thr, uint32_t nxt, uint32_t win,
                                             uint32_t *lnext, uint32_t *qnext, uint32_t *qn_cnt, uint32_t *far,
                                             uint32_t *far_cnt, uint32_t *far_min, uint32_t *err,
                                             Counters &ct)
{
    constexpr int RB = T::RB;
    ct.c[STAT_RELAX] += e1 - e0;
    for (uint32_t eb = e0; eb < e1; eb += RB) {
        uint2 a[RB];
        bool ok[RB];
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            /* past the row's end: its last arc again (a valid index, e1 > eb),
             * so the block's loads carry no branch and go out together */
            ok[j] = eb + j < e1;
            a[j] = V.arcs[ok[j] ? eb + j : e1 - 1];
        }
        relax_arcs<T, RB>(S, du, a, ok, thr, nxt, win, lnext, qnext, qn_cnt, far, far_cnt, far_min, err, ct);
    }
}

/* Phase 1: near-far delta-stepping.  On return the distances are final.
 *
 * Round r expands the near queue counted by qcnt[r % 3] and fills the one
 * counted by qcnt[(r + 1) % 3]; its leader clears qcnt[(r + 2) % 3], last
 * read two rounds earlier, so one team barrier per round suffices.  Far-list
 * counters and window minima alternate by window parity in the same way. */
template <class T>
__device__ __forceinline__ void phase_distances(const View &V, const Slot S, const T &t, uint32_t src,
                                uint32_t delta, uint32_t sort_cap, uint32_t *err, Counters &ct)
{
    constexpr int SC = T::SC;
    const uint32_t r0 = t.rank(), R = t.size();
    Ctl &c = *t.ctl;
    for (uint32_t v = r0; v < V.n; v += R) {
        S.key[v] = v == src ? NEAR_TAG | 1u : (uint64_t)DINF << 32;   /* src: distance 0, queued for round 1 */
    }
    if (t.leader()) {
        ast32<SC>(&c.qcnt[0], 0u);
        ast32<SC>(&c.qcnt[1], 1u);
        ast32<SC>(&c.qcnt[2], 0u);
        ast32<SC>(&c.fcnt[0], 0u);
        ast32<SC>(&c.fcnt[1], 0u);
        ast32<SC>(&c.minv[0], DINF);
        ast32<SC>(&c.minv[1], DINF);
        ast32<SC>(&c.minv[2], DINF);
        ast32<SC>(&c.wide, 0u);
        if (T::LQ)
            t.lq[0] = src;
        else
            S.qa[0] = src;
    }
    Local &L = *t.loc;
    if (t.lleader()) L.nheavy[0] = L.nheavy[1] = 0;
    t.sync();

    uint32_t *qa = S.qa, *qb = S.qb, *fa = S.fa, *fb = S.fb;
    uint32_t *lqa = t.lq, *lqb = t.lq + T::LQ;   /* the queues' LDS heads */
    uint32_t near = 1, rnd = 1, win = 1;
    uint32_t thr = delta;
    for (;;) {
        while (near) {
            const uint32_t nxt = rnd + 1;
            uint32_t *qn_cnt = &c.qcnt[nxt % 3];
            uint32_t *far_cnt = &c.fcnt[win & 1];
            uint32_t *far_min = &c.minv[win % 3];
            uint32_t *nh_cnt = &L.nheavy[rnd & 1];
            if (t.leader()) {
                ++ct.c[STAT_ROUNDS];
                ast32<SC>(&c.qcnt[(rnd + 2) % 3], 0u);
            }
            if (t.lleader()) L.nheavy[(rnd + 1) & 1] = 0;
            if constexpr (IsWG<T>::value)
                if (near > 64 && sort_cap) {
                    /* the queue's LDS head, then its remainder in HBM */
                    const uint32_t sh = sort_shift(V.n), head = min(near, T::LQ);
                    for (uint32_t c0 = 0; c0 < head; c0 += sort_cap)
                        bucket_vertices(lqa + c0, min(sort_cap, head - c0), sh, t.sortbuf);
                    for (uint32_t c0 = T::LQ; c0 < near; c0 += sort_cap)
                        bucket_vertices(qa + c0, min(sort_cap, near - c0), sh, t.sortbuf);
                }
            for (uint32_t i = r0; i < near; i += R) {
                const uint32_t u = near_entry<T>(lqa, qa, i);
                ++ct.c[STAT_VISITS];
                const uint32_t du = dist_of<SC>(S, u);
                if (V.ell) {
                    /* a short row: its arcs in one 16-byte load, no row starts */
                    /* relaxed whole and before the test for a wide row (which
                     * has no arc to take: ELL_WIDE, then ELL_NONE), so the
                     * row is one 16-byte load rather than a word, a test and
                     * a dependent load of the rest */
                    const uint4 r = V.ell[u];
                    const uint32_t e[ELL_W] = {r.x, r.y, r.z, r.w};
                    uint2 a[ELL_W];
                    bool ok[ELL_W];
#pragma unroll
                    for (int j = 0; j < (int)ELL_W; ++j) {
                        ok[j] = e[j] < ELL_WIDE;
                        a[j] = make_uint2(u + (uint32_t)(int32_t)(int16_t)(e[j] & 0xFFFFu), e[j] >> 16);
                        ct.c[STAT_RELAX] += ok[j];
                    }
                    relax_arcs<T, ELL_W>(S, du, a, ok, thr, nxt, win, lqb, qb, qn_cnt, fa, far_cnt, far_min, err, ct);
                    if (r.x != ELL_WIDE) continue;
                }
                const uint32_t e0 = V.rp[u], e1 = V.rp[u + 1];
                if (e1 - e0 > (uint32_t)HEAVY_DEG) {
                    const uint32_t h = atomicAdd(nh_cnt, 1u);
                    if (h < (uint32_t)HEAVY_CAP) {
                        L.heavy[h] = u;
                        continue;
                    }
                }
                relax_vertex<T>(V, S, du, e0, e1, thr, nxt, win, lqb, qb, qn_cnt, fa, far_cnt, far_min, err, ct);
            }
            t.lsync();
            const uint32_t nh = *nh_cnt < (uint32_t)HEAVY_CAP ? *nh_cnt : (uint32_t)HEAVY_CAP;
            for (uint32_t h = 0; h < nh; ++h) {   /* heavy vertices: the local group splits the arcs */
                const uint32_t u = L.heavy[h];
                const uint32_t du = dist_of<SC>(S, u);
                const uint32_t e1 = V.rp[u + 1];
                for (uint32_t e = V.rp[u] + t.lrank(); e < e1; e += t.lsize())
                    relax_arc<T>(V, S, du, e, thr, nxt, win, lqb, qb, qn_cnt, fa, far_cnt, far_min, err, ct);
            }
            t.sync();
            near = ald32<SC>(qn_cnt);
            uint32_t *tq = qa; qa = qb; qb = tq;
            uint32_t *tl = lqa; lqa = lqb; lqb = tl;
            rnd = nxt;
            if (t.aborted()) return;
        }
        /* window [.., thr) is settled; move on to the nearest far vertex.  The
         * near queue of round rnd is empty and its counter is 0.  minv[win % 3]
         * is the smallest distance any far relaxation (or carried entry) of
         * this window set: a lower bound of the far distances (an entry may
         * since have become near), and at least thr, so the next window
         * [thr, m + delta) is never empty of progress.  Far entries are
         * unique within a window (the key's window tag), so each is queued at
         * most once and needs no round tag. */
        const uint32_t nfar = ald32<SC>(&c.fcnt[win & 1]);
        if (nfar == 0) break;
        const uint32_t m = ald32<SC>(&c.minv[win % 3]);
        if (t.leader()) {
            ++ct.c[STAT_WINDOWS];
            ast32<SC>(&c.fcnt[(win + 1) & 1], 0u);   /* last read in the previous window */
            ast32<SC>(&c.minv[(win + 2) % 3], DINF); /* last read in the previous window */
        }
        if (m == DINF) break;
        const uint64_t nthr64 = (uint64_t)m + delta;
        const uint32_t nthr = nthr64 >= DINF ? DINF : (uint32_t)nthr64;
        const uint32_t nwin = win + 1;
        for (uint32_t i = r0; i < nfar; i += R) {
            const uint32_t w = fa[i];
            const uint32_t d = dist_of<SC>(S, w);   /* no relaxation runs during the split */
            if (d < thr) continue;            /* settled in this window */
            if (d < nthr) {
                push_near<T>(lqa, qa, &c.qcnt[rnd % 3], w, S.agg);
            } else {
                /* carried to the next window's list, tagged as on it */
                amin32<SC>(&c.minv[nwin % 3], d);
                ast64<SC>(&S.key[w], ((uint64_t)d << 32) | nwin);
                fb[S.agg ? team_append<T>(&c.fcnt[nwin & 1]) : aadd32<SC>(&c.fcnt[nwin & 1], 1u)] = w;
            }
        }
        t.sync();
        near = ald32<SC>(&c.qcnt[rnd % 3]);
        uint32_t *tf = fa; fa = fb; fb = tf;
        thr = nthr;
        win = nwin;
        if (t.aborted()) return;
    }
}

/* Kahn arrivals of one block: sigma(u) = su joins each listed successor
 * w[k] (ok[k]) and one of its predecessors is taken off, the atomics of the
 * block in flight together; a successor whose last predecessor this was joins
 * the next level, at lbase + its place in that level.  Packed form: a sigma
 * carry into the count, or more arrivals than the level list holds (only
 * after such a carry or a count beyond PK_CMAX), raises *wide. */
template <bool PACKED, class T, int RB>
__device__ __forceinline__ void arrive(const Slot S, uint32_t n, const uint32_t (&w)[RB], const bool (&ok)[RB],
                                       uint64_t su, uint32_t lbase, uint32_t *lcnt, uint32_t *lqn,
                                       uint32_t *wide, uint32_t *err)
{
    constexpr int SC = T::SC;
    constexpr uint32_t LQ = T::LQ;
    if constexpr (PACKED) {
        uint64_t old[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) old[k] = ok[k] ? aadd64<SC>(&S.sd[2 * (uint64_t)w[k]], su - PK_ONE) : 0;
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            if (!ok[k]) continue;
            if ((old[k] & PK_SIGMA) + su > PK_SIGMA) ast32<SC>(wide, 1u);
            if ((old[k] >> 48) == 1u) {
                const uint32_t x = S.agg ? team_append<T>(lcnt) : aadd32<SC>(lcnt, 1u), i = lbase + x;
                if (i < n) {
                    S.qa[i] = w[k];
                    if (LQ && x < LQ) lqn[x] = w[k];
                } else {
                    ast32<SC>(wide, 1u);
                }
            }
        }
    } else {
        uint64_t old[RB], cnt[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            old[k] = ok[k] ? aadd64<SC>(&S.sd[2 * (uint64_t)w[k]], su) : 0;
            cnt[k] = ok[k] ? asub64<SC>(&S.sd[2 * (uint64_t)w[k] + 1], PRED_ONE) : 0;
        }
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            if (!ok[k]) continue;
            if (old[k] + su < old[k])
                atomicOr(err, ERR_SIGMA);
            else if (old[k] + su >= (1ull << 53))
                atomicOr(err, ERR_SIGMA_INEXACT);   /* the reference's FP64 sigma rounds here */
            if ((cnt[k] >> 32) == 1u) {
                const uint32_t x = S.agg ? team_append<T>(lcnt) : aadd32<SC>(lcnt, 1u);
                S.qa[lbase + x] = w[k];
                if (LQ && x < LQ) lqn[x] = w[k];
            }
        }
    }
    (void)n;
    (void)wide;
    (void)err;
    (void)lqn;
}

/* Push sigma(u) along u's tight arcs, given by its mask (bit j: arc e0 + j),
 * RB at a time. */
template <bool PACKED, class T>
__device__ __forceinline__ void push_sigma_masked(const View &V, const Slot S, uint32_t u, uint64_t su, uint32_t e0,
                                                  uint32_t mask, uint32_t lbase, uint32_t *lcnt, uint32_t *lqn,
                                                  uint32_t *wide, uint32_t *err)
{
    constexpr int RB = T::MB;   /* unused slots load arc e0 again: no branches */
    const uint4 row = lds_row(V, u);
    const bool lr = row.x != ELL_WIDE;
    uint32_t m = mask;
    while (m) {
        uint32_t j[RB];
        bool ok[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            ok[k] = m != 0;
            j[k] = ok[k] ? (uint32_t)__builtin_ctz(m) : 0u;
            m &= m - 1;
        }
        uint32_t w[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k)   /* j[k] = 0 when not ok: arc e0 exists (mask != 0) */
            w[k] = lr ? row_arc(row, u, j[k]).x : V.arcs[e0 + j[k]].x;
        arrive<PACKED, T, RB>(S, V.n, w, ok, su, lbase, lcnt, lqn, wide, err);
    }
}

/* Push sigma(u) along the DAG arcs among [e0, e1) (every step-th), testing
 * each against the distances, RB at a time with the loads and atomics of a
 * block in flight together. */
template <bool PACKED, class T>
__device__ __forceinline__ void push_sigma(const View &V, const Slot S, uint32_t du, uint64_t su,
                                           uint32_t e0, uint32_t e1, uint32_t step, uint32_t lbase,
                                           uint32_t *lcnt, uint32_t *lqn, uint32_t *wide, uint32_t *err)
{
    constexpr int SC = T::SC, RB = T::RB;
    for (uint32_t eb = e0; eb < e1; eb += RB * step) {
        uint2 a[RB];
#pragma unroll
        for (int j = 0; j < RB; ++j) a[j] = V.arcs[eb + j * step < e1 ? eb + j * step : eb];   /* eb < e1: valid */
        bool tight[RB];
        uint32_t w[RB], dw[RB];
#pragma unroll
        for (int j = 0; j < RB; ++j) dw[j] = dist_of<SC>(S, a[j].x);   /* no branch: in flight together */
#pragma unroll
        for (int j = 0; j < RB; ++j) {
            tight[j] = eb + j * step < e1 && (uint64_t)du + a[j].y == dw[j];
            w[j] = a[j].x;
        }
        arrive<PACKED, T, RB>(S, V.n, w, tight, su, lbase, lcnt, lqn, wide, err);
    }
}

/* Phase 2 in one sigma form: DAG in-degrees and tight-arc masks, then the
 * Kahn sweep with exact sigma.  Returns the number of levels; the level list
 * is S.qa, the level starts S.fa.  The sweep walks the vertices in order, so
 * the distance tests it makes read neighbouring lines, and it writes every
 * vertex's (sigma, mask) pair whole (packed: with the row start beside the
 * mask); the Kahn sweep and the dependency phase, which visit vertices
 * scattered, then touch only the successors' sigma lines.  Level l's arrivals
 * are counted in lcnt[(l + 1) % 3], which every thread reads after the
 * level's barrier and the leader clears two levels later, so the threads
 * agree on every level's extent with one barrier per level.  In the packed
 * form the team's Ctl::wide flag reports an overflow (the result is then
 * void). */
template <bool PACKED, class T>
__device__ __forceinline__ uint32_t sigma_form(const View &V, const Slot S, const T &t, uint32_t src,
                                               uint32_t sort_cap, bool count_te, uint32_t *err, Counters &ct)
{
    constexpr int SC = T::SC;
    const uint32_t r0 = t.rank(), R = t.size();
    Ctl &c = *t.ctl;
    unsigned long long te = 0;
    /* SU vertices per thread and step, their out- and in-arcs SB at a time.
     * The loads of a block carry no branch -- past a row's end they read the
     * row's last arc again (or, for an empty row beside a non-empty one, an
     * arc next to the row: the view then has arcs), a step's vertex past the
     * view's end is its last vertex again, and a slot's distance load reads
     * v's own -- so the step's arc loads, then its distance loads, are in
     * flight together */
    constexpr int SU = T::SWEEP_U, SB = 4;
    for (uint32_t v0 = r0; v0 < V.n; v0 += SU * R) {
        uint32_t vv[SU], dv[SU], e0[SU], e1[SU], f0[SU], f1[SU], os[SU], is[SU], cin[SU], mask[SU];
        bool in[SU];
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            in[k] = v0 + k * R < V.n;
            vv[k] = in[k] ? v0 + k * R : V.n - 1;
            dv[k] = dist_of<SC>(S, vv[k]);
            e0[k] = V.rp[vv[k]];
            e1[k] = V.rp[vv[k] + 1];
            f0[k] = V.rrp[vv[k]];
            f1[k] = V.rrp[vv[k] + 1];
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            cin[k] = 0;
            mask[k] = 0;
            if (!in[k] || dv[k] == DINF) {
                e1[k] = e0[k];   /* unreached (or no vertex): nothing to test */
                f1[k] = f0[k];
            } else {
                te += e1[k] - e0[k];
                if (vv[k] == src) f1[k] = f0[k];   /* the source has no predecessors */
            }
            os[k] = e1[k] > e0[k] ? e1[k] - 1 : e0[k] > 0 ? e0[k] - 1 : 0u;
            is[k] = f1[k] > f0[k] ? f1[k] - 1 : f0[k] > 0 ? f0[k] - 1 : 0u;
        }
        for (uint32_t b = 0;; b += SB) {
            bool any = false;
#pragma unroll
            for (int k = 0; k < SU; ++k) any = any || e0[k] + b < e1[k] || f0[k] + b < f1[k];
            if (!any) break;
            /* a vertex with no arc left in this block reads arc os / is, which
             * exists when any row of the step is non-empty only if its own
             * view has arcs: clamp it to a step vertex's safe arc that does */
            uint32_t osafe = 0, isafe = 0;
#pragma unroll
            for (int k = SU - 1; k >= 0; --k) {
                if (e1[k] > e0[k] || f1[k] > f0[k]) {
                    osafe = os[k];
                    isafe = is[k];
                }
            }
            uint2 ao[SU][SB], ai[SU][SB];
            bool oo[SU][SB], oi[SU][SB];
#pragma unroll
            for (int k = 0; k < SU; ++k)
#pragma unroll
                for (int jj = 0; jj < SB; ++jj) {
                    oo[k][jj] = e0[k] + b + jj < e1[k];
                    oi[k][jj] = f0[k] + b + jj < f1[k];
                    ao[k][jj] = V.arcs[oo[k][jj] ? e0[k] + b + jj : osafe];
                    ai[k][jj] = V.rarcs[oi[k][jj] ? f0[k] + b + jj : isafe];
                }
            uint32_t dout[SU][SB], din[SU][SB];
#pragma unroll
            for (int k = 0; k < SU; ++k)
#pragma unroll
                for (int jj = 0; jj < SB; ++jj) {
                    dout[k][jj] = dist_of<SC>(S, oo[k][jj] ? ao[k][jj].x : vv[k]);
                    din[k][jj] = dist_of<SC>(S, oi[k][jj] ? ai[k][jj].x : vv[k]);
                }
#pragma unroll
            for (int k = 0; k < SU; ++k)
#pragma unroll
                for (int jj = 0; jj < SB; ++jj) {
                    if (oo[k][jj] && dout[k][jj] != DINF && (uint64_t)dv[k] + ao[k][jj].y == dout[k][jj] && b + jj < 32)
                        mask[k] |= 1u << (b + jj);
                    cin[k] += oi[k][jj] && din[k][jj] != DINF && (uint64_t)din[k][jj] + ai[k][jj].y == dv[k];
                }
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            if (!in[k]) break;
            ulonglong2 pair = make_ulonglong2(0, 0);   /* unreached: sigma 0 */
            if (dv[k] != DINF) {
                const uint32_t mk = e1[k] - e0[k] > MASK_ARCS ? MASK_NONE : mask[k];
                if (PACKED) {
                    if (cin[k] > PK_CMAX) ast32<SC>(&c.wide, 1u);
                    pair.x = vv[k] == src ? 1u : (uint64_t)cin[k] << 48;
                    pair.y = ((uint64_t)e0[k] << 32) | mk;   /* the row start rides along */
                } else {
                    pair.x = vv[k] == src ? 1u : 0u;
                    pair.y = ((uint64_t)cin[k] << 32) | mk;
                }
            }
            reinterpret_cast<ulonglong2 *>(S.sd)[vv[k]] = pair;
        }
    }
    Local &L = *t.loc;
    if (te && count_te) atomicAdd(&L.te, te);
    if (t.leader()) {
        S.qa[0] = src;
        if (T::LQ) t.lq[0] = src;
        S.fa[0] = 0;
        S.fa[1] = 1;
        ast32<SC>(&c.lcnt[1], 0u);
    }
    if (t.lleader()) L.nheavy[0] = L.nheavy[1] = 0;
    t.sync();
    /* a level's first LQ entries are also kept in LDS (the team's queue
     * heads, free in this phase), so the next level reads them without a
     * dependent global load; the dependency phase reads S.qa */
    uint32_t *lqc = t.lq, *lqn = T::LQ ? t.lq + T::LQ : nullptr;
    uint32_t lb = 0, le = 1, nlev = 1;
    for (uint32_t lvl = 0; lb < le; ++lvl) {
        uint32_t *nh_cnt = &L.nheavy[lvl & 1];
        uint32_t *lc = &c.lcnt[(lvl + 1) % 3];
        if (t.lleader()) L.nheavy[(lvl + 1) & 1] = 0;
        if (t.leader()) ast32<SC>(&c.lcnt[(lvl + 2) % 3], 0u);
        bool sorted = false;   /* a sorted level is read back from S.qa */
        if constexpr (IsWG<T>::value)
            if (le - lb > 64 && sort_cap) {   /* every entry of a level is in S.qa */
                const uint32_t sh = sort_shift(V.n);
                for (uint32_t c0 = lb; c0 < le; c0 += sort_cap)
                    bucket_vertices(S.qa + c0, min(sort_cap, le - c0), sh, t.sortbuf);
                sorted = true;
            }
        for (uint32_t i = lb + r0; i < le; i += R) {
            const uint32_t u = T::LQ && !sorted && i - lb < T::LQ ? lqc[i - lb] : S.qa[i];
            /* u's sigma, mask (and, packed, row start): final since u joined
             * the list */
            uint64_t su, y;
            if constexpr (SC == __HIP_MEMORY_SCOPE_WORKGROUP) {
                const ulonglong2 pr = sd_pair(S, u);   /* one 16-byte load */
                su = pr.x;
                y = pr.y;
            } else {
                su = sigma_of<SC>(S, u);
                y = ald64<SC>(&S.sd[2 * (uint64_t)u + 1]);
            }
            const uint32_t mask = (uint32_t)y;
            if (PACKED && mask != MASK_NONE) {   /* at most MASK_ARCS arcs: never heavy */
                push_sigma_masked<PACKED, T>(V, S, u, su, (uint32_t)(y >> 32), mask, le, lc, lqn, &c.wide, err);
                continue;
            }
            const uint32_t e0 = V.rp[u], e1 = V.rp[u + 1];
            if (e1 - e0 > (uint32_t)HEAVY_DEG) {
                const uint32_t h = atomicAdd(nh_cnt, 1u);
                if (h < (uint32_t)HEAVY_CAP) {
                    L.heavy[h] = u;
                    continue;
                }
            }
            if (mask != MASK_NONE)
                push_sigma_masked<PACKED, T>(V, S, u, su, e0, mask, le, lc, lqn, &c.wide, err);
            else
                push_sigma<PACKED, T>(V, S, dist_of<SC>(S, u), su, e0, e1, 1, le, lc, lqn, &c.wide, err);
        }
        t.lsync();
        const uint32_t nh = *nh_cnt < (uint32_t)HEAVY_CAP ? *nh_cnt : (uint32_t)HEAVY_CAP;
        for (uint32_t h = 0; h < nh; ++h) {
            const uint32_t u = L.heavy[h];
            push_sigma<PACKED, T>(V, S, dist_of<SC>(S, u), sigma_of<SC>(S, u), V.rp[u] + t.lrank(), V.rp[u + 1],
                                  t.lsize(), le, lc, lqn, &c.wide, err);
        }
        t.sync();
        if (T::LQ) {
            uint32_t *tq = lqc;
            lqc = lqn;
            lqn = tq;
        }
        lb = le;
        const uint32_t ne = le + ald32<SC>(lc);
        le = ne < V.n ? ne : V.n;   /* more only after a packed overflow */
        if (le > lb) {
            ++nlev;
            if (t.leader()) S.fa[nlev] = le;
        }
        if (t.aborted()) return 0;
    }
    return nlev;
}

/* Phase 2: the packed sigma form, and the wide one when the packed form
 * overflowed.  The overflow flag is read after the packed form's last
 * barrier, when no thread sets it any more; the next source's phase 1 clears
 * it. */
template <class T>
__device__ __forceinline__ uint32_t phase_sigma(const View &V, const Slot S, const T &t, uint32_t src,
                                                uint32_t sort_cap, uint32_t *err, Counters &ct)
{
    uint32_t nlev = sigma_form<true>(V, S, t, src, sort_cap, true, err, ct);
    if (ald32<T::SC>(&t.ctl->wide)) {
        if (t.leader()) ++ct.c[STAT_SIGMA_WIDE];
        t.sync();
        nlev = sigma_form<false>(V, S, t, src, sort_cap, false, err, ct);
    }
    if (t.leader()) ct.c[STAT_LEVELS] += nlev;
    return nlev;
}

/* Contribution sigma_v / sigma_w * (1 + delta_w), FP64, reference order of
 * operations (DijkstraSSBrandesBC.h:69). */
__device__ __forceinline__ double contrib64(double sv, uint64_t sw, uint64_t dw_bits)
{
    const double dw = __longlong_as_double((long long)dw_bits);
    return sv / (double)sw * (1.0 + dw);
}

template <int PREC>
struct Acc;

template <>
struct Acc<0> {
    double v = 0.0;
    __device__ void add(uint64_t sv_int, double sv, uint64_t sw, uint64_t dw_bits)
    {
        (void)sv_int;
        v += contrib64(sv, sw, dw_bits);
    }
    __device__ uint64_t bits() const { return (uint64_t)__double_as_longlong(v); }
    __device__ double value() const { return v; }
    static __device__ double value_of(uint64_t b) { return __longlong_as_double((long long)b); }
};

template <>
struct Acc<1> {
    F2 v = {0.0f, 0.0f};
    __device__ void add(uint64_t sv_int, double sv, uint64_t sw, uint64_t dw_bits)
    {
        (void)sv;
        const F2 ratio = f2_div(f2_from_u64(sv_int), f2_from_u64(sw));
        const F2 dw = f2_unpack(dw_bits);
        const F2 one_plus = f2_add({1.0f, 0.0f}, dw);
        v = f2_add(v, f2_mul(ratio, one_plus));
    }
    __device__ uint64_t bits() const { return f2_pack(v); }
    __device__ double value() const { return (double)v.hi + (double)v.lo; }
    static __device__ double value_of(uint64_t b)
    {
        const F2 x = f2_unpack(b);
        return (double)x.hi + (double)x.lo;
    }
};

/* Dependency of v: fold over DAG successors.  'ordered' sums them in
 * descending (d(w), w).  Distances, sigma and the successors' deltas are
 * final here, so plain loads see them after the level barrier.  With a
 * tight-arc mask the successors are known: d(w) = d(v) + w(v, w) and only
 * their (sigma, delta) pairs are read. */
template <int PREC, bool ORDERED, int RB, int SUCC_REG, int MB>
__device__ __forceinline__ Acc<PREC> fold_successors(const View &V, const Slot S, uint32_t v, bool packed)
{
    Acc<PREC> acc;
    const ulonglong2 own = sd_pair(S, v);
    const uint64_t sv_int = own.x;
    const uint32_t mask = (uint32_t)own.y;
    const double sv = (double)sv_int;
    /* the packed sigma form left a masked row's start beside its mask (the
     * wide form's high half is its spent predecessor count) */
    const uint32_t e0 = packed && mask != MASK_NONE ? (uint32_t)(own.y >> 32) : V.rp[v];
    if (!ORDERED && mask != MASK_NONE) {
        /* ascending arc order, MB successors' loads in flight together */
        uint32_t m = mask;
        while (m) {
            uint32_t j[MB];
            bool ok[MB];
#pragma unroll
            for (int k = 0; k < MB; ++k) {
                ok[k] = m != 0;
                j[k] = ok[k] ? (uint32_t)__builtin_ctz(m) : 0u;
                m &= m - 1;
            }
            uint32_t w[MB];
#pragma unroll
            for (int k = 0; k < MB; ++k) w[k] = V.arcs[e0 + j[k]].x;   /* j[k] = 0 when not ok: arc e0 exists */
            ulonglong2 sw[MB];
#pragma unroll
            for (int k = 0; k < MB; ++k) sw[k] = sd_pair(S, w[k]);   /* loads without branches: in flight together */
#pragma unroll
            for (int k = 0; k < MB; ++k)
                if (ok[k]) acc.add(sv_int, sv, sw[k].x, sw[k].y);
        }
        return acc;
    }
    const uint32_t dv = dist_final(S, v);   /* the masked unordered fold needs no distance */
    if (ORDERED && mask != MASK_NONE) {
        if (__popc(mask) <= SUCC_REG) {
            /* the successors' (sigma, delta) loads go out together; only the
             * sums follow the order */
            if (mask == 0) return acc;   /* no DAG successor */
            const uint4 row = lds_row(V, v);
            const bool lr = row.x != ELL_WIDE;
            /* slots past the successors repeat the first one (a valid arc and
             * vertex), so the arc loads, then the (sigma, delta) loads, carry
             * no branch and are in flight together */
            const uint32_t j0 = (uint32_t)__builtin_ctz(mask);
            uint32_t m = mask, aj[SUCC_REG];
            bool ok[SUCC_REG];
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j) {
                ok[j] = m != 0;
                aj[j] = ok[j] ? (uint32_t)__builtin_ctz(m) : j0;
                m &= m - 1;
            }
            uint2 a[SUCC_REG];
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j) a[j] = lr ? row_arc(row, v, aj[j]) : V.arcs[e0 + aj[j]];
            uint64_t key[SUCC_REG];
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j) key[j] = ok[j] ? ((uint64_t)(dv + a[j].y) << 32) | a[j].x : 0;
            ulonglong2 sw[SUCC_REG];
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j) sw[j] = sd_pair(S, a[j].x);
#pragma unroll
            for (int r = 0; r < SUCC_REG; ++r) {
                uint64_t best = 0;
                int bj = -1;
#pragma unroll
                for (int j = 0; j < SUCC_REG; ++j)
                    if (key[j] > best) { best = key[j]; bj = j; }
                if (bj < 0) break;
                ulonglong2 b = make_ulonglong2(0, 0);
#pragma unroll
                for (int j = 0; j < SUCC_REG; ++j)
                    if (j == bj) {
                        b = sw[j];
                        key[j] = 0;
                    }
                acc.add(sv_int, sv, b.x, b.y);
            }
            return acc;
        }
        /* more successors than registers: repeated selection over the mask */
        uint64_t prev = ~uint64_t(0);
        for (;;) {
            uint64_t best = 0;
            for (uint32_t m = mask; m; m &= m - 1) {
                const uint2 a = V.arcs[e0 + __builtin_ctz(m)];
                const uint64_t k = ((uint64_t)(dv + a.y) << 32) | a.x;
                if (k < prev && k > best) best = k;
            }
            if (best == 0) break;
            const ulonglong2 sw = sd_pair(S, (uint32_t)best);
            acc.add(sv_int, sv, sw.x, sw.y);
            prev = best;
        }
        return acc;
    }
    const uint32_t e1 = V.rp[v + 1];
    if (!ORDERED) {
        /* ascending arc order, RB successors' loads in flight together */
        for (uint32_t eb = e0; eb < e1; eb += RB) {
            uint2 a[RB];
#pragma unroll
            /* past the row's end: arc eb again (valid), so the arc, distance
             * and (sigma, delta) loads carry no branch */
            for (int j = 0; j < RB; ++j) a[j] = V.arcs[eb + j < e1 ? eb + j : eb];
            uint32_t dw[RB];
#pragma unroll
            for (int j = 0; j < RB; ++j) dw[j] = dist_final(S, a[j].x);
            bool tight[RB];
#pragma unroll
            for (int j = 0; j < RB; ++j) tight[j] = eb + j < e1 && (uint64_t)dv + a[j].y == dw[j];
            ulonglong2 sw[RB];
#pragma unroll
            for (int j = 0; j < RB; ++j) sw[j] = sd_pair(S, a[j].x);
#pragma unroll
            for (int j = 0; j < RB; ++j)
                if (tight[j]) acc.add(sv_int, sv, sw[j].x, sw[j].y);
        }
        return acc;
    }
    if (e1 == e0) return acc;   /* no out-arc */
    if (e1 - e0 <= (uint32_t)SUCC_REG) {
        /* past the row's end: arc e0 again (valid), so the loads carry no
         * branch */
        uint2 a[SUCC_REG];
#pragma unroll
        for (int j = 0; j < SUCC_REG; ++j) a[j] = V.arcs[e0 + j < e1 ? e0 + j : e0];
        uint32_t dwj[SUCC_REG];
#pragma unroll
        for (int j = 0; j < SUCC_REG; ++j) dwj[j] = dist_final(S, a[j].x);
        uint64_t key[SUCC_REG];
#pragma unroll
        for (int j = 0; j < SUCC_REG; ++j)
            key[j] = e0 + j < e1 && (uint64_t)dv + a[j].y == dwj[j] ? ((uint64_t)dwj[j] << 32) | a[j].x : 0;
        /* keys are distinct and nonzero for successors (dw >= 1); the low
         * half of the best key is the successor's id */
#pragma unroll
        for (int r = 0; r < SUCC_REG; ++r) {
            uint64_t best = 0;
            int bj = -1;
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j)
                if (key[j] > best) { best = key[j]; bj = j; }
            if (bj < 0) break;
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j)
                if (j == bj) key[j] = 0;
            const ulonglong2 b = sd_pair(S, (uint32_t)best);
            acc.add(sv_int, sv, b.x, b.y);
        }
        return acc;
    }
    /* wide rows: repeated selection of the next smaller key */
    uint64_t prev = ~uint64_t(0);
    for (;;) {
        uint64_t best = 0;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint2 a = V.arcs[e];
            const uint32_t dw = dist_final(S, a.x);
            if ((uint64_t)dv + a.y != dw) continue;
            const uint64_t k = ((uint64_t)dw << 32) | a.x;
            if (k < prev && k > best) best = k;
        }
        if (best == 0) break;
        const ulonglong2 sw = sd_pair(S, (uint32_t)best);
        acc.add(sv_int, sv, sw.x, sw.y);
        prev = best;
    }
    return acc;
}

/* Phase 3: dependencies, the levels in reverse, then the mode's output in
 * one pass in vertex order: the levels visit vertices scattered over the
 * view, while the output pass reads the dependencies back with coalesced
 * loads and writes (or atomically adds) whole lines. */
template <int MODE, int PREC, bool ORDERED, class T>
__device__ __forceinline__ void phase_dependencies(const BrandesParams &p, const View &V, const Slot S, const T &t,
                                   uint32_t task, uint32_t c, uint32_t src, uint32_t nlev)
{
    const uint32_t r0 = t.rank(), R = t.size();
    const double wgt = MODE == MODE_PIVOT ? (double)p.task_aux[task] : 1.0;
    const bool packed = ald32<T::SC>(&t.ctl->wide) == 0;   /* which sigma form ran */
    for (int k = (int)nlev - 1; k >= 0; --k) {
        const uint32_t i0 = S.fa[k], i1 = S.fa[k + 1];
        for (uint32_t i = i0 + r0; i < i1; i += R) {
            const uint32_t v = S.qa[i];
            if (v == src) continue;
            const Acc<PREC> acc = fold_successors<PREC, ORDERED, T::RB, T::SUCC, T::MB>(V, S, v, packed);
            S.sd[2 * (uint64_t)v + 1] = acc.bits();
        }
        t.sync();
        if (t.aborted()) return;
    }
    /* the source has no dependency (its delta slot holds its mask); the sigma
     * sweep left an unreached vertex's slot 0, a zero dependency in either
     * precision.  OU vertices per thread have their loads in flight
     * together */
    constexpr int OU = 4;
    for (uint32_t v0 = r0; v0 < V.n; v0 += OU * R) {
        uint64_t b[OU];
#pragma unroll
        for (int k = 0; k < OU; ++k) {
            const uint32_t v = v0 + k * R;
            b[k] = S.sd[2 * (uint64_t)(v < V.n ? v : V.n - 1) + 1];   /* unconditional: in flight together */
        }
#pragma unroll
        for (int k = 0; k < OU; ++k) {
            const uint32_t v = v0 + k * R;
            if (v >= V.n) break;
            const double dv = v != src ? Acc<PREC>::value_of(b[k]) : 0.0;
            if (MODE == MODE_DEP) {
                p.out_dep[(uint64_t)task * V.n + v] = dv;
            } else if (MODE == MODE_EXACT) {
                if (dv != 0.0) unsafeAtomicAdd(&p.bc[v], dv);
            } else if (MODE == MODE_PIVOT) {
                const double x = dv * wgt;
                if (x != 0.0) unsafeAtomicAdd(&p.bc[v], x);
            } else if (MODE == MODE_CLUSTER) {
                p.dmat[p.cl_doff[c] + (uint64_t)src * V.n + v] = dv;
            }
        }
    }
}

template <class T>
__device__ __forceinline__ void flush_counters(const BrandesParams &p, const T &t, const Counters &ct)
{
    t.lsync();
    if (t.lleader() && t.loc->te) atomicAdd(p.te, t.loc->te);
    if (p.stats)
        for (int i = 0; i < STAT_COUNT; ++i)
            if (ct.c[i]) atomicAdd(&p.stats[i], ct.c[i]);
}

/* One (view, source) task of the batched kernels, run by team t on slot S;
 * q is its queue index (which trial set it belongs to in a trial launch). */
template <int MODE, int PREC, bool ORDERED, class T>
__device__ __forceinline__ void run_task(const BrandesParams &p, const Slot S, const T &t, uint32_t task,
                                         uint32_t q, Counters &ct, const uint4 *ell_view = nullptr)
{
    View V;
    uint32_t c = 0;
    if (MODE == MODE_CLUSTER) {
        c = p.task_aux[task];
        const uint64_t ro = p.cl_rpoff[c];
        V.n = p.cl_n[c];
        V.rp = p.rp + ro;
        V.rrp = p.rrp + ro;
        /* the cluster's short rows: staged in LDS by the LDS-view kernel */
        V.ell = ell_view ? ell_view : p.ell ? p.ell + p.cl_voff[c] : nullptr;
    } else {
        V.n = p.n;
        V.rp = p.rp;
        V.rrp = p.rrp;
        V.ell = p.ell;
    }
    V.lrows = ell_view;
    V.arcs = p.arcs;
    V.rarcs = p.rarcs;
    const uint32_t src = p.task_src[task];
    const uint32_t r0 = t.rank(), R = t.size();
    if (src >= V.n) {   /* the same for the whole team: it skips the task together */
        if (t.leader()) atomicOr(p.err, ERR_SOURCE);
        return;
    }

    const uint32_t set = p.trial_sets ? q % p.trial_sets : 0u;
    const uint32_t delta = p.trial_sets ? p.trial_delta[set] : p.delta;
    const unsigned long long te0 = t.loc->te;   /* stable: the last task ended with barriers */
    const uint64_t t0 = wall_clock64();
    phase_distances(V, S, t, src, delta, p.sort_cap, p.err, ct);
    const uint64_t t1 = wall_clock64();
    const uint32_t nlev = phase_sigma(V, S, t, src, p.level_sort, p.err, ct);
    const uint64_t t2 = wall_clock64();
    if (t.leader()) {
        ct.c[STAT_T_DIST] += t1 - t0;
        ct.c[STAT_T_SIGMA] += t2 - t1;
    }

    if (MODE == MODE_SSSP) {
        uint64_t *od = p.out_dist + (uint64_t)task * V.n;
        uint64_t *os = p.out_sigma + (uint64_t)task * V.n;
        for (uint32_t v = r0; v < V.n; v += R) {
            const uint32_t d = dist_final(S, v);
            od[v] = d == DINF ? ~uint64_t(0) : d;
            os[v] = S.sd[2 * (uint64_t)v];
        }
        t.sync();
        return;
    }
    if (MODE == MODE_CLUSTER) {
        const uint32_t b0 = p.cl_bptr[c], nb = p.cl_bptr[c + 1] - b0;
        const uint64_t row = p.cl_vioff[c] + (uint64_t)src * nb;
        for (uint32_t j = r0; j < nb; j += R) {
            const uint32_t b = p.cl_blocal[b0 + j];
            const uint32_t d = dist_final(S, b);
            p.vi_len[row + j] = d == DINF ? 0 : d;
            /* the reference keeps (int)sigma: INT_MIN beyond the int range */
            const uint64_t sg = S.sd[2 * (uint64_t)b];
            p.vi_cnt[row + j] = sg > 0x7FFFFFFFull ? (int32_t)0x80000000u : (int32_t)sg;
        }
    }
    phase_dependencies<MODE, PREC, ORDERED>(p, V, S, t, task, c, src, nlev);
    if (t.leader()) {
        const uint64_t t3 = wall_clock64();
        ct.c[STAT_T_DEP] += t3 - t2;
        if (p.trial_sets) {   /* te is final: the sigma sweep's adds ended before its barrier */
            atomicAdd(&p.trial_stats[2 * set], t.loc->te - te0);
            atomicAdd(&p.trial_stats[2 * set + 1], (unsigned long long)(t3 - t0));
        }
    }
}

/* ---- the batched kernels: a device work queue of (view, source) tasks ------ */

/* One workgroup per source. */
template <int MODE, int PREC, bool ORDERED>
__global__ __launch_bounds__(WG) void k_brandes(BrandesParams p)
{
    __shared__ Shared sh;
    Slot S;
    slot_layout(p.nmax, p.scratch + (uint64_t)blockIdx.x * p.slot_bytes, &S);
    S.agg = p.ballot_append != 0;
    const TeamWG t{&sh.ctl, &sh.loc, sh.sortbuf, sh.lq};
    if (threadIdx.x == 0) sh.loc.te = 0;
    Counters ct;
    for (;;) {
        if (threadIdx.x == 0) sh.task = atomicAdd(p.queue_head, 1u);
        __syncthreads();
        const uint32_t task = sh.task;
        __syncthreads();
        if (task >= p.ntasks) break;
        run_task<MODE, PREC, ORDERED>(p, S, t, p.task_base + task, task, ct);
    }
    flush_counters(p, t, ct);
}

/* One wavefront per source: WAVES slots per workgroup, each wavefront pulls
 * its own tasks.  Four wavefronts per SIMD at least (128 VGPRs): these
 * launches are latency-bound, and the ordered cluster form would otherwise
 * take 130 and run three. */
template <int MODE, int PREC, bool ORDERED>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_brandes_wave(BrandesParams p)
{
    __shared__ SharedWave sw;
    const uint32_t w = threadIdx.x / 64;
    if (blockIdx.x * WAVES + w >= p.nslots) return;   /* no slot behind this wavefront */
    Slot S;
    slot_layout(p.nmax, p.scratch + ((uint64_t)blockIdx.x * WAVES + w) * p.slot_bytes, &S);
    S.agg = p.ballot_append != 0;
    const TeamWave t{&sw.ctl[w], &sw.loc[w], sw.lq[w]};
    if (t.lleader()) sw.loc[w].te = 0;
    Counters ct;
    for (;;) {
        uint32_t task = 0;
        if (t.lleader()) task = atomicAdd(p.queue_head, 1u);
        task = __shfl(task, 0, 64);
        if (task >= p.ntasks) break;
        run_task<MODE, PREC, ORDERED>(p, S, t, p.task_base + task, task, ct);
    }
    flush_counters(p, t, ct);
}

/* One half wavefront per source: HALVES slots per workgroup; each half pulls
 * its own tasks and broadcasts them from its first lane. */
template <int MODE, int PREC, bool ORDERED>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4))) void k_brandes_half(BrandesParams p)
{
    __shared__ SharedHalf sh;
    const uint32_t h = threadIdx.x / 32;
    if (blockIdx.x * HALVES + h >= p.nslots) return;   /* no slot behind this half */
    Slot S;
    slot_layout(p.nmax, p.scratch + ((uint64_t)blockIdx.x * HALVES + h) * p.slot_bytes, &S);
    S.agg = p.ballot_append != 0;
    const TeamHalf t{&sh.ctl[h], &sh.loc[h], sh.lq[h]};
    if (t.lleader()) sh.loc[h].te = 0;
    Counters ct;
    for (;;) {
        uint32_t task = 0;
        if (t.lleader()) task = atomicAdd(p.queue_head, 1u);
        task = __shfl(task, (int)(threadIdx.x & 32), 64);
        if (task >= p.ntasks) break;
        run_task<MODE, PREC, ORDERED>(p, S, t, p.task_base + task, task, ct);
    }
    flush_counters(p, t, ct);
}

/* ---- the LDS-view kernel: cluster views whose short rows sit in LDS ------- */

/* Per-workgroup LDS of the LDS-view kernel besides the rows: NT teams of
 * LQN-entry queue heads; the workgroup's current run of tasks (end in the
 * high 32 bits, next in the low 32; end RUN_BUSY while a team claims the next
 * run); the staged view's state word: the staged cluster (high 32 bits;
 * VIEW_NONE while empty or being staged) and the number of teams using it
 * (low 32 bits). */
template <int NT, uint32_t LQN>
struct SharedLds {
    Ctl ctl[NT];
    Local loc[NT];
    uint32_t lq[NT][2 * LQN];
    unsigned long long run;
    unsigned long long view;
};
constexpr uint32_t VIEW_NONE = 0xFFFFFFFFu;
constexpr uint32_t RUN_BUSY = 0xFFFFFFFFu;

/* The next task of the workgroup's run (team leader only; ntasks or more
 * when the queue is empty).  A team that finds the run used up claims the
 * next one from the device queue: a quarter of the remaining tasks per
 * workgroup, at most run_max and at least one.  A team that finds another
 * claiming takes a single task from the device queue instead of waiting:
 * the two half-wavefront teams of a wavefront run in lockstep, so one must
 * never wait for the other. */
__device__ __forceinline__ uint32_t run_next(unsigned long long *run, uint32_t *queue_head, uint32_t ntasks,
                                             uint32_t run_max)
{
    unsigned long long s = __hip_atomic_load(run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
        const uint32_t next = (uint32_t)s, end = (uint32_t)(s >> 32);
        if (end == RUN_BUSY) return atomicAdd(queue_head, 1u);
        if (next < end) {
            if (__hip_atomic_compare_exchange_strong(run, &s, s + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP))
                return next;
            continue;
        }
        const unsigned long long busy = (unsigned long long)RUN_BUSY << 32;
        if (!__hip_atomic_compare_exchange_strong(run, &s, busy, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP))
            continue;
        const uint32_t h0 = __hip_atomic_load(queue_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t rem = h0 < ntasks ? ntasks - h0 : 0u;
        uint32_t len = rem / (4u * gridDim.x);
        len = len < 1u ? 1u : len > run_max ? run_max : len;
        const uint32_t h = atomicAdd(queue_head, len);
        const uint32_t e = h >= ntasks ? ntasks : ntasks - h < len ? ntasks : h + len;
        const uint32_t b = h >= ntasks ? ntasks : h;
        /* this team takes b; the run holds [b + 1, e) */
        __hip_atomic_store(run, ((unsigned long long)e << 32) | (b < e ? b + 1 : e), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
        return b;
    }
}
/* queue heads of an LDS-view team: wavefronts 128 entries, half wavefronts 64 */
constexpr uint32_t LQ_LDS_WAVE = 128, LQ_LDS_HALF = 64;
/* longest run of tasks a workgroup claims at once */
constexpr uint32_t LDS_RUN_MAX = 256;

/* The team leader's claim on cluster c's rows in LDS: true when they are
 * staged and its use is counted; when no team uses the staged view (or none
 * is staged) it takes the view over and returns *stage = true, the view
 * marked as being staged.  Otherwise false: the team reads its rows from
 * HBM. */
__device__ __forceinline__ bool view_acquire(unsigned long long *view, uint32_t c, bool *stage)
{
    *stage = false;
    unsigned long long s = __hip_atomic_load(view, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
        const uint32_t vc = (uint32_t)(s >> 32), users = (uint32_t)s;
        unsigned long long want;
        if (vc == c)
            want = s + 1;
        else if (users == 0)
            want = ((unsigned long long)VIEW_NONE << 32) | 1ull;   /* staging: one user, the stager */
        else
            return false;
        if (__hip_atomic_compare_exchange_strong(view, &s, want, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP)) {
            *stage = vc != c;
            return true;
        }
    }
}

/* Cluster evaluation with the cluster's short rows in LDS (dynamic shared
 * memory, p.lds_rows_max rows).  One workgroup of LWG threads per CU holds
 * NT teams (wavefronts, or half wavefronts when HALF) -- the occupancy of the
 * other team kernels -- and one staged view that all its teams may share.
 * Tasks go out cluster by cluster; a workgroup claims runs of consecutive
 * tasks (run_next), so its teams work on one cluster for a while, and each
 * team takes the run's next task when it finishes one.  A task whose
 * cluster is staged reads its rows (distance phase, Kahn step, ordered fold)
 * from LDS; when no team of the workgroup uses the staged view, the next
 * team to take a task of a cluster that fits restages the view to that
 * cluster; tasks that find the view staged to another cluster in use, or
 * being staged, read their rows from HBM.  No team waits for another's
 * task. */
template <int MODE, int PREC, bool ORDERED, bool HALF>
__global__ __launch_bounds__(LWG) void k_brandes_lds(BrandesParams p)
{
    using T = typename std::conditional<HALF, TeamHalfT<LQ_LDS_HALF>, TeamWaveT<LQ_LDS_WAVE>>::type;
    constexpr int NT = LWG / (int)T::LANES;
    static_assert(NT == (HALF ? LDS_TEAMS_HALF : LDS_TEAMS_WAVE), "teams per LDS-view workgroup");
    extern __shared__ uint4 lds_rows[];
    __shared__ SharedLds<NT, T::LQ> sh;
    if (threadIdx.x == 0) {
        sh.view = (unsigned long long)VIEW_NONE << 32;
        sh.run = 0;   /* used up: the first team claims a run */
    }
    __syncthreads();
    const uint32_t tm = threadIdx.x / T::LANES;
    if (blockIdx.x * (uint32_t)NT + tm >= p.nslots) return;   /* no slot behind this team */
    Slot S;
    slot_layout(p.nmax, p.scratch + ((uint64_t)blockIdx.x * NT + tm) * p.slot_bytes, &S);
    S.agg = p.ballot_append != 0;
    const T t{&sh.ctl[tm], &sh.loc[tm], sh.lq[tm]};
    if (t.lleader()) sh.loc[tm].te = 0;
    const int lead = (int)(threadIdx.x & (64u - T::LANES));   /* the team's first lane in its wavefront */
    Counters ct;
    for (;;) {
        uint32_t task = 0;
        if (t.lleader()) task = run_next(&sh.run, p.queue_head, p.ntasks, LDS_RUN_MAX);
        task = __shfl(task, lead, 64);
        if (task >= p.ntasks) break;
        const uint32_t c = p.task_aux[p.task_base + task];
        const uint32_t nc = p.cl_n[c];
        uint32_t mode = 0;   /* 0: HBM rows, 1: staged rows, 2: stage them first */
        if (nc <= p.lds_rows_max && t.lleader()) {
            bool stage = false;
            if (view_acquire(&sh.view, c, &stage)) mode = stage ? 2u : 1u;
        }
        mode = __shfl(mode, lead, 64);
        if (mode == 2) {
            const uint4 *src = p.ell + p.cl_voff[c];
            for (uint32_t i = t.lrank(); i < nc; i += T::LANES) lds_rows[i] = src[i];
            t.lsync();
            if (t.lleader())
                __hip_atomic_store(&sh.view, ((unsigned long long)c << 32) | 1ull, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        run_task<MODE, PREC, ORDERED>(p, S, t, p.task_base + task, task, ct, mode ? lds_rows : nullptr);
        if (mode) {
            t.lsync();   /* every lane is done with the rows */
            if (t.lleader())
                __hip_atomic_fetch_sub(&sh.view, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    flush_counters(p, t, ct);
}

/* ---- the plain kernel: one source per cooperative launch over the grid ---- */

template <int MODE>
__global__ __launch_bounds__(WG) void k_plain(BrandesParams p, uint32_t task)
{
    __shared__ Shared sh;
    Slot S;
    slot_layout(p.nmax, p.scratch, &S);
    S.agg = p.ballot_append != 0;
    const TeamGrid t{reinterpret_cast<Ctl *>(p.grid_ctl), &sh.loc, nullptr};
    if (threadIdx.x == 0) sh.loc.te = 0;
    Counters ct;
    View V;
    V.n = p.n;
    V.rp = p.rp;
    V.rrp = p.rrp;
    V.ell = p.ell;
    V.lrows = nullptr;
    V.arcs = p.arcs;
    V.rarcs = p.rarcs;
    const uint32_t src = p.task_src[task];
    const uint64_t t0 = wall_clock64();
    phase_distances(V, S, t, src, p.delta, 0, p.err, ct);
    const uint64_t t1 = wall_clock64();
    const uint32_t nlev = phase_sigma(V, S, t, src, 0, p.err, ct);
    const uint64_t t2 = wall_clock64();
    phase_dependencies<MODE, 0, true>(p, V, S, t, task, 0, src, nlev);
    if (t.leader()) {
        ct.c[STAT_T_DIST] += t1 - t0;
        ct.c[STAT_T_SIGMA] += t2 - t1;
        ct.c[STAT_T_DEP] += wall_clock64() - t2;
    }
    flush_counters(p, t, ct);
}

/* Largest dynamic LDS of an LDS-view workgroup: a workgroup's LDS limit
 * (the whole 160 KB of a gfx950 CU) less the kernel's static part. */
template <int MODE, int PREC, bool ORDERED, bool HALF>
static hipError_t lds_capacity(uint32_t *bytes)
{
    const void *f = reinterpret_cast<const void *>(&k_brandes_lds<MODE, PREC, ORDERED, HALF>);
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, f);
    if (e != hipSuccess) return e;
    int dev = 0, lds = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
        return e;
    *bytes = lds > (int)a.sharedSizeBytes ? (uint32_t)(lds - (int)a.sharedSizeBytes) : 0u;
    return hipSuccess;
}

template <int MODE, int PREC, bool ORDERED, bool HALF>
static hipError_t launch_lds(const BrandesParams &p, uint32_t slots, hipStream_t s)
{
    const void *f = reinterpret_cast<const void *>(&k_brandes_lds<MODE, PREC, ORDERED, HALF>);
    constexpr uint32_t NT = HALF ? LDS_TEAMS_HALF : LDS_TEAMS_WAVE;
    const uint32_t bytes = p.lds_rows_max * (uint32_t)sizeof(uint4);
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_brandes_lds<MODE, PREC, ORDERED, HALF>), dim3((slots + NT - 1) / NT), dim3(LWG), bytes, s,
                       p);
    return hipGetLastError();
}

template <int MODE, int PREC, bool ORDERED>
static hipError_t launch_one(const BrandesParams &p, uint32_t slots, hipStream_t s)
{
    const int kind = team_kind(p.nmax, p.half_max);
    if constexpr (MODE == MODE_CLUSTER) {
        if (p.lds_rows_max)
            return kind == TEAM_HALF ? launch_lds<MODE, PREC, ORDERED, true>(p, slots, s)
                                     : launch_lds<MODE, PREC, ORDERED, false>(p, slots, s);
    }
    if (kind == TEAM_HALF)
        hipLaunchKernelGGL((k_brandes_half<MODE, PREC, ORDERED>), dim3((slots + HALVES - 1) / HALVES), dim3(WG), 0, s,
                           p);
    else if (kind == TEAM_WAVE)
        hipLaunchKernelGGL((k_brandes_wave<MODE, PREC, ORDERED>), dim3((slots + WAVES - 1) / WAVES), dim3(WG), 0, s, p);
    else
        hipLaunchKernelGGL((k_brandes<MODE, PREC, ORDERED>), dim3(slots), dim3(WG), 0, s, p);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mode(int prec, bool ordered, const BrandesParams &p, uint32_t slots,
                              hipStream_t s)
{
    if (prec == 0)
        return ordered ? launch_one<MODE, 0, true>(p, slots, s) : launch_one<MODE, 0, false>(p, slots, s);
    return ordered ? launch_one<MODE, 1, true>(p, slots, s) : launch_one<MODE, 1, false>(p, slots, s);
}

hipError_t launch_brandes(int mode, int prec, bool ordered, const BrandesParams &p,
                          uint32_t slots, hipStream_t stream)
{
    switch (mode) {
    case MODE_SSSP: return launch_one<MODE_SSSP, 0, false>(p, slots, stream);
    case MODE_DEP: return launch_mode<MODE_DEP>(prec, ordered, p, slots, stream);
    case MODE_EXACT: return launch_mode<MODE_EXACT>(prec, ordered, p, slots, stream);
    case MODE_PIVOT: return launch_mode<MODE_PIVOT>(prec, ordered, p, slots, stream);
    case MODE_CLUSTER: return launch_one<MODE_CLUSTER, 0, true>(p, slots, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t lds_view_capacity(bool half, uint32_t *bytes)
{
    return half ? lds_capacity<MODE_CLUSTER, 0, true, true>(bytes) : lds_capacity<MODE_CLUSTER, 0, true, false>(bytes);
}

static const void *plain_kernel(int mode)
{
    switch (mode) {
    case MODE_DEP: return reinterpret_cast<const void *>(&k_plain<MODE_DEP>);
    case MODE_PIVOT: return reinterpret_cast<const void *>(&k_plain<MODE_PIVOT>);
    default: return nullptr;
    }
}

hipError_t plain_max_blocks(int mode, int cu_count, uint32_t *blocks)
{
    const void *f = plain_kernel(mode);
    if (!f) return hipErrorInvalidValue;
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, WG, 0);
    if (e != hipSuccess) return e;
    *blocks = (uint32_t)(per_cu * cu_count);
    return hipSuccess;
}

hipError_t launch_plain(int mode, const BrandesParams &p, uint32_t task, uint32_t blocks,
                        hipStream_t stream)
{
    const void *f = plain_kernel(mode);
    if (!f) return hipErrorInvalidValue;
    BrandesParams pc = p;
    uint32_t tk = task;
    void *args[] = {&pc, &tk};
    return hipLaunchCooperativeKernel(f, dim3(blocks), dim3(WG), args, 0, stream);
}

/* ---- clearing control words (a kernel, not a memset node) ------------------ */

__global__ void k_clear_words(uint32_t *p, uint32_t nwords)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) p[i] = 0u;
}

hipError_t launch_clear_words(uint32_t *p, uint32_t nwords, hipStream_t stream)
{
    if (!nwords) return hipSuccess;
    hipLaunchKernelGGL(k_clear_words, dim3((nwords + 255) / 256), dim3(256), 0, stream, p, nwords);
    return hipGetLastError();
}

/* ---- cluster reduction and pivot subtraction ------------------------------- */

__global__ void k_intra_reduce(BrandesParams p, uint64_t vbegin, uint64_t vcount,
                               const uint32_t *v_cluster, double *intra)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vcount) return;
    const uint64_t vi = vbegin + i;               /* index into the concatenated vertex blocks */
    const uint32_t c = v_cluster[vi];
    if (p.cl_doff[c] == ~uint64_t(0)) return;
    const uint32_t nc = p.cl_n[c];
    const uint32_t v = (uint32_t)(vi - p.cl_voff[c]);
    const double *col = p.dmat + p.cl_doff[c] + v;
    double acc = 0.0;
    for (uint32_t s = 0; s < nc; ++s) acc += col[(uint64_t)s * nc];
    const uint32_t g = p.cl_gid[vi];
    intra[g] = intra[g] + acc;
}

hipError_t launch_intra_reduce(const BrandesParams &p, uint64_t vbegin, uint64_t vcount,
                               const uint32_t *v_cluster, double *intra, hipStream_t stream)
{
    if (!vcount) return hipSuccess;
    const uint32_t blocks = (uint32_t)((vcount + 255) / 256);
    hipLaunchKernelGGL(k_intra_reduce, dim3(blocks), dim3(256), 0, stream, p, vbegin, vcount,
                       v_cluster, intra);
    return hipGetLastError();
}

__global__ void k_pivot_subtract(uint64_t nv, const uint32_t *v_gid, const uint32_t *v_cluster,
                                 const uint32_t *cpiv_ptr, const uint32_t *piv_w,
                                 const double *intra, double *bc)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const uint32_t g = v_gid[i];
    const uint32_t c = v_cluster[i];
    const double x = intra[g];
    double b = bc[g];
    for (uint32_t k = cpiv_ptr[c]; k < cpiv_ptr[c + 1]; ++k) b -= x * (double)piv_w[k];
    bc[g] = b;
}

__global__ void k_gather_add(uint32_t n, const uint32_t *perm, const double *src, double *dst)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) dst[v] += src[perm[v]];
}

hipError_t launch_gather_add(uint32_t n, const uint32_t *perm, const double *src, double *dst, hipStream_t stream)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_gather_add, dim3((n + 255) / 256), dim3(256), 0, stream, n, perm, src, dst);
    return hipGetLastError();
}

hipError_t launch_pivot_subtract(uint64_t nv, const uint32_t *v_gid, const uint32_t *v_cluster,
                                 const uint32_t *cpiv_ptr, const uint32_t *piv_w,
                                 const double *intra, double *bc, hipStream_t stream)
{
    if (!nv) return hipSuccess;
    const uint32_t blocks = (uint32_t)((nv + 255) / 256);
    hipLaunchKernelGGL(k_pivot_subtract, dim3(blocks), dim3(256), 0, stream, nv, v_gid, v_cluster,
                       cpiv_ptr, piv_w, intra, bc);
    return hipGetLastError();
}

} // namespace fbcdev
