/*
 * Persistent single-source Brandes kernel for gfx950.
 *
 * One workgroup (4 wavefronts) owns one source at a time and pulls the next
 * (view, source) task from a device work queue, so every source of a run --
 * all vertices (--exact), all cluster vertices, or all pivots -- goes through
 * one launch.  Per source:
 *
 *   1. distances: near-far delta-stepping.  Relaxations use 32-bit atomicMin
 *      on the tentative distance; vertices whose distance improves below the
 *      window limit are deduplicated per round with an atomicMax round stamp
 *      and appended to the next near queue, the others to the far queue.
 *   2. sigma: a Kahn sweep over the shortest-path DAG (arc u->w with
 *      d(u) + w = d(w)).  In-degrees are pulled over the reverse CSR; sigma is
 *      pushed with 64-bit integer atomics (exact, overflow checked) and every
 *      vertex is appended to the level list when its last DAG predecessor is
 *      done.  The level list is a topological order of the DAG.
 *   3. dependencies: the levels in reverse; every vertex pulls
 *      sigma(v) / sigma(w) * (1 + delta(w)) from its DAG successors.  In
 *      'ordered' mode the successors are summed in descending (d(w), w), the
 *      order in which the reference's backward stack visits them
 *      (DijkstraSSBrandesBC.h:61-78), so the FP64 result equals the
 *      reference's bit for bit.
 *
 * Views are CSRs in HBM (the whole graph, or one cluster's local CSR whose
 * local ids follow ascending global ids).  Per-slot state (distance, stamps,
 * queues, sigma, delta) lives in a private scratch region of the slot.
 */
#include "fbc_internal.h"

namespace fbcdev {

constexpr int HEAVY_DEG = 96;   /* out-degree handled by the whole workgroup */
constexpr int HEAVY_CAP = 64;
constexpr int SUCC_REG = 8;     /* successors kept in registers by the ordered fold */

/* ---- memory helpers ------------------------------------------------------ */

/* Per-slot state is touched by one workgroup only, so its atomics and the
 * loads of words other threads update atomically use workgroup scope: they
 * stay in the XCD's L2 instead of going to the device-coherent memory side.
 * Only the shared outputs (bc) and the task queue use device scope. */
__device__ __forceinline__ uint32_t ld_a32(const uint32_t *p)
{
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t ld_a64(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t wg_min(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t wg_max(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t wg_sub(uint32_t *p, uint32_t v)
{
    return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t wg_add64(uint64_t *p, uint64_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__host__ __device__ inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

struct Slot {
    uint32_t *dist, *stamp, *fstamp, *qa, *qb, *fa, *fb;
    uint64_t *sigma;
    uint64_t *delta;   /* double, or a float pair, bit-cast */
};

__host__ __device__ inline uint64_t slot_layout(uint32_t nmax, uint8_t *base, Slot *s)
{
    const uint64_t w32 = align256((uint64_t)(nmax + 1) * 4);
    const uint64_t w64 = align256((uint64_t)(nmax + 1) * 8);
    if (s) {
        s->dist = (uint32_t *)(base);
        s->stamp = (uint32_t *)(base + w32);
        s->fstamp = (uint32_t *)(base + 2 * w32);
        s->qa = (uint32_t *)(base + 3 * w32);
        s->qb = (uint32_t *)(base + 4 * w32);
        s->fa = (uint32_t *)(base + 5 * w32);
        s->fb = (uint32_t *)(base + 6 * w32);
        s->sigma = (uint64_t *)(base + 7 * w32);
        s->delta = (uint64_t *)(base + 7 * w32 + w64);
    }
    return 7 * w32 + 2 * w64;
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
__device__ __forceinline__ F2 f2_from_u64(uint64_t v)
{
    float hi = (float)v;
    double rem = (double)v - (double)hi;   /* exact below 2^53, close above */
    return {hi, (float)rem};
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
    uint32_t n;
};

/* Per-thread work counters, flushed into BrandesParams::stats. */
struct Counters {
    unsigned long long c[STAT_COUNT] = {};
};

struct Shared {
    uint32_t task;
    uint32_t qcnt[2];
    uint32_t fcnt, fcnt2;
    uint32_t minv;
    uint32_t tail;
    uint32_t nheavy;
    uint32_t heavy[HEAVY_CAP];
    unsigned long long te;
};

/* Relax arc e of u (distance du).  Pushes into the next near queue or the far
 * list. */
__device__ __forceinline__ void relax_arc(const View &V, const Slot &S, Shared &sh,
                                          uint32_t du, uint32_t e, uint32_t thr,
                                          uint32_t nxt, uint32_t win, uint32_t *qnext,
                                          uint32_t *far, uint32_t *err, Counters &ct)
{
    ++ct.c[STAT_RELAX];
    const uint2 a = V.arcs[e];
    const uint64_t nd64 = (uint64_t)du + a.y;
    if (nd64 >= DINF) {
        atomicOr(err, ERR_DIST);
        return;
    }
    const uint32_t nd = (uint32_t)nd64;
    if (nd >= ld_a32(&S.dist[a.x])) return;
    const uint32_t old = wg_min(&S.dist[a.x], nd);
    if (nd >= old) return;
    ++ct.c[STAT_IMPROVE];
    if (nd < thr) {
        if (wg_max(&S.stamp[a.x], nxt) < nxt) qnext[atomicAdd(&sh.qcnt[nxt & 1], 1u)] = a.x;
    } else {
        if (wg_max(&S.fstamp[a.x], win) < win) {
            far[atomicAdd(&sh.fcnt, 1u)] = a.x;
            ++ct.c[STAT_FARPUSH];
        }
    }
}

/* Phase 1: near-far delta-stepping.  On return dist[] is final. */
__device__ void phase_distances(const View &V, const Slot &S, Shared &sh, uint32_t src,
                                uint32_t delta, uint32_t *err, Counters &ct)
{
    const int tid = threadIdx.x;
    for (uint32_t v = tid; v < V.n; v += WG) {
        S.dist[v] = DINF;
        S.stamp[v] = 0;
        S.fstamp[v] = 0;
        S.sigma[v] = 0;
    }
    if (tid == 0) {
        sh.qcnt[0] = sh.qcnt[1] = 0;
        sh.fcnt = sh.fcnt2 = 0;
        sh.nheavy = 0;
    }
    __syncthreads();
    if (tid == 0) {
        S.dist[src] = 0;
        S.stamp[src] = 1;
        S.qa[0] = src;
    }
    __syncthreads();

    uint32_t *qa = S.qa, *qb = S.qb, *fa = S.fa, *fb = S.fb;
    uint32_t near = 1, rnd = 1, win = 1;
    uint32_t thr = delta;
    for (;;) {
        while (near) {
            const uint32_t nxt = rnd + 1;
            if (tid == 0) ++ct.c[STAT_ROUNDS];
            for (uint32_t i = tid; i < near; i += WG) {
                const uint32_t u = qa[i];
                ++ct.c[STAT_VISITS];
                const uint32_t du = ld_a32(&S.dist[u]);
                const uint32_t e0 = V.rp[u], e1 = V.rp[u + 1];
                if (e1 - e0 > (uint32_t)HEAVY_DEG) {
                    uint32_t h = atomicAdd(&sh.nheavy, 1u);
                    if (h < (uint32_t)HEAVY_CAP) {
                        sh.heavy[h] = u;
                        continue;
                    }
                }
                for (uint32_t e = e0; e < e1; ++e)
                    relax_arc(V, S, sh, du, e, thr, nxt, win, qb, fa, err, ct);
            }
            __syncthreads();
            const uint32_t nh = sh.nheavy < (uint32_t)HEAVY_CAP ? sh.nheavy : (uint32_t)HEAVY_CAP;
            if (nh) {
                for (uint32_t h = 0; h < nh; ++h) {
                    const uint32_t u = sh.heavy[h];
                    const uint32_t du = ld_a32(&S.dist[u]);
                    const uint32_t e1 = V.rp[u + 1];
                    for (uint32_t e = V.rp[u] + tid; e < e1; e += WG)
                        relax_arc(V, S, sh, du, e, thr, nxt, win, qb, fa, err, ct);
                }
                __syncthreads();
            }
            near = sh.qcnt[nxt & 1];
            uint32_t *t = qa; qa = qb; qb = t;
            rnd = nxt;
            __syncthreads();
            if (tid == 0) {
                sh.qcnt[(rnd + 1) & 1] = 0;
                sh.nheavy = 0;
            }
            __syncthreads();
        }
        /* window [.., thr) is settled; move on to the nearest far vertex */
        const uint32_t nfar = sh.fcnt;
        if (nfar == 0) break;
        if (tid == 0) ++ct.c[STAT_WINDOWS];
        if (tid == 0) sh.minv = DINF;
        __syncthreads();
        uint32_t mn = DINF;
        for (uint32_t i = tid; i < nfar; i += WG) {
            const uint32_t d = ld_a32(&S.dist[fa[i]]);
            if (d >= thr && d < mn) mn = d;
        }
        if (mn != DINF) atomicMin(&sh.minv, mn);
        __syncthreads();
        const uint32_t m = sh.minv;
        if (m == DINF) break;
        const uint64_t nthr64 = (uint64_t)m + delta;
        const uint32_t nthr = nthr64 >= DINF ? DINF : (uint32_t)nthr64;
        const uint32_t nwin = win + 1;
        for (uint32_t i = tid; i < nfar; i += WG) {
            const uint32_t w = fa[i];
            const uint32_t d = ld_a32(&S.dist[w]);
            if (d < thr) continue;            /* settled in an earlier window */
            if (d < nthr) {
                if (wg_max(&S.stamp[w], rnd) < rnd) qa[atomicAdd(&sh.qcnt[rnd & 1], 1u)] = w;
            } else {
                if (wg_max(&S.fstamp[w], nwin) < nwin) fb[atomicAdd(&sh.fcnt2, 1u)] = w;
            }
        }
        __syncthreads();
        near = sh.qcnt[rnd & 1];
        const uint32_t kept = sh.fcnt2;
        uint32_t *t = fa; fa = fb; fb = t;
        thr = nthr;
        win = nwin;
        __syncthreads();
        if (tid == 0) {
            sh.fcnt = kept;
            sh.fcnt2 = 0;
        }
        __syncthreads();
    }
}

/* Phase 2: DAG in-degrees, then the Kahn sweep with exact sigma.  Returns the
 * number of levels; the level list is S.qa, the level starts S.fa. */
__device__ uint32_t phase_sigma(const View &V, const Slot &S, Shared &sh, uint32_t src,
                                uint32_t *err, Counters &ct)
{
    const int tid = threadIdx.x;
    uint32_t *cnt = S.stamp;          /* stamps are no longer needed */
    unsigned long long te = 0;
    for (uint32_t v = tid; v < V.n; v += WG) {
        const uint32_t dv = ld_a32(&S.dist[v]);
        uint32_t c = 0;
        if (dv != DINF) {
            te += V.rp[v + 1] - V.rp[v];
            if (v != src) {
                const uint32_t r1 = V.rrp[v + 1];
                for (uint32_t r = V.rrp[v]; r < r1; ++r) {
                    const uint2 a = V.rarcs[r];
                    const uint32_t du = ld_a32(&S.dist[a.x]);
                    if (du != DINF && (uint64_t)du + a.y == dv) ++c;
                }
            }
        }
        cnt[v] = c;
    }
    if (te) atomicAdd(&sh.te, te);
    if (tid == 0) {
        S.sigma[src] = 1;
        S.qa[0] = src;
        S.fa[0] = 0;
        S.fa[1] = 1;
        sh.tail = 1;
        sh.nheavy = 0;
    }
    __syncthreads();
    uint32_t lb = 0, le = 1, nlev = 1;
    while (lb < le) {
        for (uint32_t i = lb + tid; i < le; i += WG) {
            const uint32_t u = S.qa[i];
            const uint32_t du = ld_a32(&S.dist[u]);
            const uint64_t su = ld_a64(&S.sigma[u]);
            const uint32_t e0 = V.rp[u], e1 = V.rp[u + 1];
            if (e1 - e0 > (uint32_t)HEAVY_DEG) {
                uint32_t h = atomicAdd(&sh.nheavy, 1u);
                if (h < (uint32_t)HEAVY_CAP) {
                    sh.heavy[h] = u;
                    continue;
                }
            }
            for (uint32_t e = e0; e < e1; ++e) {
                const uint2 a = V.arcs[e];
                if ((uint64_t)du + a.y != ld_a32(&S.dist[a.x])) continue;
                const uint64_t old = wg_add64(&S.sigma[a.x], su);
                if (old + su < old) atomicOr(err, ERR_SIGMA);
                if (wg_sub(&cnt[a.x], 1u) == 1u) S.qa[atomicAdd(&sh.tail, 1u)] = a.x;
            }
        }
        __syncthreads();
        const uint32_t nh = sh.nheavy < (uint32_t)HEAVY_CAP ? sh.nheavy : (uint32_t)HEAVY_CAP;
        if (nh) {
            for (uint32_t h = 0; h < nh; ++h) {
                const uint32_t u = sh.heavy[h];
                const uint32_t du = ld_a32(&S.dist[u]);
                const uint64_t su = ld_a64(&S.sigma[u]);
                const uint32_t e1 = V.rp[u + 1];
                for (uint32_t e = V.rp[u] + tid; e < e1; e += WG) {
                    const uint2 a = V.arcs[e];
                    if ((uint64_t)du + a.y != ld_a32(&S.dist[a.x])) continue;
                    const uint64_t old = wg_add64(&S.sigma[a.x], su);
                    if (old + su < old) atomicOr(err, ERR_SIGMA);
                    if (wg_sub(&cnt[a.x], 1u) == 1u) S.qa[atomicAdd(&sh.tail, 1u)] = a.x;
                }
            }
            __syncthreads();
        }
        lb = le;
        le = sh.tail;
        if (le > lb) {
            ++nlev;
            if (tid == 0) S.fa[nlev] = le;
        }
        __syncthreads();
        if (tid == 0) sh.nheavy = 0;
        __syncthreads();
    }
    if (tid == 0) ct.c[STAT_LEVELS] += nlev;
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
};

/* Dependency of v: fold over DAG successors.  'ordered' sums them in
 * descending (d(w), w). */
template <int PREC, bool ORDERED>
__device__ __forceinline__ Acc<PREC> fold_successors(const View &V, const Slot &S, uint32_t v)
{
    Acc<PREC> acc;
    const uint32_t dv = ld_a32(&S.dist[v]);
    const uint64_t sv_int = ld_a64(&S.sigma[v]);
    const double sv = (double)sv_int;
    const uint32_t e0 = V.rp[v], e1 = V.rp[v + 1];
    if (!ORDERED) {
        for (uint32_t e = e0; e < e1; ++e) {
            const uint2 a = V.arcs[e];
            if ((uint64_t)dv + a.y != ld_a32(&S.dist[a.x])) continue;
            acc.add(sv_int, sv, ld_a64(&S.sigma[a.x]), S.delta[a.x]);
        }
        return acc;
    }
    if (e1 - e0 <= (uint32_t)SUCC_REG) {
        uint64_t key[SUCC_REG];
        uint32_t who[SUCC_REG];
#pragma unroll
        for (int j = 0; j < SUCC_REG; ++j) {
            key[j] = 0;
            who[j] = 0;
            if (e0 + j < e1) {
                const uint2 a = V.arcs[e0 + j];
                const uint32_t dw = ld_a32(&S.dist[a.x]);
                if ((uint64_t)dv + a.y == dw) {
                    key[j] = ((uint64_t)dw << 32) | a.x;
                    who[j] = a.x;
                }
            }
        }
        /* keys are distinct and nonzero for successors (dw >= 1) */
#pragma unroll
        for (int r = 0; r < SUCC_REG; ++r) {
            uint64_t best = 0;
            uint32_t bw = 0;
            int bj = -1;
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j)
                if (key[j] > best) { best = key[j]; bw = who[j]; bj = j; }
            if (bj < 0) break;
#pragma unroll
            for (int j = 0; j < SUCC_REG; ++j)
                if (j == bj) key[j] = 0;
            acc.add(sv_int, sv, ld_a64(&S.sigma[bw]), S.delta[bw]);
        }
        return acc;
    }
    /* wide rows: repeated selection of the next smaller key */
    uint64_t prev = ~uint64_t(0);
    for (;;) {
        uint64_t best = 0;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint2 a = V.arcs[e];
            const uint32_t dw = ld_a32(&S.dist[a.x]);
            if ((uint64_t)dv + a.y != dw) continue;
            const uint64_t k = ((uint64_t)dw << 32) | a.x;
            if (k < prev && k > best) best = k;
        }
        if (best == 0) break;
        const uint32_t w = (uint32_t)best;
        acc.add(sv_int, sv, ld_a64(&S.sigma[w]), S.delta[w]);
        prev = best;
    }
    return acc;
}

/* ---- the kernel ------------------------------------------------------------ */

template <int MODE, int PREC, bool ORDERED>
__global__ __launch_bounds__(WG) void k_brandes(BrandesParams p)
{
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    Slot S;
    slot_layout(p.nmax, p.scratch + (uint64_t)blockIdx.x * p.slot_bytes, &S);
    if (tid == 0) sh.te = 0;
    Counters ct;

    for (;;) {
        if (tid == 0) sh.task = atomicAdd(p.queue_head, 1u);
        __syncthreads();
        const uint32_t task = sh.task;
        __syncthreads();
        if (task >= p.ntasks) break;

        View V;
        uint32_t c = 0;
        if (MODE == MODE_CLUSTER) {
            c = p.task_aux[task];
            const uint64_t ro = p.cl_rpoff[c];
            V.n = p.cl_n[c];
            V.rp = p.rp + ro;
            V.rrp = p.rrp + ro;
            V.arcs = p.arcs;
            V.rarcs = p.rarcs;
        } else {
            V.n = p.n;
            V.rp = p.rp;
            V.rrp = p.rrp;
            V.arcs = p.arcs;
            V.rarcs = p.rarcs;
        }
        const uint32_t src = p.task_src[task];

        phase_distances(V, S, sh, src, p.delta, p.err, ct);
        const uint32_t nlev = phase_sigma(V, S, sh, src, p.err, ct);

        if (MODE == MODE_SSSP) {
            uint64_t *od = p.out_dist + (uint64_t)task * V.n;
            uint64_t *os = p.out_sigma + (uint64_t)task * V.n;
            for (uint32_t v = tid; v < V.n; v += WG) {
                const uint32_t d = ld_a32(&S.dist[v]);
                od[v] = d == DINF ? ~uint64_t(0) : d;
                os[v] = ld_a64(&S.sigma[v]);
            }
            continue;
        }
        if (MODE == MODE_CLUSTER) {
            const uint32_t b0 = p.cl_bptr[c], nb = p.cl_bptr[c + 1] - b0;
            const uint64_t row = p.cl_vioff[c] + (uint64_t)src * nb;
            for (uint32_t j = tid; j < nb; j += WG) {
                const uint32_t b = p.cl_blocal[b0 + j];
                const uint32_t d = ld_a32(&S.dist[b]);
                p.vi_len[row + j] = d == DINF ? 0 : d;
                p.vi_cnt[row + j] = ld_a64(&S.sigma[b]);
            }
        }

        /* phase 3: dependencies, levels in reverse */
        const double wgt = MODE == MODE_PIVOT ? (double)p.task_aux[task] : 1.0;
        for (int k = (int)nlev - 1; k >= 0; --k) {
            const uint32_t i0 = S.fa[k], i1 = S.fa[k + 1];
            for (uint32_t i = i0 + tid; i < i1; i += WG) {
                const uint32_t v = S.qa[i];
                if (v == src) continue;
                const Acc<PREC> acc = fold_successors<PREC, ORDERED>(V, S, v);
                S.delta[v] = acc.bits();
                const double dv = acc.value();
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
            __syncthreads();
        }
    }
    __syncthreads();
    if (tid == 0 && sh.te) atomicAdd(p.te, sh.te);
    if (p.stats)
        for (int i = 0; i < STAT_COUNT; ++i)
            if (ct.c[i]) atomicAdd(&p.stats[i], ct.c[i]);
}

template <int MODE, int PREC, bool ORDERED>
static hipError_t launch_one(const BrandesParams &p, uint32_t slots, hipStream_t s)
{
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
