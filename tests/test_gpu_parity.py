"""Parity of the HIP path (called through the C ABI) against the CPU oracle.

Bars: distances and sigma bit-exact; FP64 dependencies in the reference's
order bit-exact; sums whose order differs from the reference's (OpenMP
reductions over sources) within 1e-9 relative; the FP32-compensated form
within 1e-6 relative of FP64.
"""
import os

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN, read_arcs

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)
U64MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def both_graphs(ctx, n, src, dst, w):
    g = F.Graph(ctx, n, *F.csr_from_arcs(n, src, dst, w))
    og = O.Graph.from_arcs(src, dst, np.asarray(w, np.float64))
    assert og.n == n
    return g, og


@pytest.fixture(scope="module")
def dwg9(gpu_ctx):
    s, d, w = read_arcs(os.path.join(GOLDEN, "dwg9.txt"))
    order = np.lexsort((d, s))
    return both_graphs(gpu_ctx, 9, s[order], d[order], w[order])


@pytest.fixture(scope="module")
def cfg0(gpu_ctx):
    n, (s, d, w) = F.generate_config(0)
    return both_graphs(gpu_ctx, n, s, d, w)


def assert_plain_ran(ctx, nsrc):
    """The last call ran every source on k_plain (cooperative
    launches), none on the batched kernel."""
    c = ctx.last_counters()
    assert c["plain_fallback_sources"] == 0, f"cooperative launch refused (HIP status {c['plain_refusal']})"
    assert c["plain_sources"] == nsrc, c


def check_sssp(g, og, sources):
    dist, sig = g.sssp_sigma(sources)
    for i, s in enumerate(sources):
        od, osg, _ = og.sssp(int(s))
        reached = np.isfinite(od)
        assert np.array_equal(dist[i][~reached], np.full((~reached).sum(), U64MAX))
        assert np.array_equal(dist[i][reached], od[reached].astype(np.uint64))
        assert np.array_equal(sig[i], osg.astype(np.uint64))


def test_dwg9_sssp_sigma_dep(dwg9):
    g, og = dwg9
    check_sssp(g, og, list(range(9)))
    for s in range(9):
        assert np.array_equal(g.single_source_dep(s), og.single_source_dep(s))


def test_dwg9_cluster_known_answer(dwg9):
    """test/brandes/DijkstraClusterEvaluator.cpp:38-74 on cluster {0..4}."""
    g, _ = dwg9
    cl = g.clusters(communities=[[0, 1, 2, 3, 4], [5, 6, 7, 8]])
    assert list(cl.borders(0)) == [3, 4]
    assert cl.edges(0) == 7
    bc, L, C = cl.evaluate()
    # cluster {5..8} is bordered too, so only vertices 0..4 carry the known answer
    assert list(bc[:5]) == [2.0, 2.0, 0.5, 1.0, 0.0]
    assert cl.vi_rows(L, 0).tolist() == [[5, 7], [1, 4], [8, 4], [0, 3], [0, 0]]
    assert cl.vi_rows(C, 0).tolist() == [[1, 2], [1, 1], [1, 1], [1, 1], [0, 1]]


def test_cfg0_sssp_sigma_bitexact(cfg0):
    g, og = cfg0
    rng = np.random.default_rng(0)
    check_sssp(g, og, rng.choice(g.n, 48, replace=False))


def test_cfg0_dep_bitexact(cfg0):
    g, og = cfg0
    srcs = np.random.default_rng(1).choice(g.n, 32, replace=False)
    dep = g.dep_batch(srcs)
    for i, s in enumerate(srcs):
        assert np.array_equal(dep[i], og.single_source_dep(int(s))), f"source {s}"


def test_cfg0_exact_bc(cfg0):
    g, og = cfg0
    bc = g.exact_bc()
    ref = og.exact_bc(THREADS)
    rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-12 * np.abs(ref).max())
    assert rel.max() <= 1e-9, rel.max()
    gold = os.path.join(GOLDEN, "cfg0_exact_bc.npy")
    if os.path.exists(gold):
        r2 = np.load(gold)
        assert (np.abs(bc - r2) / np.maximum(np.abs(r2), 1e-12 * np.abs(r2).max())).max() <= 1e-9


@pytest.mark.skip(reason="the grouping of rounds and levels is off by default: a build with 4,096 buckets and "
                         "2,048-entry chunks faulted the GPU on its first cfg2 pivot launch, cause not found; "
                         "the grouped path stays off the GPU until one is")
def test_sorted_frontier_bitexact(gpu_ctx):
    """Grouping the near queues and DAG levels by vertex id (workgroup teams:
    views of more than 65,536 vertices) changes only the order in which
    vertices are expanded: distances, sigma and FP64 dependencies stay
    exact with the grouping off, on (chunks of 4,096 entries)
    and in chunks of 100 (a round's or level's list grouped piece by piece,
    the LDS head and the HBM remainder of a near queue separately)."""
    n, (s, d, w) = F.generate_grid(300, 300, 1000, 7, road=True)
    assert n > 65536
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    srcs = np.random.default_rng(5).choice(n, 6, replace=False)
    ref = [og.single_source_dep(int(x)) for x in srcs]
    try:
        for chunk in (0, 100, 4096):
            gpu_ctx.set_option("frontier_sort", chunk)
            gpu_ctx.set_option("level_sort", chunk)
            check_sssp(g, og, srcs[:3])
            dep = g.dep_batch(srcs)
            for i, x in enumerate(srcs):
                assert np.array_equal(dep[i], ref[i]), f"chunk {chunk}, source {x}"
    finally:
        gpu_ctx.set_option("frontier_sort", 0)
        gpu_ctx.set_option("level_sort", 0)


def test_cfg0_dep_f32c(cfg0):
    g, _ = cfg0
    srcs = np.arange(0, g.n, 257)
    d64 = g.dep_batch(srcs, F.PREC_F64)
    d32 = g.dep_batch(srcs, F.PREC_F32C)
    rel = np.abs(d32 - d64) / np.maximum(np.abs(d64), 1e-300)
    rel[d64 == 0] = np.abs(d32[d64 == 0])
    assert rel.max() <= 1e-6, rel.max()


def test_cfg0_cluster_eval_bitexact(cfg0):
    g, og = cfg0
    comm, k = og.louvain([1, 2, 3, 4], 0.01)
    cl = g.clusters(comm_of=comm)
    bc, L, C = cl.evaluate()
    ref = np.zeros(g.n)
    for c in range(k):
        verts = cl.members(c)
        obc, ob, oL, oC = og.cluster_eval(verts)
        assert np.array_equal(ob, cl.borders(c))
        ref[verts] = obc[verts]
        assert np.array_equal(cl.vi_rows(L, c), oL.astype(np.uint64))
        assert np.array_equal(cl.vi_rows(C, c).astype(np.int64), oC.astype(np.int64))
    assert np.array_equal(bc, ref)


@pytest.mark.parametrize("half_max", [0, 65536])
def test_cluster_eval_team_forms_agree(gpu_ctx, half_max):
    """The cluster launch in its wavefront form (half teams off) and with
    every view on half wavefronts gives the same intra BC and VertexInfo
    rows as the default (cfg0's clusters of a few hundred vertices run on
    half wavefronts by default)."""
    n, (s, d, w) = F.generate_config(0)
    comm, _ = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    cl = g.clusters(comm_of=comm)
    bc, L, C = cl.evaluate()
    gpu_ctx.set_option("half_team_max", half_max)
    try:
        bc2, L2, C2 = cl.evaluate()
    finally:
        gpu_ctx.set_option("half_team_max", 2048)
    assert np.array_equal(bc, bc2) and np.array_equal(L, L2) and np.array_equal(C, C2)


def test_cfg0_pivot_phase(cfg0):
    g, og = cfg0
    r = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=0.5, threads=1)
    cl = g.clusters(comm_of=r["comm"])
    out = cl.pivot_phase(r["pivots"], r["weights"], r["pivot_cluster"], r["intra"], r["intra"])
    ref = r["bc"]
    # magnitude of the summed terms bounds the rounding of any summation order
    dep = g.dep_batch(r["pivots"])
    mag = np.abs(r["intra"]) + (np.abs(dep) * r["weights"][:, None]).sum(0)
    wsum = np.zeros(cl.ncl)
    np.add.at(wsum, r["pivot_cluster"], r["weights"])
    mag += wsum[r["comm"]] * np.abs(r["intra"])
    err = np.abs(out - ref)
    assert (err <= 1e-9 * np.abs(ref) + 1e-13 * mag).all(), (err / np.maximum(mag, 1)).max()


def test_cfg0_plain_single_source_bitexact(cfg0, gpu_ctx):
    """The plain form (one cooperative launch per source, FP64 in
    the reference's successor order) equals the oracle bit for bit."""
    g, og = cfg0
    for s in (0, 17, 2049, g.n - 1):
        dep = g.single_source_dep(s)
        assert_plain_ran(gpu_ctx, 1)
        assert np.array_equal(dep, og.single_source_dep(s)), f"source {s}"


def test_rmat_plain_single_source_heavy_rows(gpu_ctx):
    """R-MAT hubs exceed the per-thread degree limit, so their arcs go through
    the workgroup-split path of every phase."""
    n, (s, d, w) = F.generate_rmat(10, 16, seed=11)
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    deg = np.bincount(s, minlength=n)
    assert deg.max() > 96
    for src in (int(deg.argmax()), 0, n - 1):
        dep = g.single_source_dep(src)
        assert_plain_ran(gpu_ctx, 1)
        assert np.array_equal(dep, og.single_source_dep(src)), f"source {src}"


def test_cfg0_pivot_phase_plain_matches_batched(cfg0, gpu_ctx):
    g, og = cfg0
    r = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=0.5, threads=1)
    cl = g.clusters(comm_of=r["comm"])
    args = (r["pivots"], r["weights"], r["pivot_cluster"], r["intra"], r["intra"])
    plain = cl.pivot_phase(*args, plain=True)
    assert_plain_ran(gpu_ctx, len(r["pivots"]))
    batched = cl.pivot_phase(*args)
    ref = r["bc"]
    dep = g.dep_batch(r["pivots"])
    mag = np.abs(r["intra"]) + (np.abs(dep) * r["weights"][:, None]).sum(0)
    wsum = np.zeros(cl.ncl)
    np.add.at(wsum, r["pivot_cluster"], r["weights"])
    mag += wsum[r["comm"]] * np.abs(r["intra"])
    for out in (plain, batched):
        err = np.abs(out - ref)
        assert (err <= 1e-9 * np.abs(ref) + 1e-13 * mag).all(), (err / np.maximum(mag, 1)).max()


def interleave(L, C):
    """VertexInfo rows as k-means features: length then count per border."""
    L, C = np.asarray(L, float), np.asarray(C, float)
    f = np.empty((L.shape[0], 2 * L.shape[1]))
    f[:, 0::2], f[:, 1::2] = L, C
    return f


def test_kmeans_batch_matches_reference(gpu_ctx):
    gold = np.load(os.path.join(GOLDEN, "kmeans_cases.npz"))
    sets = [(int(gold[f"k{t}"][0]), interleave(gold[f"L{t}"], gold[f"C{t}"]), gold[f"W{t}"]) for t in range(24)]
    for t, (cent, cw) in enumerate(gpu_ctx.kmeans_batch(sets)):
        assert np.array_equal(cent, gold[f"cent{t}"]) and np.array_equal(cw, gold[f"cw{t}"]), t


@pytest.mark.parametrize("fresh", [True, False])
def test_kmeans_batch_small_gram_budget(gpu_ctx, fresh):
    """A 1 MiB Gram budget: a context that has run no Brandes launch holds no
    slot memory and runs its sets in groups of about one each; one that has
    (the shared context, after the tests above) puts the Gram matrices in
    its idle slot memory.  Either way the reference's centroids and weights."""
    gold = np.load(os.path.join(GOLDEN, "kmeans_cases.npz"))
    sets = [(int(gold[f"k{t}"][0]), interleave(gold[f"L{t}"], gold[f"C{t}"]), gold[f"W{t}"]) for t in range(24)]
    ctx = F.Context(0) if fresh else gpu_ctx
    ctx.set_option("kmeans_budget_mb", 1)
    try:
        got = ctx.kmeans_batch(sets)
    finally:
        ctx.set_option("kmeans_budget_mb", 0)
        if fresh:
            ctx.close()
    for t, (cent, cw) in enumerate(got):
        assert np.array_equal(cent, gold[f"cent{t}"]) and np.array_equal(cw, gold[f"cw{t}"]), t


def test_kmeans_batch_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(77)
    sets, expect = [], []
    for t in range(12):
        P, nb = int(rng.integers(2, 700)), int(rng.integers(1, 40))
        k = int(rng.integers(1, P))
        L = rng.integers(0, 3000 if t % 2 else 4, (P, nb)).astype(float)   # ties when values are small
        C = rng.integers(0, 3, (P, nb)).astype(np.int32)
        W = rng.integers(1, 9, P).astype(np.int32)
        sets.append((k, interleave(L, C), W))
        expect.append(O.kmeans(k, L, C, W))
    for (cent, cw), (oc, ow) in zip(gpu_ctx.kmeans_batch(sets), expect):
        assert np.array_equal(cent, oc) and np.array_equal(cw, ow)


def test_sigma_overflow_detected(gpu_ctx):
    """Unit-weight bidirected 70x70 grid: sigma to the far corner is
    C(138, 69) > 2^64."""
    n, (s, d, w) = F.generate_grid(70, 70, 1, 7, road=False)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    with pytest.raises(F.FbcError) as e:
        g.sssp_sigma([0])
    assert e.value.code == -5


def test_unit_grid_sigma_closed_form(gpu_ctx):
    """sigma(0 -> (r, c)) = C(r + c, r) on a unit-weight bidirected grid."""
    from math import comb
    R = 30
    n, (s, d, w) = F.generate_grid(R, R, 1, 7, road=False)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    dist, sig = g.sssp_sigma([0])
    # C(58, 29) ~ 3.0e16 is beyond 2^53: exact here, rounded in the reference's FP64
    assert gpu_ctx.last_counters()["sigma_inexact"] == 1
    r, c = np.divmod(np.arange(n), R)
    assert np.array_equal(dist[0], (r + c).astype(np.uint64))
    assert all(int(sig[0][v]) == comb(int(r[v] + c[v]), int(r[v])) for v in range(n))


def test_unit_strip_sigma_beyond_48_bits(gpu_ctx):
    """A 5 x 14000 unit-weight bidirected strip: 70,000 vertices (a workgroup
    per source) and sigma(0 -> (r, c)) = C(r + c, r) up to C(14003, 4), about
    1.6e15: beyond 48 bits, so the sigma sweep starts packed and runs again in
    64 bits; below 2^53, so the oracle's FP64 dependencies are exact too.  The
    pivot phase over the same sources (one cluster, no intra-cluster term)
    equals the weighted sum of their dependency rows."""
    from math import comb
    C = 14000
    n, (s, d, w) = F.generate_grid(5, C, 1, 7, road=False)
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    dist, sig = g.sssp_sigma([0])
    assert gpu_ctx.last_counters()["sigma_wide"] == 1
    assert gpu_ctx.last_counters()["sigma_inexact"] == 0   # 1.6e15 < 2^53
    r, c = np.divmod(np.arange(n), C)
    assert np.array_equal(dist[0], (r + c).astype(np.uint64))
    assert all(int(sig[0][v]) == comb(int(r[v] + c[v]), int(r[v])) for v in range(n))
    check_sssp(g, og, [0, 3 * C + C // 2, n - 1])
    srcs = np.array([0, 5, 2 * C + 7, n - 1], np.uint32)
    dep = g.dep_batch(srcs)
    for i, src in enumerate(srcs):
        ref = og.single_source_dep(int(src))
        assert np.array_equal(dep[i], ref), f"source {src}"
        assert np.array_equal(g.single_source_dep(int(src)), ref), f"plain, source {src}"
        assert_plain_ran(gpu_ctx, 1)
    wts = np.array([1, 3, 2, 5], np.uint32)
    cl = g.clusters(comm_of=np.zeros(n, np.int64))
    bc = cl.pivot_phase(srcs, wts, np.zeros(len(srcs), np.uint32), np.zeros(n), np.zeros(n))
    expect = (dep * wts[:, None].astype(float)).sum(0)
    assert np.allclose(bc, expect, rtol=1e-12, atol=0), np.abs(bc - expect).max()


def test_rmat_exact_bc(gpu_ctx):
    n, (s, d, w) = F.generate_rmat(10, 16, seed=11)
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    check_sssp(g, og, [0, 1, 5, n - 1])
    bc = g.exact_bc()
    ref = og.exact_bc(THREADS)
    rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-12 * np.abs(ref).max())
    assert rel.max() <= 1e-9, rel.max()


def test_cfg2_dist_sigma_properties(gpu_ctx):
    """Full-size road grid (1,048,576 vertices): Bellman optimality of the
    distances and sigma(v) = sum of sigma over DAG predecessors."""
    n, (s, d, w) = F.generate_config(2)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    srcs = [0, 523777, n - 1]
    dist, sig = g.sssp_sigma(srcs)
    s64, d64, w64 = s.astype(np.int64), d.astype(np.int64), w.astype(np.int64)
    for i, src in enumerate(srcs):
        D = dist[i].astype(np.int64)
        assert D[src] == 0 and (dist[i] != U64MAX).all()
        assert (D[d64] <= D[s64] + w64).all()
        tight = D[s64] + w64 == D[d64]
        has_pred = np.zeros(n, bool)
        has_pred[d64[tight]] = True
        has_pred[src] = True
        assert has_pred.all()
        acc = np.zeros(n, np.uint64)
        np.add.at(acc, d64[tight], sig[i][s64[tight]])
        acc[src] = 1
        assert np.array_equal(acc, sig[i])


def test_ballot_append_same_results(gpu_ctx):
    """Wavefront-aggregated appends (one atomic per wavefront for the near
    queues, far lists and DAG levels) change only the order of the lists:
    distances, sigma and the ordered FP64 dependencies stay bit-exact, for
    wavefront teams (cfg0), half-wavefront teams (cfg0's cluster views),
    workgroup teams (a 1M-vertex grid) and the plain form."""
    n, (s, d, w) = F.generate_config(0)
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    comm, _ = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    cl = g.clusters(comm_of=comm)
    srcs = np.random.default_rng(9).choice(n, 24, replace=False)
    n2, arcs2 = F.generate_config(2)
    g2 = F.Graph(gpu_ctx, n2, *F.csr_from_arcs(n2, *arcs2))
    big = np.array([0, 523777], np.uint32)
    base = (g.dep_batch(srcs), cl.evaluate(), g2.dep_batch(big))
    gpu_ctx.set_option("ballot_append", 1)
    try:
        agg = (g.dep_batch(srcs), cl.evaluate(), g2.dep_batch(big))
        plain = g.single_source_dep(int(srcs[0]))
        assert_plain_ran(gpu_ctx, 1)
        check_sssp(g, og, srcs[:6])
    finally:
        gpu_ctx.set_option("ballot_append", 0)
    assert np.array_equal(agg[0], base[0]) and np.array_equal(agg[2], base[2])
    assert all(np.array_equal(a, b) for a, b in zip(agg[1], base[1]))
    assert np.array_equal(plain, og.single_source_dep(int(srcs[0])))


def test_wide_weight_rows_bitexact(gpu_ctx):
    """Weights beyond a short row's 16 bits (0xFFFE) mark every row wide, so
    the distance phase relaxes each short row as empty and takes the CSR
    path, and the LDS view holds only wide rows: distances, sigma and FP64
    dependencies stay bit-exact against the oracle, and the cluster
    evaluation with the rows in LDS equals the HBM-row kernel."""
    n, (s, d, w) = F.generate_config(0)
    rng = np.random.default_rng(7)
    w = (w.astype(np.uint64) * 1000 + rng.integers(0, 1000, len(w))).astype(np.uint32)   # 1,000 .. 100,999
    assert w.max() > 0xFFFE
    g, og = both_graphs(gpu_ctx, n, s, d, w)
    srcs = rng.choice(n, 24, replace=False)
    check_sssp(g, og, srcs)
    dep = g.dep_batch(srcs)
    for i, x in enumerate(srcs):
        assert np.array_equal(dep[i], og.single_source_dep(int(x))), f"source {x}"
    comm, _ = og.louvain([1, 2, 3, 4], 0.01)
    cl = g.clusters(comm_of=comm)
    gpu_ctx.set_option("lds_view", 0)
    try:
        bc0, L0, C0 = cl.evaluate()
        gpu_ctx.set_option("lds_view", 1)
        bc1, L1, C1 = cl.evaluate()
    finally:
        gpu_ctx.set_option("lds_view", 2)
    assert np.array_equal(bc1, bc0) and np.array_equal(L1, L0) and np.array_equal(C1, C0)
    c0 = 5
    verts = cl.members(c0)
    obc, ob, oL, oC = og.cluster_eval(verts)
    assert np.array_equal(bc0[verts], obc[verts])
