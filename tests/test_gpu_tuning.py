"""The autotuner (fbc_set_option "autotune"): a large batched run tunes delta
(a walk over powers of two of the mean arc weight, 1/16 to 8) and, for views
small enough for half-wavefront teams, the team form, on interleaved trial
sets of its own tasks; it keeps the choice for the graph, runs every task
exactly once (the results match the oracle) and changes no result beyond
the summation order.  Capping the slots makes cfg0's and cfg1's runs large
enough to tune."""
import os

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


def test_autotune_exact_bc(gpu_ctx):
    n, (s, d, w) = F.generate_config(0)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    mean = float(np.mean(w))
    ref = O.Graph.from_arcs(s, d, w.astype(np.float64)).exact_bc(THREADS)
    floor = 1e-12 * np.abs(ref).max()
    gpu_ctx.set_tuning(0, 64)
    try:
        bc1 = g.exact_bc()
        c1 = gpu_ctx.last_counters()
        bc2 = g.exact_bc()
        c2 = gpu_ctx.last_counters()
        gpu_ctx.set_option("autotune", 0)
        bc0 = g.exact_bc()
        c0 = gpu_ctx.last_counters()
    finally:
        gpu_ctx.set_option("autotune", 1)
        gpu_ctx.set_tuning(0, 0)
    print(f"tuned delta {c1['delta']} (mean weight {mean:.1f}); cached {c2['delta']}; fixed {c0['delta']}", flush=True)
    assert (c1["tuned"], c2["tuned"], c0["tuned"]) == (1, 2, 0)
    walk = {max(1, int(mean * 2.0 ** k)) for k in range(-4, 4)}   # 1/16 .. 8 mean weights
    assert c1["delta"] == c2["delta"] and c1["delta"] in walk
    assert c0["delta"] == int(mean * 4)
    for bc in (bc1, bc2, bc0):
        assert (np.abs(bc - ref) / np.maximum(np.abs(ref), floor)).max() <= 1e-9


def test_tune_log_lines(gpu_ctx, capfd):
    """tune_log prints the trial launches on stderr and changes no result."""
    n, (s, d, w) = F.generate_config(0)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))   # a new graph: tuned afresh
    ref = O.Graph.from_arcs(s, d, w.astype(np.float64)).exact_bc(THREADS)
    floor = 1e-12 * np.abs(ref).max()
    gpu_ctx.set_tuning(0, 64)
    gpu_ctx.set_option("tune_log", 1)
    try:
        bc = g.exact_bc()
        c = gpu_ctx.last_counters()
    finally:
        gpu_ctx.set_option("tune_log", 0)
        gpu_ctx.set_tuning(0, 0)
    err = capfd.readouterr().err
    lines = [x for x in err.splitlines() if x.startswith("[fbc tune]")]
    print("\n".join(lines), flush=True)
    assert c["tuned"] == 1
    assert any(" region: " in x for x in lines)
    assert sum("candidate delta" in x for x in lines) >= 2
    assert any(f" rest: delta {c['delta']}," in x for x in lines)
    assert (np.abs(bc - ref) / np.maximum(np.abs(ref), floor)).max() <= 1e-9


def test_pivot_walk_starts_from_cluster_choice(gpu_ctx, capfd):
    """After the cluster evaluation has tuned on a graph, the pivot phase
    tunes without a region launch, walking from the cluster's delta, and its
    sums equal an untuned run's."""
    n, (s, d, w) = F.generate_config(0)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))   # a new graph: nothing tuned yet
    comm, k = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    cl = g.clusters(comm_of=comm)
    piv = np.arange(0, n, 2, dtype=np.uint32)
    ones = np.ones(len(piv), np.uint32)
    gpu_ctx.set_tuning(0, 64)
    gpu_ctx.set_option("tune_log", 1)
    try:
        intra = cl.evaluate_resident()
        cc = gpu_ctx.last_counters()
        out = cl.pivot_phase(piv, ones, comm[piv], intra, np.zeros(n), prec=F.PREC_F64)
        cp = gpu_ctx.last_counters()
        gpu_ctx.set_option("autotune", 0)
        ref = cl.pivot_phase(piv, ones, comm[piv], intra, np.zeros(n), prec=F.PREC_F64)
    finally:
        gpu_ctx.set_option("autotune", 1)
        gpu_ctx.set_option("tune_log", 0)
        gpu_ctx.set_tuning(0, 0)
    err = capfd.readouterr().err
    lines = [x for x in err.splitlines() if x.startswith("[fbc tune]")]
    print("\n".join(lines), flush=True)
    print(f"cluster delta {cc['delta']}, pivot delta {cp['delta']}", flush=True)
    assert cc["tuned"] == 1 and cp["tuned"] == 1
    assert any("mode 3: no region launch" in x and f"choice {cc['delta']})" in x for x in lines)
    assert not any(x.startswith("[fbc tune] mode 3 region") for x in lines)
    assert any(x.startswith("[fbc tune] mode 4 region") for x in lines)
    scale = np.abs(ref).max()
    assert np.abs(out - ref).max() <= 1e-12 * scale


def test_autotune_cluster_eval_identical(gpu_ctx):
    """Cluster evaluation tuned (half-wavefront or wavefront teams, three
    deltas) equals the fixed configuration bit for bit: the intra-cluster
    sums run in a fixed order whatever launched the rows."""
    n, (s, d, w) = F.generate_config(1)
    comm, _ = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    cl = g.clusters(comm_of=comm)
    gpu_ctx.set_tuning(0, 256)
    try:
        bc1, L1, C1 = cl.evaluate()
        c1 = gpu_ctx.last_counters()
        # tuned again on the LDS-view kernel: its trial launches (mixed-delta
        # sets, then one delta each) run tasks whose views are staged in LDS
        g2 = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
        cl2 = g2.clusters(comm_of=comm)
        gpu_ctx.set_option("lds_view", 1)
        bc2, L2, C2 = cl2.evaluate()
        c2 = gpu_ctx.last_counters()
        gpu_ctx.set_option("lds_view", 2)
        gpu_ctx.set_option("autotune", 0)
        bc0, L0, C0 = cl.evaluate()
    finally:
        gpu_ctx.set_option("lds_view", 2)
        gpu_ctx.set_option("autotune", 1)
        gpu_ctx.set_tuning(0, 0)
    print(f"cluster views up to {int(np.bincount(comm).max())} vertices: tuned delta {c1['delta']}, "
          f"half-team limit {c1['half_team_max']}; on the LDS view: delta {c2['delta']}, "
          f"half-team limit {c2['half_team_max']}", flush=True)
    assert c1["tuned"] == 1 and c2["tuned"] == 1 and c2["lds_view"] == 1
    assert np.array_equal(bc1, bc0) and np.array_equal(L1, L0) and np.array_equal(C1, C0)
    assert np.array_equal(bc2, bc0) and np.array_equal(L2, L0) and np.array_equal(C2, C0)


@pytest.mark.parametrize("cfg", [1, 2])
def test_lds_view_cluster_eval_identical(gpu_ctx, cfg):
    """Cluster evaluation with the clusters' short rows staged in LDS (option
    lds_view=1) equals the HBM-row kernels bit for bit: intra BC, VertexInfo
    lengths and counts.  cfg1's views (up to ~1,200 vertices) run on
    half-wavefront teams and all fit; cfg2's (1,159 to 9,870 vertices) run on
    wavefront teams and the largest read their rows from HBM."""
    n, (s, d, w) = F.generate_config(cfg)
    comm, _ = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    cl = g.clusters(comm_of=comm)
    gpu_ctx.set_option("autotune", 0)
    try:
        gpu_ctx.set_option("lds_view", 0)
        bc0, L0, C0 = cl.evaluate()
        ms0, _ = gpu_ctx.last_stats()
        assert gpu_ctx.last_counters()["lds_view"] == 0
        gpu_ctx.set_option("lds_view", 1)
        bc1, L1, C1 = cl.evaluate()
        ms1, _ = gpu_ctx.last_stats()
        assert gpu_ctx.last_counters()["lds_view"] == 1
        # the default rule: LDS rows when the views that fit carry half the
        # work and every team gets 32 tasks (cfg2: 81% fit, 1M tasks; cfg1:
        # 65K tasks for 8,192 half-wavefront teams)
        gpu_ctx.set_option("lds_view", 2)
        bc2, L2, C2 = cl.evaluate()
        assert gpu_ctx.last_counters()["lds_view"] == (1 if cfg == 2 else 0)
    finally:
        gpu_ctx.set_option("lds_view", 2)
        gpu_ctx.set_option("autotune", 1)
    print(f"cfg{cfg}: cluster launch {ms0:.1f} ms (HBM rows) vs {ms1:.1f} ms (LDS rows), views up to "
          f"{int(np.bincount(comm).max())} vertices", flush=True)
    assert np.array_equal(bc1, bc0) and np.array_equal(L1, L0) and np.array_equal(C1, C0)
    assert np.array_equal(bc2, bc0) and np.array_equal(L2, L0) and np.array_equal(C2, C0)


def test_lds_view_irregular_partition(gpu_ctx):
    """A partition of cfg1's grid into row-major strips of 1 to ~5,000
    vertices (singletons and pairs included, so views change size from task
    to task and the staged view is replaced often): the LDS-view kernel,
    forced on, equals the HBM-row kernel bit for bit."""
    n, (s, d, w) = F.generate_config(1)
    rng = np.random.default_rng(11)
    sizes = []
    while sum(sizes) < n:
        sizes.append(int(rng.choice([1, 2, 3, 17, 250, 1000, 2047, 2049, 5000])))
    comm = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)[:n]
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    cl = g.clusters(comm_of=comm)
    gpu_ctx.set_option("autotune", 0)
    try:
        gpu_ctx.set_option("lds_view", 0)
        bc0, L0, C0 = cl.evaluate()
        gpu_ctx.set_option("lds_view", 1)
        bc1, L1, C1 = cl.evaluate()
        assert gpu_ctx.last_counters()["lds_view"] == 1
    finally:
        gpu_ctx.set_option("lds_view", 2)
        gpu_ctx.set_option("autotune", 1)
    print(f"{len(sizes)} strips of 1..5000 vertices", flush=True)
    assert np.array_equal(bc1, bc0) and np.array_equal(L1, L0) and np.array_equal(C1, C0)
