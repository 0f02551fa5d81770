"""The HIP path at the full sizes of the benchmark configurations, against
fixed answers written by the reference's own code (tests/golden/
make_golden.py) and against properties that hold at any size.

  1,048,576-vertex road grid (-k 0.5): the Louvain partition equals the
      reference's; on its three smallest communities the borders, the
      intra-cluster BC (bit for bit) and the VertexInfo rows (digests) equal
      DijkstraClusterEvaluator's, and the pivots and weights follow the
      reference's PlusPlusKMeans on their exact classes; the dependency rows
      of four pivots equal DijkstraSSBrandesBC's bit for bit (digests); the
      whole selection equals the oracle's (all 211 communities).
  32,768-vertex R-MAT graph (--exact): BC within 1e-9 (FP64) and 1e-6
      (FP32 two-sum) of ExactBrandesBC's; distances and sigma bit-exact.
  4,194,304-vertex road grid (-k 0.25): the partition's digest equals the
      reference's, dependency rows equal the reference's (digests), Bellman
      and sigma recurrences hold, and the FP32 two-sum pivot phase stays
      within 1e-6 of the FP64 one on a strided sample of the pivots.
  65,536-vertex road grid: the plain form (one cooperative launch
      per pivot) over every pivot of the run, against the batched form.
"""
import hashlib
import os

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)
U64MAX = np.uint64(0xFFFFFFFFFFFFFFFF)
SEEDS = [1, 2, 3, 4]


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


def assert_plain_ran(ctx, nsrc):
    c = ctx.last_counters()
    assert c["plain_fallback_sources"] == 0, f"cooperative launch refused (HIP status {c['plain_refusal']})"
    assert c["plain_sources"] == nsrc, c


def check_properties(g, arcs, sources):
    """Bellman optimality of the distances and sigma(v) = sum of sigma over
    the DAG predecessors, at full size."""
    n = g.n
    s64, d64, w64 = (a.astype(np.int64) for a in arcs)
    dist, sig = g.sssp_sigma(sources)
    for i, src in enumerate(sources):
        D = dist[i].astype(np.int64)
        reached = dist[i] != U64MAX
        assert D[src] == 0 and reached.all()
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


def magnitude(b, intra, comm, pivot_cluster, weights):
    """|intra| + sum_p w_p dep_p + W_c |intra| per vertex, from the result
    itself: the terms are non-negative, so sum_p w_p dep_p = b - intra +
    W_c intra (ClusteredBrandesBC.h:151-164)."""
    wsum = np.zeros(int(comm.max()) + 1)
    np.add.at(wsum, pivot_cluster, np.asarray(weights, np.float64))
    W = np.where(comm >= 0, wsum[np.maximum(comm, 0)], 0.0)
    return np.abs(b) + 2 * W * np.abs(intra)


@pytest.fixture(scope="module")
def cfg2():
    n, arcs = F.generate_config(2)
    sel = F.clustered_bc(*arcs, SEEDS, 0.01, kfrac=0.5, form=F.FORM_SELECT_ONLY)
    return n, arcs, sel


def test_cfg2_partition_and_small_clusters(gpu_ctx, cfg2):
    n, arcs, sel = cfg2
    ref = np.load(os.path.join(GOLDEN, "cfg2_reference.npz"))
    comm_ref = ref["comm"].astype(np.int32)
    assert np.array_equal(sel["comm"], comm_ref), "Louvain partition differs from the reference's"
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, *arcs))
    members = [np.where(comm_ref == c)[0] for c in ref["clusters"]]
    cl = g.clusters(communities=members)
    bc, L, C = cl.evaluate()
    for i, (c, verts) in enumerate(zip(ref["clusters"], members)):
        print(f"cluster {c}: {len(verts)} vertices, {len(ref[f'borders{i}'])} borders", flush=True)
        assert np.array_equal(cl.borders(i), ref[f"borders{i}"])
        assert np.array_equal(bc[verts], ref[f"intra{i}"]), "intra-cluster BC differs from the reference's"
        assert np.array_equal(sel["intra"][verts], ref[f"intra{i}"])
        Li, Ci = cl.vi_rows(L, i), cl.vi_rows(C, i)
        assert Li.shape == tuple(ref[f"shape{i}"])
        assert np.array_equal(sha(Li.astype(np.uint32)), ref[f"len_sha{i}"]), "VertexInfo lengths differ"
        assert np.array_equal(sha(Ci.astype(np.int32)), ref[f"cnt_sha{i}"]), "VertexInfo counts differ"
        # exact classes (the oracle's selector, pinned by the reference's
        # known answer), the reference's k-means centroids over them, then
        # the first occurrence of each pivot
        exact, _, _ = O.select_pivots(verts, Li, Ci, ref[f"intra{i}"], cl.borders(i))
        assert len(exact) == int(ref[f"classes{i}"][0])
        expect_p, expect_w, seen = [], [], set()
        for k, w in zip(ref[f"km_cent{i}"], ref[f"km_w{i}"]):
            p = int(exact[k])
            if p not in seen:
                seen.add(p)
                expect_p.append(p)
                expect_w.append(int(w))
        mine = sel["pivot_cluster"] == c
        assert sel["pivots"][mine].tolist() == expect_p
        assert sel["weights"][mine].tolist() == expect_w


def test_cfg2_selection_matches_oracle(cfg2):
    path = os.path.join(GOLDEN, "cfg2_oracle_pivots.npz")
    if not os.path.exists(path):
        pytest.skip("no oracle selection fixture")
    _, _, sel = cfg2
    ref = np.load(path)
    assert np.array_equal(sha(sel["comm"].astype(np.int32)), ref["comm_sha"])
    assert np.array_equal(sha(sel["intra"]), ref["intra_sha"]), "intra-cluster BC differs from the oracle's"
    assert np.array_equal(sel["pivots"], ref["pivots"])
    assert np.array_equal(sel["weights"], ref["weights"])
    assert np.array_equal(sel["pivot_cluster"], ref["pivot_cluster"])
    print(f"{len(sel['pivots'])} pivots identical to the oracle's", flush=True)


def test_cfg2_dependency_rows_match_reference(gpu_ctx, cfg2):
    n, arcs, sel = cfg2
    ref = np.load(os.path.join(GOLDEN, "cfg2_reference.npz"))
    src = ref["dep_src"]
    assert np.isin(src, sel["pivots"]).all()
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, *arcs))
    dep = g.dep_batch(src)
    for i, s in enumerate(src):
        assert np.array_equal(sha(dep[i]), ref["dep_sha"][i]), f"source {s}: sum {dep[i].sum()!r} vs {ref['dep_sum'][i]!r}"
    # the same rows through the plain form (one cooperative launch each)
    for i, s in enumerate(src[:2]):
        assert np.array_equal(sha(g.single_source_dep(int(s))), ref["dep_sha"][i])
        assert_plain_ran(gpu_ctx, 1)
    check_properties(g, arcs, [int(src[0]), n - 1])


def test_cfg2_fp32c_pivot_phase_error(gpu_ctx, cfg2):
    """FP32 two-sum against FP64 over every pivot of the run, identical
    pivots (the selection does not depend on the dependency arithmetic)."""
    n, arcs, sel = cfg2
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, *arcs))
    cl = g.clusters(comm_of=sel["comm"])
    args = (sel["pivots"], sel["weights"], sel["pivot_cluster"], sel["intra"], sel["intra"])
    b64 = cl.pivot_phase(*args, prec=F.PREC_F64)
    b32 = cl.pivot_phase(*args, prec=F.PREC_F32C)
    mag = magnitude(b64, sel["intra"], sel["comm"], sel["pivot_cluster"], sel["weights"])
    err = np.abs(b32 - b64)
    print(f"FP32C vs FP64: max |diff| / magnitude {float((err / np.maximum(mag, 1e-300)).max()):.3e}, "
          f"max relative {float((err / np.maximum(np.abs(b64), 1e-300))[b64 != 0].max()):.3e}", flush=True)
    assert (err <= 1e-6 * np.abs(b64) + 1e-13 * mag).all()


def test_cfg3_exact_against_reference(gpu_ctx):
    gold = np.load(os.path.join(GOLDEN, "cfg3_exact.npz"))["bc"]
    n, (s, d, w) = F.generate_config(3)
    assert len(gold) == n
    floor = 1e-12 * np.abs(gold).max()
    for prec, tol in ((F.PREC_F64, 1e-9), (F.PREC_F32C, 1e-6)):
        bc = F.exact_bc_arcs(s, d, w, prec=prec)
        rel = np.abs(bc - gold) / np.maximum(np.abs(gold), floor)
        print(f"cfg3 exact prec {prec}: max relative error vs reference {rel.max():.3e}", flush=True)
        assert rel.max() <= tol
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    hub = int(np.bincount(s, minlength=n).argmax())
    srcs = [0, hub, 12345, n - 1]
    dist, sig = g.sssp_sigma(srcs)
    for i, src in enumerate(srcs):
        od, osg, _ = og.sssp(src)
        reached = np.isfinite(od)
        assert np.array_equal(dist[i][reached], od[reached].astype(np.uint64))
        assert (dist[i][~reached] == U64MAX).all()
        assert np.array_equal(sig[i], osg.astype(np.uint64))


def test_cfg4_full_size(gpu_ctx):
    ref = np.load(os.path.join(GOLDEN, "cfg4_reference.npz"))
    n, arcs = F.generate_config(4)
    sel = F.clustered_bc(*arcs, SEEDS, 0.01, kfrac=0.25, form=F.FORM_SELECT_ONLY)
    print(f"cfg4: {int(sel['comm'].max()) + 1} communities, {len(sel['pivots'])} pivots, "
          f"selection {sel['phases']['total_s']:.1f} s", flush=True)
    assert int(sel["comm"].max()) + 1 == int(ref["ncomm"][0])
    assert np.array_equal(sha(sel["comm"].astype(np.int32)), ref["comm_sha"]), "partition differs from the reference's"
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, *arcs))
    dep = g.dep_batch(ref["dep_src"])
    for i, s in enumerate(ref["dep_src"]):
        assert np.array_equal(sha(dep[i]), ref["dep_sha"][i]), f"source {s}"
    del dep
    check_properties(g, arcs, [int(sel["pivots"][0]), n // 2 + 77])
    # FP32 two-sum against FP64 on a strided sample of the run's pivots
    cl = g.clusters(comm_of=sel["comm"])
    k = max(1, len(sel["pivots"]) // 1024)
    piv, wts, pc = sel["pivots"][::k], sel["weights"][::k], sel["pivot_cluster"][::k]
    b64 = cl.pivot_phase(piv, wts, pc, sel["intra"], sel["intra"], prec=F.PREC_F64)
    b32 = cl.pivot_phase(piv, wts, pc, sel["intra"], sel["intra"], prec=F.PREC_F32C)
    mag = magnitude(b64, sel["intra"], sel["comm"], pc, wts)
    err = np.abs(b32 - b64)
    print(f"cfg4 FP32C vs FP64 on {len(piv)} pivots: max |diff| / magnitude "
          f"{float((err / np.maximum(mag, 1e-300)).max()):.3e}", flush=True)
    assert (err <= 1e-6 * np.abs(b64) + 1e-13 * mag).all()


def test_cfg1_plain_form(gpu_ctx):
    """Every pivot of the run through the plain form: one cooperative launch
    of 16 workgroups per pivot (the grid the benchmark's plain figure uses),
    in chunks of 256 pivots with a progress line each (the pivot terms add
    up, so chunks compose), against the batched FP64 form."""
    n, (s, d, w) = F.generate_config(1)
    sel = F.clustered_bc(s, d, w, SEEDS, 0.01, kfrac=0.5, form=F.FORM_SELECT_ONLY)
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    cl = g.clusters(comm_of=sel["comm"])
    piv, wts, pc = sel["pivots"], sel["weights"], sel["pivot_cluster"]
    print(f"cfg1: {len(piv)} pivots", flush=True)
    plain = sel["intra"].copy()
    gpu_ctx.set_tuning(0, 16)
    try:
        for i in range(0, len(piv), 256):
            plain = cl.pivot_phase(piv[i:i + 256], wts[i:i + 256], pc[i:i + 256], sel["intra"], plain, plain=True)
            assert_plain_ran(gpu_ctx, len(piv[i:i + 256]))
            ms, _ = gpu_ctx.last_stats()
            print(f"  plain pivots {i}..{min(i + 256, len(piv))}: {ms:.0f} ms", flush=True)
    finally:
        gpu_ctx.set_tuning(0, 0)
    batched = cl.pivot_phase(piv, wts, pc, sel["intra"], sel["intra"], prec=F.PREC_F64)
    mag = magnitude(batched, sel["intra"], sel["comm"], pc, wts)
    err = np.abs(plain - batched)
    print(f"cfg1 plain vs batched FP64: max |diff| / magnitude {float((err / np.maximum(mag, 1e-300)).max()):.3e}",
          flush=True)
    assert (err <= 1e-9 * np.abs(batched) + 1e-13 * mag).all()
