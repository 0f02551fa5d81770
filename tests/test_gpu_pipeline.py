"""End-to-end clustered betweenness on the GPU (include/fbc_host.h entry
points) against the oracle: identical partition, bit-identical intra-cluster
BC, identical pivots and weights, final BC within 1e-9 relative (scaled by the
magnitude of the summed terms where the reference's formula cancels)."""
import os

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN, read_arcs

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


def check_bc(out, ref, mag, rtol=1e-9):
    err = np.abs(out - ref)
    bound = rtol * np.abs(ref) + 1e-13 * mag
    assert (err <= bound).all(), float((err / np.maximum(bound, 1e-300)).max())


def term_magnitude(g_gpu, r):
    """|intra| + sum_p w_p |dep_p| + W_c |intra| per vertex (the reference's
    formula, ClusteredBrandesBC.h:151-164), from GPU dependencies that the
    parity tests pin bit-exactly."""
    n = len(r["bc"])
    mag = np.abs(r["intra"]).copy()
    piv = np.asarray(r["pivots"])
    wts = np.asarray(r["weights"], np.float64)
    for i in range(0, len(piv), 256):
        dep = g_gpu.dep_batch(piv[i:i + 256])
        mag += (np.abs(dep) * wts[i:i + 256, None]).sum(0)
    wsum = np.zeros(int(np.max(r["comm"])) + 1)
    np.add.at(wsum, np.asarray(r["pivot_cluster"]), wts)
    mag += wsum[r["comm"]] * np.abs(r["intra"])
    return mag


@pytest.mark.parametrize("kfrac", [0.0, 0.5])
def test_cfg0_clustered_end_to_end(gpu_ctx, kfrac):
    n, (s, d, w) = F.generate_config(0)
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    ref = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=kfrac, threads=1)
    out = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac)
    assert np.array_equal(out["comm"], ref["comm"])
    assert np.array_equal(out["intra"], ref["intra"])
    assert np.array_equal(out["pivots"], ref["pivots"].astype(np.uint32))
    assert np.array_equal(out["weights"], ref["weights"].astype(np.uint32))
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    check_bc(out["bc"], ref["bc"], term_magnitude(g, ref))
    f32 = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac, prec=F.PREC_F32C)
    nz = np.abs(ref["bc"]) > 1e-9 * np.abs(ref["bc"]).max()
    print(f"cfg0 kfrac {kfrac}: FP32C max relative error vs the FP64 oracle "
          f"{float((np.abs(f32['bc'] - ref['bc'])[nz] / np.abs(ref['bc'][nz])).max()):.3e}", flush=True)
    check_bc(f32["bc"], ref["bc"], term_magnitude(g, ref), rtol=1e-6)
    plain = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac, form=F.FORM_PLAIN)
    check_bc(plain["bc"], ref["bc"], term_magnitude(g, ref))
    sel = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac, form=F.FORM_SELECT_ONLY)
    assert np.array_equal(sel["pivots"], out["pivots"]) and np.array_equal(sel["bc"], ref["intra"])


def test_cfg1_clustered_pivots_and_bc(gpu_ctx):
    n, (s, d, w) = F.generate_config(1)
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    ref = O.clustered_pivots(og, [1, 2, 3, 4], 0.01, kfrac=0.5, threads=THREADS)
    out = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=0.5)
    assert np.array_equal(out["comm"], ref["comm"])
    assert np.array_equal(out["intra"], ref["intra"])
    assert np.array_equal(out["pivots"], ref["pivots"].astype(np.uint32))
    assert np.array_equal(out["weights"], ref["weights"].astype(np.uint32))
    # the pivot phase against the formula assembled from pinned dependencies
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, s, d, w))
    expect = out["intra"].copy()
    piv, wts = out["pivots"], out["weights"].astype(np.float64)
    for i in range(0, len(piv), 256):
        expect += (g.dep_batch(piv[i:i + 256]) * wts[i:i + 256, None]).sum(0)
    wsum = np.zeros(int(out["comm"].max()) + 1)
    np.add.at(wsum, out["pivot_cluster"], wts)
    expect -= wsum[out["comm"]] * out["intra"]
    mag = term_magnitude(g, out)
    check_bc(out["bc"], expect, mag)
    f32 = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=0.5, prec=F.PREC_F32C)
    nz = np.abs(expect) > 1e-9 * np.abs(expect).max()
    print(f"cfg1: FP32C max relative error vs the FP64 oracle-pinned sum "
          f"{float((np.abs(f32['bc'] - expect)[nz] / np.abs(expect[nz])).max()):.3e}", flush=True)
    check_bc(f32["bc"], expect, mag, rtol=1e-6)


def test_exact_through_host_abi(gpu_ctx):
    s, d, w = read_arcs(os.path.join(GOLDEN, "dwg9.txt"))
    bc = F.exact_bc_arcs(s, d, w)
    assert np.allclose(bc, O.Graph.from_file(os.path.join(GOLDEN, "dwg9.txt")).exact_bc(), rtol=1e-12)
    gold = np.load(os.path.join(GOLDEN, "cfg0_exact.npz"))["bc"]
    n, (s, d, w) = F.generate_config(0)
    bc = F.exact_bc_arcs(s, d, w)
    assert (np.abs(bc - gold) / np.maximum(np.abs(gold), 1e-12 * gold.max())).max() <= 1e-9


def test_rmat_exact_f32c(gpu_ctx):
    n, (s, d, w) = F.generate_rmat(11, 16, seed=4)
    ref = O.Graph.from_arcs(s, d, w.astype(np.float64)).exact_bc(THREADS)
    for prec, tol in ((F.PREC_F64, 1e-9), (F.PREC_F32C, 1e-6)):
        bc = F.exact_bc_arcs(s, d, w, prec=prec)
        rel = np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-12 * ref.max())
        assert rel.max() <= tol, (prec, rel.max())


def random_strong_arcs(seed, n):
    rng = np.random.default_rng(seed)
    arcs = {(v, (v + 1) % n): int(rng.integers(1, 4)) for v in range(n)}
    for _ in range(2 * n):
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            arcs[(a, b)] = int(rng.integers(1, 4))
    a = np.array([(x, y, c) for (x, y), c in sorted(arcs.items())])
    return a[:, 0], a[:, 1], a[:, 2].astype(np.float64)


@pytest.mark.parametrize("form,prec", [(F.FORM_BATCHED, F.PREC_F64), (F.FORM_BATCHED, F.PREC_F32C),
                                       (F.FORM_PLAIN, F.PREC_F64)])
def test_composition_closed_forms(gpu_ctx, form, prec):
    """The product's composition with caller-given partitions (the cases of
    tests/test_composition.py, which pin the oracle by the same closed
    forms): a path community cut from a directed cycle,
    (N-1)(N-2)/2 + (1-k) i(k-1-i); borderless communities, 0; singleton
    communities, exact BC; a class of two vertices, the oracle."""
    tol = 1e-12 if prec == F.PREC_F64 else 1e-6
    for k in (3, 8, 40):
        N = k + 1
        s = np.arange(N)
        out = F.clustered_bc(s, (s + 1) % N, np.ones(N), [1], partition=[0] * k + [1], form=form, prec=prec)
        i = np.arange(k)
        expect = np.full(N, (N - 1) * (N - 2) / 2.0)
        expect[:k] += (1 - k) * i * (k - 1 - i)
        assert sorted(out["pivots"].tolist()) == list(range(N)) and (out["weights"] == 1).all()
        assert np.allclose(out["bc"], expect, rtol=tol, atol=tol * np.abs(expect).max()), (k, out["bc"], expect)
    arcs = []
    for base, k in ((0, 5), (5, 4), (9, 6)):
        arcs += [(base + j, base + (j + 1) % k, 1 + j % 3) for j in range(k)]
    a = np.array(arcs)
    for kfrac in (0.0, 0.5):
        out = F.clustered_bc(a[:, 0], a[:, 1], a[:, 2], [1], partition=np.repeat([0, 1, 2], [5, 4, 6]), form=form,
                             prec=prec, kfrac=kfrac)
        assert len(out["pivots"]) == 0 and not out["bc"].any()
    s, d, w = random_strong_arcs(3, 9)
    og = O.Graph.from_arcs(s, d, w)
    out = F.clustered_bc(s, d, w, [1], partition=np.arange(9), form=form, prec=prec)
    ex = og.exact_bc()
    assert np.allclose(out["bc"], ex, rtol=tol, atol=tol * ex.max())
    s, d = np.array([0, 0, 1, 2, 3, 4]), np.array([1, 2, 3, 3, 4, 0])
    w = np.ones(6)
    part = np.array([0, 0, 0, 0, 1])
    out = F.clustered_bc(s, d, w, [1], partition=part, form=form, prec=prec)
    ref = O.clustered_bc_partition(O.Graph.from_arcs(s, d, w), part)
    assert out["pivots"].tolist() == [0, 1, 3, 4] and out["weights"].tolist() == [1, 2, 1, 1]
    assert np.array_equal(out["intra"], ref["intra"])
    assert np.allclose(out["bc"], ref["bc"], rtol=tol, atol=tol * np.abs(ref["bc"]).max())


def test_partition_entry_matches_louvain_run(gpu_ctx):
    """cfg0 with the reference's Louvain partition handed in equals the run
    that computes it (same communities, pivots and BC)."""
    n, (s, d, w) = F.generate_config(0)
    r = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=0.5)
    p = F.clustered_bc(s, d, w, [], partition=r["comm"], kfrac=0.5)
    assert np.array_equal(p["comm"], r["comm"]) and np.array_equal(p["pivots"], r["pivots"])
    assert np.array_equal(p["intra"], r["intra"]) and np.array_equal(p["weights"], r["weights"])
    assert np.allclose(p["bc"], r["bc"], rtol=1e-12, atol=1e-9)
