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
    check_bc(out["bc"], expect, term_magnitude(g, out))


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
