"""Graphs of unusual shape -- a 2-cycle, vertex ids with no arcs, two
components, a star, a random graph with self-loops, repeated arcs (which the
loader sums, DirectedWeightedGraph.h:154-165) and vertices no arc reaches --
against the reference's own outputs (tests/golden/degenerate.npz, written by
make_golden.py --part degenerate from oracle/_ref).

CPU tests pin the oracle on these shapes; GPU tests run the HIP path through
the C ABI: distances and sigma bit-exact, FP64 dependency rows bit-exact,
exact BC within 1e-12 relative (the reference sums over sources in its own
order), the clustered pipeline with the reference's partition, intra-cluster
BC bit for bit, the oracle's pivots (the reference's selectors need spdlog
and are not built here), and BC within 1e-9 relative of the composition
assembled from the reference's pieces (ClusteredBrandesBC.h:130-166)."""
import os
import subprocess

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN

Z = np.load(os.path.join(GOLDEN, "degenerate.npz"))
CASES = ["two_cycle", "isolated_ids", "two_components", "star", "loops_duplicates"]
KFRACS = [("k0", 0.0), ("k5", 0.5)]
U64MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def arcs(name):
    return Z[f"{name}_src"], Z[f"{name}_dst"], Z[f"{name}_w"]


def oracle_graph(name):
    s, d, w = arcs(name)
    return O.Graph.from_arcs(s, d, w.astype(np.float64))


def close(out, ref, rtol, atol=1e-12):
    err = np.abs(out - ref)
    assert (err <= rtol * np.abs(ref) + atol).all(), float(err.max())


def expected_clustered(name, tag):
    """intra + sum_p w_p dep_p - W_c intra, from the reference's intra-cluster
    BC and its dependency sum over the pivots; and the magnitude of the
    summed terms (dependencies are non-negative)."""
    comm, intra = Z[f"{name}_comm"], Z[f"{name}_intra"]
    ssb = Z[f"{name}_{tag}_ssb"]
    wsum = np.zeros(int(comm.max()) + 1)
    np.add.at(wsum, Z[f"{name}_{tag}_pcl"], Z[f"{name}_{tag}_weights"].astype(np.float64))
    expect = intra + ssb - wsum[comm] * intra
    mag = np.abs(intra) + ssb + wsum[comm] * np.abs(intra)
    return expect, mag


# ---- the oracle on these shapes (CPU) ---------------------------------------

@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    og = oracle_graph(name)
    comm, _ = og.louvain([1, 2, 3, 4], 0.01)
    assert np.array_equal(comm, Z[f"{name}_comm"])
    close(og.exact_bc(), Z[f"{name}_exact"], 1e-12)
    for tag, kf in KFRACS:
        r = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=kf)
        assert np.array_equal(r["intra"], Z[f"{name}_intra"])
        assert np.array_equal(r["pivots"], Z[f"{name}_{tag}_pivots"])
        assert np.array_equal(r["weights"], Z[f"{name}_{tag}_weights"])
        ssb = np.zeros(og.n)
        for p, w in zip(r["pivots"], r["weights"]):
            ssb += og.single_source_dep(int(p)) * w
        close(ssb, Z[f"{name}_{tag}_ssb"], 1e-12)
        expect, mag = expected_clustered(name, tag)
        close(r["bc"], expect, 1e-9, 1e-12 * mag.max())


def test_fixture_shapes():
    """The cases keep the shapes they are meant to have."""
    s, d, _ = arcs("isolated_ids")
    assert set(range(10)) - set(s) - set(d) == {1, 2, 3, 4, 6, 7, 8}
    s, d, w = arcs("loops_duplicates")
    assert (s == d).sum() >= 8
    keys = s.astype(np.int64) * 1000 + d
    assert len(np.unique(keys)) < len(keys)
    assert len(set(range(60)) - set(d)) > 0   # some vertex no arc reaches
    assert Z["star_exact"][0] == 20 * 19


# ---- the HIP path -------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_sssp_rows_exact(name, gpu_ctx):
    og = oracle_graph(name)
    rp, col, w = og.csr()
    g = F.Graph(gpu_ctx, og.n, rp, col, w.astype(np.uint32))
    srcs = np.arange(og.n, dtype=np.uint32)
    dist, sig = g.sssp_sigma(srcs)
    dep = g.dep_batch(srcs)
    for v in range(og.n):
        od, osg, _ = og.sssp(v)
        reached = np.isfinite(od)
        assert np.array_equal(dist[v][~reached], np.full((~reached).sum(), U64MAX)), f"source {v}"
        assert np.array_equal(dist[v][reached], od[reached].astype(np.uint64)), f"source {v}"
        assert np.array_equal(sig[v], osg.astype(np.uint64)), f"source {v}"
        assert np.array_equal(dep[v], og.single_source_dep(v)), f"source {v}"
    for v in (0, og.n - 1):
        assert np.array_equal(g.single_source_dep(v), dep[v])
        c = gpu_ctx.last_counters()
        assert c["plain_fallback_sources"] == 0 and c["plain_sources"] == 1, c
    s, d, w = arcs(name)
    close(F.exact_bc_arcs(s, d, w), Z[f"{name}_exact"], 1e-12)
    close(F.exact_bc_arcs(s, d, w, prec=F.PREC_F32C), Z[f"{name}_exact"], 1e-6)
    close(g.exact_bc(), Z[f"{name}_exact"], 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("tag,kfrac", KFRACS)
@pytest.mark.parametrize("name", CASES)
def test_gpu_clustered(name, tag, kfrac, gpu_ctx):
    s, d, w = arcs(name)
    expect, mag = expected_clustered(name, tag)
    for prec, form, rtol in ((F.PREC_F64, F.FORM_BATCHED, 1e-9), (F.PREC_F32C, F.FORM_BATCHED, 1e-6),
                             (F.PREC_F64, F.FORM_PLAIN, 1e-9)):
        out = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac, prec=prec, form=form)
        assert np.array_equal(out["comm"], Z[f"{name}_comm"])
        assert np.array_equal(out["intra"], Z[f"{name}_intra"])
        assert np.array_equal(out["pivots"], Z[f"{name}_{tag}_pivots"].astype(np.uint32))
        assert np.array_equal(out["weights"], Z[f"{name}_{tag}_weights"].astype(np.uint32))
        close(out["bc"], expect, rtol, 1e-12 * mag.max())


# ---- VertexInfo counts beyond the int range ----------------------------------
# A 20x20 unit-weight grid community whose corner-to-corner path count,
# C(38,19), exceeds INT_MAX: the reference's (int) conversion of it
# (DijkstraClusterEvaluator.h:172-175) gives INT_MIN on x86-64, and the pivot
# classes and k-means points are built from those counts.

BIG_PART = Z["bigsig_part"]


def bigsig_arcs():
    s, d = Z["bigsig_src"], Z["bigsig_dst"]
    return s, d, np.ones(len(s), np.int32)


def test_oracle_big_sigma_counts():
    s, d, w = bigsig_arcs()
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    verts = np.where(BIG_PART == 0)[0]
    bc, b, L, C = og.cluster_eval(verts)
    assert np.array_equal(b, Z["bigsig_borders"])
    assert np.array_equal(bc[verts], Z["bigsig_intra"])
    assert np.array_equal(L, Z["bigsig_len"])
    assert np.array_equal(C, Z["bigsig_cnt"])
    assert (C == np.iinfo(np.int32).min).sum() == 13


@pytest.mark.gpu
def test_gpu_big_sigma_counts(gpu_ctx):
    s, d, w = bigsig_arcs()
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    rp, col, ww = og.csr()
    g = F.Graph(gpu_ctx, og.n, rp, col, ww.astype(np.uint32))
    cl = g.clusters(comm_of=BIG_PART)
    bc, L, C = cl.evaluate()
    assert np.array_equal(cl.borders(0), Z["bigsig_borders"].astype(np.uint32))
    assert np.array_equal(cl.vi_rows(L, 0).astype(np.float64), Z["bigsig_len"])
    assert np.array_equal(cl.vi_rows(C, 0), Z["bigsig_cnt"])
    assert np.array_equal(bc[cl.members(0)], Z["bigsig_intra"])
    for kfrac in (0.0, 0.5):
        ref = O.clustered_bc_partition(og, BIG_PART, kfrac=kfrac)
        mag = np.abs(ref["intra"]).max() * (1 + len(ref["pivots"]))
        for prec, form, rtol in ((F.PREC_F64, F.FORM_BATCHED, 1e-9), (F.PREC_F32C, F.FORM_BATCHED, 1e-6),
                                 (F.PREC_F64, F.FORM_PLAIN, 1e-9)):
            out = F.clustered_bc(s, d, w, [1], kfrac=kfrac, prec=prec, form=form, partition=BIG_PART)
            assert np.array_equal(out["intra"], ref["intra"])
            assert np.array_equal(out["pivots"], ref["pivots"].astype(np.uint32))
            assert np.array_equal(out["weights"], ref["weights"].astype(np.uint32))
            close(out["bc"], ref["bc"], rtol, 1e-12 * mag)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "2-clustered"])
def test_cli_degenerate_text(tmp_path, mode):
    """The fbc command line on the text of the random graph: the loader sums
    repeated arcs and counts each distinct arc once (the log line), and
    bc.txt writes the negative clustered values as 0 (main.cpp:282-294)."""
    s, d, w = arcs("loops_duplicates")
    g = str(tmp_path / "g.txt")
    with open(g, "w") as f:
        f.writelines(f"{a} {b} {c}\n" for a, b, c in zip(s, d, w))
    out, full = str(tmp_path / "bc.txt"), str(tmp_path / "bc_full.txt")
    args = ["--exact"] if mode == "exact" else ["-s", "1,2,3,4", "-e", "4", "-p", "0.01", "-k", "0.5"]
    r = subprocess.run([F.FBC_BIN, *args, "-o", out, "--full-output", full, g], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    og = oracle_graph("loops_duplicates")
    assert f"Loaded graph contains 60 vertices and {og.edges} edges" in r.stdout
    vals = np.loadtxt(full)
    if mode == "exact":
        close(vals, Z["loops_duplicates_exact"], 1e-12)
    else:
        expect, mag = expected_clustered("loops_duplicates", "k5")
        assert (expect < 0).any()
        close(vals, expect, 1e-9, 1e-12 * mag.max())
    assert open(out).read() == "".join(("%g" % x if x >= 0 else "0") + "\n" for x in vals)
