"""Pivot classes and k-means on the device (fbc_cluster_classes /
fbc_cluster_kmeans) against the host selector and the batch k-means, which
the parity tests pin against the reference: identical classes, pivots,
weights and centroids.  Hash collisions (provoked by keeping a few hash bits)
must be reported, and the pipeline must then select on the host with the same
result."""
import os

import numpy as np
import pytest

import fastbc_amd as F

pytestmark = pytest.mark.gpu


def clusters_of(ctx, cfg):
    n, (s, d, w) = F.generate_config(cfg)
    comm, _ = F.louvain(s, d, w, [1, 2, 3, 4], 0.01)
    g = F.Graph(ctx, n, *F.csr_from_arcs(n, s, d, w))
    return g, g.clusters(comm_of=comm)


def host_classes(cl, intra, L, C):
    out = []
    for c in range(cl.ncl):
        b = cl.borders(c)
        if len(b) == 0:
            out.append((np.zeros(0, np.int32), np.zeros(0, np.int32), None))
            continue
        m = cl.members(c)
        isb = np.isin(m, b).astype(np.uint8)
        rows, wts, norm = F.select_pivots(cl.vi_rows(L, c), cl.vi_rows(C, c), intra[m], isb)
        out.append((rows, wts, norm))
    return out


@pytest.mark.parametrize("cfg", [0, 1])
def test_device_classes_match_host_selector(gpu_ctx, cfg):
    print(f"device classes cfg{cfg}", flush=True)
    g, cl = clusters_of(gpu_ctx, cfg)
    intra, L, C = cl.evaluate()
    intra_r = cl.evaluate_resident()
    assert np.array_equal(intra, intra_r)
    L_r, C_r = cl.rows()
    assert np.array_equal(L, L_r) and np.array_equal(C, C_r)
    ptr, row, wt = cl.classes(intra)
    ref = host_classes(cl, intra, L, C)
    assert int(ptr[-1]) == sum(len(r[0]) for r in ref)
    for c, (rows, wts, _) in enumerate(ref):
        assert np.array_equal(row[ptr[c]:ptr[c + 1]], rows.astype(np.uint32)), c
        assert np.array_equal(wt[ptr[c]:ptr[c + 1]], wts.astype(np.uint32)), c


def test_device_kmeans_matches_batch(gpu_ctx):
    print("device k-means cfg1", flush=True)
    g, cl = clusters_of(gpu_ctx, 1)
    intra, L, C = cl.evaluate()
    cl.evaluate_resident()
    ptr, row, wt = cl.classes(intra)
    ref = host_classes(cl, intra, L, C)
    kptr, cent, cw = cl.kmeans(0.5)
    sets, which = [], []
    for c in range(cl.ncl):
        P = int(ptr[c + 1] - ptr[c])
        if P == 0:
            assert kptr[c + 1] == kptr[c]
            continue
        f = cl.class_features(c, P)
        rows, _, norm = ref[c]
        # the points are the normalized rows of the class pivots, length then count per border
        assert np.array_equal(f[:, 0::2], norm[rows].astype(np.int32))
        assert np.array_equal(f[:, 1::2], cl.vi_rows(C, c)[rows])
        sets.append((max(int(P * 0.5), 1), f, wt[ptr[c]:ptr[c + 1]].astype(np.int32)))
        which.append(c)
    expect = gpu_ctx.kmeans_batch(sets)
    for c, (oc, ow) in zip(which, expect):
        assert np.array_equal(cent[kptr[c]:kptr[c + 1]], oc.astype(np.uint32)), c
        assert np.array_equal(cw[kptr[c]:kptr[c + 1]], ow.astype(np.int32)), c


def test_device_kmeans_gram_in_slot_memory(gpu_ctx):
    """With a 1 MiB Gram budget the device k-means of cfg1's clusters puts its
    Gram matrices in the slot memory the cluster evaluation left idle, and
    chooses the same centroids and weights as with the automatic budget."""
    print("device k-means cfg1, Gram matrices in the slot memory", flush=True)
    g, cl = clusters_of(gpu_ctx, 1)
    intra = cl.evaluate_resident()
    cl.classes(intra)
    ref = cl.kmeans(0.5)
    gpu_ctx.set_option("kmeans_budget_mb", 1)
    try:
        got = cl.kmeans(0.5)
    finally:
        gpu_ctx.set_option("kmeans_budget_mb", 0)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_class_hash_collision_reported(gpu_ctx):
    print("class hash collisions cfg0", flush=True)
    g, cl = clusters_of(gpu_ctx, 0)
    intra = cl.evaluate_resident()
    gpu_ctx.set_option("class_key_bits", 3)
    try:
        with pytest.raises(F.FbcError, match="collision"):
            cl.classes(intra)
    finally:
        gpu_ctx.set_option("class_key_bits", 0)
    ptr, _, _ = cl.classes(intra)
    assert ptr[-1] > 0


@pytest.mark.parametrize("kfrac", [0.0, 0.5])
def test_pipeline_host_fallback_same_result(kfrac, monkeypatch):
    print(f"pipeline fallback kfrac={kfrac}", flush=True)
    n, (s, d, w) = F.generate_config(0)
    out = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac)
    monkeypatch.setenv("FBC_OPTIONS", "class_key_bits=3")
    fb = F.clustered_bc(s, d, w, [1, 2, 3, 4], 0.01, kfrac=kfrac)
    for k in ("pivots", "weights", "pivot_cluster", "intra"):
        assert np.array_equal(out[k], fb[k]), k
    # the pivot phase adds in launch order: equal up to rounding
    assert np.allclose(out["bc"], fb["bc"], rtol=1e-12, atol=1e-9)
