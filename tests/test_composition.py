"""Pins of the clustered composition (ClusteredBrandeBC::computeBC,
ClusteredBrandesBC.h:63-169) and of KMeansPivotSelector's duplicate removal
(KMeansPivotSelector.h:103-129) by cases whose answer the mathematics fixes,
run on the oracle with caller-given partitions (the reference takes any
IGraphPartition, ClusteredBrandesBC.h:80-82):

  * every community borderless: every cluster is skipped
    (FASTBC_BRANDES_CLUSTERED_IGNORE_UNCONNECTED, main.cpp:6), so BC == 0;
  * singleton communities: every vertex is its own pivot with weight 1 and
    no intra-cluster term, so clustered BC == exact BC (brute force);
  * communities of at most two vertices: no intra-cluster BC, so
    BC == sum_p w_p * dep_p with dependencies from the definition;
  * a directed cycle cut into a path community and one vertex: intra BC is
    the path's i(k-1-i), every vertex is a pivot of weight 1, and the
    result is (N-1)(N-2)/2 + (1-k) i(k-1-i) -- the subtract term in closed
    form;
  * a diamond community whose two middle vertices share a class (weight 2):
    the weighted sum with every term computed from the definition;
  * k-means over identical rows: one pivot carrying the whole weight; on the
    reference's own k-means runs, dropping repeated centroids loses no
    weight (a repeated centroid never wins the first-strict-minimum
    assignment, PlusPlusKMeans.h:98-112, so its weight is 0).
"""
import itertools
import os

import numpy as np
import pytest

import oracle_lib as O
from conftest import GOLDEN


def graph(arcs, n=None):
    s = np.array([a for a, _, _ in arcs], np.uint32)
    d = np.array([b for _, b, _ in arcs], np.uint32)
    w = np.array([c for _, _, c in arcs], np.float64)
    g = O.Graph.from_arcs(s, d, w)
    if n is not None:
        assert g.n == n
    return g


def apsp(g):
    """All-pairs distances (Floyd-Warshall) and shortest-path counts."""
    n = g.n
    rp, col, w = g.csr()
    INF = float("inf")
    D = np.full((n, n), INF)
    for u in range(n):
        D[u, u] = 0
        for e in range(rp[u], rp[u + 1]):
            D[u, col[e]] = min(D[u, col[e]], w[e])
    for k in range(n):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    sig = np.zeros((n, n))
    for s in range(n):
        sig[s, s] = 1
        for x in sorted((x for x in range(n) if D[s, x] < INF and x != s), key=lambda x: D[s, x]):
            for u in range(n):
                for e in range(rp[u], rp[u + 1]):
                    if col[e] == x and D[s, u] + w[e] == D[s, x]:
                        sig[s, x] += sig[s, u]
    return D, sig


def brute_dep(D, sig, s):
    """dep_s(v) = sum over t of sigma_st(v) / sigma_st (v != s, t)."""
    n = len(D)
    dep = np.zeros(n)
    for t, v in itertools.permutations(range(n), 2):
        if s in (t, v) or not np.isfinite(D[s, t]):
            continue
        if D[s, v] + D[v, t] == D[s, t]:
            dep[v] += sig[s, v] * sig[v, t] / sig[s, t]
    return dep


def random_strong_graph(seed, n):
    rng = np.random.default_rng(seed)
    arcs = {(v, (v + 1) % n): int(rng.integers(1, 4)) for v in range(n)}
    for _ in range(2 * n):
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b:
            arcs[(a, b)] = int(rng.integers(1, 4))
    return graph([(a, b, c) for (a, b), c in arcs.items()], n)


def test_borderless_communities_give_zero():
    """Three disjoint directed cycles, one community each: no borders."""
    arcs = []
    for base, k in ((0, 5), (5, 4), (9, 6)):
        arcs += [(base + i, base + (i + 1) % k, 1 + i % 3) for i in range(k)]
    g = graph(arcs, 15)
    comm = np.repeat([0, 1, 2], [5, 4, 6])
    for kfrac in (0.0, 0.5):
        r = O.clustered_bc_partition(g, comm, kfrac)
        assert len(r["pivots"]) == 0
        assert not r["bc"].any() and not r["intra"].any()
    # one cycle split in two: only its halves are bordered, the others stay 0
    comm2 = comm.copy()
    comm2[12:15] = 3
    r = O.clustered_bc_partition(g, comm2, 0.0)
    assert not r["bc"][:9].any()
    assert r["bc"][9:].any() and set(r["pivot_cluster"]) <= {2, 3}


@pytest.mark.parametrize("seed", range(4))
def test_singleton_partition_is_exact_bc(seed):
    g = random_strong_graph(seed, 7)
    D, sig = apsp(g)
    exact = sum(brute_dep(D, sig, s) for s in range(g.n))
    for kfrac in (0.0, 0.5):
        r = O.clustered_bc_partition(g, np.arange(g.n), kfrac)
        assert list(r["pivots"]) == list(range(g.n)) and (r["weights"] == 1).all()
        assert not r["intra"].any()
        assert np.allclose(r["bc"], exact, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_pair_partition_is_weighted_dependency_sum(seed):
    g = random_strong_graph(10 + seed, 8)
    D, sig = apsp(g)
    comm = np.arange(g.n) // 2
    r = O.clustered_bc_partition(g, comm, 0.0)
    assert not r["intra"].any()
    # the classes of a community cover it
    assert np.array_equal(np.bincount(r["pivot_cluster"], weights=r["weights"]), np.full(g.n // 2, 2.0))
    expect = sum(w * brute_dep(D, sig, int(p)) for p, w in zip(r["pivots"], r["weights"]))
    assert np.allclose(r["bc"], expect, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [3, 5, 8])
def test_path_community_closed_form(k):
    """Unit directed cycle 0 -> 1 -> ... -> k -> 0 (N = k + 1), communities
    {0..k-1} and {k}."""
    N = k + 1
    g = graph([(i, (i + 1) % N, 1) for i in range(N)], N)
    comm = np.array([0] * k + [1])
    r = O.clustered_bc_partition(g, comm, 0.0)
    i = np.arange(k)
    assert sorted(r["pivots"]) == list(range(N)) and (r["weights"] == 1).all()
    assert np.array_equal(r["intra"], np.append(i * (k - 1 - i), 0.0))
    expect = np.full(N, (N - 1) * (N - 2) / 2.0)
    expect[:k] += (1 - k) * i * (k - 1 - i)
    assert np.array_equal(r["bc"], expect)


def test_diamond_class_of_two():
    """Community {a, b1, b2, c} = {0, 1, 2, 3} with a -> b1, b2 -> c, plus
    c -> x -> a (x = 4).  b1 and b2 have equal rows (one class of size 2,
    pivot b1: equal intra BC, the first wins)."""
    arcs = [(0, 1, 1), (0, 2, 1), (1, 3, 1), (2, 3, 1), (3, 4, 1), (4, 0, 1)]
    g = graph(arcs, 5)
    r = O.clustered_bc_partition(g, np.array([0, 0, 0, 0, 1]), 0.0)
    assert list(r["pivots"]) == [0, 1, 3, 4] and list(r["weights"]) == [1, 2, 1, 1]
    assert np.array_equal(r["intra"], [0, 0.5, 0.5, 0, 0])
    D, sig = apsp(g)
    expect = np.array([0, 0.5, 0.5, 0, 0])
    for p, w in zip(r["pivots"], r["weights"]):
        expect += w * brute_dep(D, sig, int(p))
    expect[:4] -= 4 * np.array([0, 0.5, 0.5, 0])   # weights of the community's pivots: 1 + 2 + 1
    assert np.allclose(r["bc"], expect, rtol=1e-12, atol=1e-12)


def test_kmeans_pivots_identical_rows():
    """All rows equal: every centroid is row 0, which takes every member."""
    L = np.full((5, 3), 4.0)
    C = np.ones((5, 3), np.int32)
    W = np.array([1, 2, 3, 4, 5], np.int32)
    for kfrac in (0.4, 0.6, 0.99):
        piv, w = O.kmeans_pivots(L, C, W, kfrac)
        assert list(piv) == [0] and list(w) == [15]


def test_kmeans_dedupe_keeps_all_weight():
    """On the reference's 24 k-means runs (duplicates included): after the
    duplicate removal the pivots are the distinct centroids and their weights
    still add up to the total point weight."""
    gold = np.load(os.path.join(GOLDEN, "kmeans_cases.npz"))
    repeated = 0
    for t in range(24):
        L, C, W = gold[f"L{t}"], gold[f"C{t}"], gold[f"W{t}"]
        k, P = int(gold[f"k{t}"][0]), len(W)
        cent = gold[f"cent{t}"]
        repeated += len(cent) - len(set(cent.tolist()))
        piv, w = O.kmeans_pivots(L, C, W, (k + 0.5) / P)
        assert sorted(piv.tolist()) == sorted(set(cent.tolist())), t
        assert len(set(piv.tolist())) == len(piv)
        assert w.sum() == W.sum(), t
    assert repeated > 0   # the fixtures do exercise the removal
