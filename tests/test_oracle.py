"""Pins the CPU oracle (the checker of every GPU test).

Each oracle function is checked by something that does not restate it:
  - the reference's own known answers (libfastbc/test/*.cpp, DWGtext.txt);
  - fixtures written by the reference's code (tests/golden/make_golden.py);
  - closed forms (paths, stars, cycles, binomial path counts on grids);
  - brute force from the definition of betweenness on tiny graphs.
"""
import ctypes
import itertools
import json
import os
from math import comb

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN


def dwg9():
    return O.Graph.from_file(os.path.join(GOLDEN, "dwg9.txt"))


def graph(arcs):
    s, d, w = (np.array(x) for x in zip(*arcs))
    return O.Graph.from_arcs(s, d, w.astype(np.float64))


# ---- reference known answers ------------------------------------------------

def test_loader_known_answers():
    """test/DirectedWeightedGraph.cpp:23-40"""
    g = dwg9()
    assert g.n == 9 and g.edges == 16
    rp, col, w = g.csr()
    fs4 = dict(zip(col[rp[4]:rp[5]].tolist(), w[rp[4]:rp[5]].tolist()))
    assert fs4 == {5: 1, 6: 5, 8: 3}
    assert g.bstar(4) == {0: 7, 2: 4, 3: 3}
    fs = lambda u: dict(zip(col[rp[u]:rp[u + 1]].tolist(), w[rp[u]:rp[u + 1]].tolist()))
    assert fs(7)[5] == 2 and fs(0)[1] == 4 and 0 not in fs(1)


def test_loader_merges_repeated_arcs_and_rejects_bad_weights(tmp_path):
    g = graph([(0, 1, 2), (1, 0, 1), (0, 1, 3)])
    rp, col, w = g.csr()
    assert g.edges == 2 and list(w[rp[0]:rp[1]]) == [5.0]
    p = tmp_path / "neg.txt"
    p.write_text("0 1 1\n1 0 -2\n")
    with pytest.raises(ValueError, match="greater than zero"):
        O.Graph.from_file(str(p))


def test_subgraph_known_answers():
    """test/SubGraph.cpp:30-46"""
    borders, edges = dwg9().subgraph([0, 1, 2, 3, 4])
    assert list(borders) == [3, 4] and edges == 7


def test_cluster_eval_known_answers():
    """test/brandes/ExactBrandesBC.cpp:31-35 and DijkstraClusterEvaluator.cpp:38-74"""
    bc, b, L, C = dwg9().cluster_eval([0, 1, 2, 3, 4])
    assert list(bc) == [2, 2, 0.5, 1, 0, 0, 0, 0, 0]
    assert list(b) == [3, 4]
    assert L.tolist() == [[5, 7], [1, 4], [8, 4], [0, 3], [0, 0]]
    assert C.tolist() == [[1, 2], [1, 1], [1, 1], [1, 1], [0, 1]]


def test_select_pivots_known_answer():
    """test/brandes/VertexInfoPivotSelector.cpp:16-74"""
    L = np.array([[1, 2, 3], [2, 1, 3], [2, 3, 4], [4, 3, 5], [5, 1, 3]], float)
    C = np.array([[2, 1, 1], [2, 2, 1], [2, 1, 1], [2, 2, 1], [1, 1, 3]], np.int32)
    bc = np.array([1, 2, 2, 1.5, 1])
    piv, w, _ = O.select_pivots([0, 1, 2, 3, 4], L, C, bc, [])
    assert sorted(piv.tolist()) == [0, 3, 4] and list(w) == [2, 2, 1]


def test_vertexinfo_distance_known_answer():
    """test/brandes/VertexInfo.cpp:94-130: squaredDistance == 2.62999964f.
    The value rules out a dropped count term (0.62999964) and |x| in place of
    x^2 (3.2999997); it does not separate orders of the eight terms, which the
    k-means fixtures (non-integer mean distances) exercise instead."""
    L = ctypes.CDLL(O.LIB_PATH)
    L.orc_vi_sqdist_f32.restype = ctypes.c_float
    a = np.array([2.3, 5.2, 1.1, 4.7], np.float32)
    ca = np.array([2, 5, 1, 4], np.int32)
    b = np.array([2, 5, 1, 4], np.int32)
    cb = np.array([3, 4, 1, 4], np.int32)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    assert np.float32(L.orc_vi_sqdist_f32(4, p(a), p(ca), p(b), p(cb))) == np.float32(2.62999964)


# ---- closed forms -------------------------------------------------------------

@pytest.mark.parametrize("n", [2, 5, 9])
def test_directed_path(n):
    g = graph([(i, i + 1, 1 + i % 3) for i in range(n - 1)] + [(n - 1, 0, 10 ** 6)])
    # the closing arc is longer than any path, so BC(i) = i (n - 1 - i) on the
    # forward chain plus the detour pairs through n-1 -> 0
    bc = g.exact_bc()
    brute = brute_bc(g)
    assert np.allclose(bc, brute, rtol=0, atol=1e-9)


@pytest.mark.parametrize("n", [3, 6, 10])
def test_bidirected_path(n):
    arcs = [(i, i + 1, 3) for i in range(n - 1)] + [(i + 1, i, 3) for i in range(n - 1)]
    bc = graph(arcs).exact_bc()
    assert np.array_equal(bc, [2.0 * i * (n - 1 - i) for i in range(n)])


@pytest.mark.parametrize("n", [4, 7, 11])
def test_wrapped_chain_closed_form(n):
    """Chain 0->1->..->n-1 closed by n-1->0: every (s, t) has one path.  For
    s < t it runs along the chain (interior s+1..t-1); for s > t it wraps
    (interior s+1..n-1 and 0..t-1), so
    BC(i) = i(n-1-i) + i(i-1)/2 + (n-1-i)(n-2-i)/2."""
    bc = graph([(i, i + 1, 2) for i in range(n - 1)] + [(n - 1, 0, 5)]).exact_bc()
    i = np.arange(n)
    assert np.array_equal(bc, i * (n - 1 - i) + i * (i - 1) / 2 + (n - 1 - i) * (n - 2 - i) / 2)


@pytest.mark.parametrize("k", [2, 5])
def test_star(k):
    arcs = [(0, i, 4) for i in range(1, k + 1)] + [(i, 0, 4) for i in range(1, k + 1)]
    bc = graph(arcs).exact_bc()
    assert bc[0] == k * (k - 1) and not bc[1:].any()


@pytest.mark.parametrize("n", [3, 8])
def test_directed_cycle(n):
    bc = graph([(i, (i + 1) % n, 1) for i in range(n)]).exact_bc()
    assert np.array_equal(bc, np.full(n, (n - 1) * (n - 2) / 2))


def test_unit_grid_sigma_binomial():
    n, (s, d, w) = F.generate_grid(12, 12, 1, 1, road=False)
    g = O.Graph.from_arcs(s, d, w.astype(np.float64))
    dist, sig, order = g.sssp(0)
    r, c = np.divmod(np.arange(n), 12)
    assert np.array_equal(dist, r + c)
    assert all(sig[v] == comb(int(r[v] + c[v]), int(r[v])) for v in range(n))
    assert np.all(np.diff(dist[order]) >= 0)


# ---- brute force from the definition ----------------------------------------

def brute_bc(g):
    """BC(v) = sum over s != v != t of sigma_st(v) / sigma_st, from all-pairs
    distances (Floyd-Warshall) and path counts over the shortest-path DAG."""
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
        order = sorted((x for x in range(n) if D[s, x] < INF), key=lambda x: D[s, x])
        sig[s, s] = 1
        for x in order:
            if x == s:
                continue
            for u in range(n):
                for e in range(rp[u], rp[u + 1]):
                    if col[e] == x and D[s, u] + w[e] == D[s, x]:
                        sig[s, x] += sig[s, u]
    bc = np.zeros(n)
    for s, t, v in itertools.permutations(range(n), 3):
        if D[s, t] < INF and D[s, v] + D[v, t] == D[s, t]:
            bc[v] += sig[s, v] * sig[v, t] / sig[s, t]
    return bc


@pytest.mark.parametrize("seed", range(12))
def test_exact_bc_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    arcs = {}
    for _ in range(int(rng.integers(n, 3 * n))):
        a, b = rng.integers(0, n, 2)
        if a != b:
            arcs[(int(a), int(b))] = int(rng.integers(1, 4))   # small weights: many ties
    arcs[(n - 1, 0)] = arcs.get((n - 1, 0), 1)
    for v in range(n - 1):
        arcs.setdefault((v, v + 1), 2)
    g = graph([(a, b, c) for (a, b), c in arcs.items()])
    assert np.allclose(g.exact_bc(), brute_bc(g), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_single_source_dep_sums_to_exact(seed):
    rng = np.random.default_rng(100 + seed)
    n = 7
    arcs = {(v, (v + 1) % n): int(rng.integers(1, 3)) for v in range(n)}
    for _ in range(10):
        a, b = rng.integers(0, n, 2)
        if a != b:
            arcs[(int(a), int(b))] = int(rng.integers(1, 3))
    g = graph([(a, b, c) for (a, b), c in arcs.items()])
    total = sum(g.single_source_dep(s) for s in range(n))
    assert np.allclose(total, brute_bc(g), rtol=1e-12, atol=1e-12)


# ---- fixtures written by the reference's code --------------------------------

def test_generator_digests():
    import hashlib
    dig = json.load(open(os.path.join(GOLDEN, "graph_digests.json")))
    for cfg in (0, 1, 3):
        n, arcs = F.generate_config(cfg)
        p = f"/tmp/fbc_digest_{os.getpid()}_{cfg}.txt"
        F.write_edge_list(n, arcs, p)
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == dig[f"cfg{cfg}"]["sha256"]
        assert (n, len(arcs[0])) == (dig[f"cfg{cfg}"]["n"], dig[f"cfg{cfg}"]["m"])
        os.remove(p)


@pytest.fixture(scope="module")
def cfg0():
    n, (s, d, w) = F.generate_config(0)
    return O.Graph.from_arcs(s, d, w.astype(np.float64))


def test_cfg0_exact_matches_reference(cfg0):
    ref = np.load(os.path.join(GOLDEN, "cfg0_exact.npz"))["bc"]
    assert np.array_equal(cfg0.exact_bc(1), ref)


def test_louvain_matches_reference(cfg0):
    gold = np.load(os.path.join(GOLDEN, "louvain.npz"))
    comm, _ = cfg0.louvain([1, 2, 3, 4], 0.01)
    assert np.array_equal(comm, gold["cfg0"])
    n, (s, d, w) = F.generate_config(1)
    comm1, _ = O.Graph.from_arcs(s, d, w.astype(np.float64)).louvain([4, 3, 2, 1], 0.01)
    assert np.array_equal(comm1, gold["cfg1"])


def test_cluster_eval_matches_reference(cfg0):
    gold = np.load(os.path.join(GOLDEN, "cfg0_clusters.npz"))
    for i in range(3):
        verts = gold[f"verts{i}"]
        bc, b, L, C = cfg0.cluster_eval(verts)
        assert np.array_equal(b, gold[f"borders{i}"])
        assert np.array_equal(bc[verts], gold[f"bc{i}"])
        assert np.array_equal(L, gold[f"len{i}"]) and np.array_equal(C, gold[f"cnt{i}"])


def test_kmeans_matches_reference():
    gold = np.load(os.path.join(GOLDEN, "kmeans_cases.npz"))
    for t in range(24):
        c, cw = O.kmeans(int(gold[f"k{t}"][0]), gold[f"L{t}"], gold[f"C{t}"], gold[f"W{t}"])
        assert np.array_equal(c, gold[f"cent{t}"]) and np.array_equal(cw, gold[f"cw{t}"]), t


def test_kmeans_special_cases():
    """P = 1; all-equal rows; k = 1 picks the member nearest the mean."""
    c, cw = O.kmeans(1, np.array([[3.0, 1.0]]), np.array([[1, 1]], np.int32), [7])
    assert list(c) == [0] and list(cw) == [7]
    c, cw = O.kmeans(3, np.ones((5, 2)), np.ones((5, 2), np.int32), [1, 2, 3, 4, 5])
    assert list(c) == [0, 0, 0] and cw[0] == 15
    # mean 4.75: the member at 5 is nearest (0.0625 < 0.5625), and the loop
    # stops once the centroid repeats
    L = np.array([[0.0], [10.0], [4.0], [5.0]])
    c, cw = O.kmeans(1, L, np.zeros((4, 1), np.int32), [1, 1, 1, 1])
    assert list(c) == [3] and list(cw) == [4]


def test_pivot_sum_matches_reference(cfg0):
    gold = np.load(os.path.join(GOLDEN, "cfg0_ssb.npz"))
    acc = np.zeros(cfg0.n)
    for p, w in zip(gold["pivots"], gold["weights"]):
        acc += cfg0.single_source_dep(int(p)) * float(w)
    assert np.allclose(acc, gold["bc"], rtol=1e-12, atol=1e-9)


def test_cfg2_small_clusters_match_reference():
    """The oracle at full cfg2 size (1,048,576 vertices) against the
    reference's own runs (cfg2_reference.npz): on the three smallest
    communities of the reference partition, borders, intra BC (bit for bit)
    and VertexInfo rows (digests) equal DijkstraClusterEvaluator's, and the
    oracle's selection fixture (cfg2_oracle_pivots.npz) holds the pivots that
    the reference's k-means gives on their exact classes."""
    import hashlib
    ref = np.load(os.path.join(GOLDEN, "cfg2_reference.npz"))
    sel = np.load(os.path.join(GOLDEN, "cfg2_oracle_pivots.npz"))
    comm = ref["comm"].astype(np.int64)

    def sha(a):
        return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)

    assert np.array_equal(sha(comm.astype(np.int32)), sel["comm_sha"])
    n, (s, d, w) = F.generate_config(2)
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    for i, c in enumerate(ref["clusters"]):
        verts = np.where(comm == c)[0]
        bc, borders, L, C = og.cluster_eval(verts)
        assert np.array_equal(borders, ref[f"borders{i}"])
        assert np.array_equal(bc[verts], ref[f"intra{i}"])
        assert np.array_equal(sha(L.astype(np.uint32)), ref[f"len_sha{i}"])
        assert np.array_equal(sha(C.astype(np.int32)), ref[f"cnt_sha{i}"])
        exact, _, _ = O.select_pivots(verts, L, C, bc[verts], borders)
        assert len(exact) == int(ref[f"classes{i}"][0])
        expect, seen = [], set()
        for k, wt in zip(ref[f"km_cent{i}"], ref[f"km_w{i}"]):
            if int(exact[k]) not in seen:
                seen.add(int(exact[k]))
                expect.append((int(exact[k]), int(wt)))
        mine = sel["pivot_cluster"] == c
        assert list(zip(sel["pivots"][mine].tolist(), sel["weights"][mine].tolist())) == expect
