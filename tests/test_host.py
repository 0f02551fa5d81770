"""Host side of the product (no GPU): Louvain, pivot selection and k-means
through the C ABI of include/fbc_host.h, checked against the oracle and the
fixtures written by the reference's code."""
import os

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN


@pytest.mark.parametrize("cfg", [0, 1])
def test_louvain_bit_identical_to_reference(cfg):
    gold = np.load(os.path.join(GOLDEN, "louvain.npz"))[f"cfg{cfg}"]
    n, (s, d, w) = F.generate_config(cfg)
    comm, k = F.louvain(s, d, w, [4, 1, 3, 2], 0.01)
    assert np.array_equal(comm, gold) and k == gold.max() + 1


@pytest.mark.parametrize("seed", range(8))
def test_louvain_matches_oracle_on_random_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 300))
    m = int(rng.integers(n, 6 * n))
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    s = np.append(s, n - 1)          # vertex n-1 has an out-arc
    d = np.append(d, 0)
    w = rng.integers(1, 20, len(s)).astype(np.float64)   # repeats are summed by both
    seeds = list(rng.choice(1000, int(rng.integers(1, 5)), replace=False))
    precision = float(rng.choice([0.0, 0.01, 0.1]))
    comm, k = F.louvain(s, d, w, seeds, precision)
    ocomm, ok = O.Graph.from_arcs(s, d, w).louvain(seeds, precision)
    assert np.array_equal(comm, ocomm) and k == ok


def test_louvain_invariants():
    """Two disjoint cliques joined by one arc pair split into >= 2 communities;
    the result is independent of the seed order."""
    arcs = []
    for base in (0, 6):
        for a in range(6):
            for b in range(6):
                if a != b:
                    arcs.append((base + a, base + b, 5))
    arcs += [(5, 6, 1), (6, 5, 1)]
    s, d, w = (np.array(x) for x in zip(*arcs))
    c1, k1 = F.louvain(s, d, w, [7, 3], 0.01)
    c2, k2 = F.louvain(s, d, w, [3, 7], 0.01)
    assert np.array_equal(c1, c2) and k1 >= 2
    assert len(set(c1[:6])) == 1 and len(set(c1[6:])) == 1 and c1[0] != c1[6]


def test_select_pivots_known_answer():
    """test/brandes/VertexInfoPivotSelector.cpp:16-74"""
    L = np.array([[1, 2, 3], [2, 1, 3], [2, 3, 4], [4, 3, 5], [5, 1, 3]], float)
    C = np.array([[2, 1, 1], [2, 2, 1], [2, 1, 1], [2, 2, 1], [1, 1, 3]], np.int32)
    rows, w, Ln = F.select_pivots(L, C, [1, 2, 2, 1.5, 1], np.zeros(5, np.uint8))
    assert rows.tolist() == [0, 3, 4] and w.tolist() == [2, 2, 1]
    assert Ln.min(1).tolist() == [0] * 5


@pytest.mark.parametrize("seed", range(10))
def test_select_pivots_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    nv, nb = int(rng.integers(1, 200)), int(rng.integers(1, 6))
    base_L = rng.integers(0, 4, (max(nv // 3, 1), nb)).astype(float)
    base_C = rng.integers(0, 3, (max(nv // 3, 1), nb)).astype(np.int32)
    pick = rng.integers(0, len(base_L), nv)                  # many repeated rows
    shift = rng.integers(0, 3, nv)[:, None].astype(float)    # equal after normalize
    L = base_L[pick] + shift
    C = base_C[pick]
    bc = rng.integers(0, 4, nv).astype(float)                # BC ties
    border = (rng.random(nv) < 0.3).astype(np.uint8)
    rows, w, Ln = F.select_pivots(L, C, bc, border)
    verts = np.arange(nv, dtype=np.int32)
    opiv, ow, oL = O.select_pivots(verts, L, C, bc, verts[border == 1])
    assert rows.tolist() == opiv.tolist() and w.tolist() == ow.tolist()
    assert np.array_equal(Ln, oL)


def test_kmeans_matches_reference():
    gold = np.load(os.path.join(GOLDEN, "kmeans_cases.npz"))
    for t in range(24):
        c, cw = F.kmeans(int(gold[f"k{t}"][0]), gold[f"L{t}"], gold[f"C{t}"], gold[f"W{t}"])
        assert np.array_equal(c, gold[f"cent{t}"]) and np.array_equal(cw, gold[f"cw{t}"]), t


@pytest.mark.parametrize("seed", range(6))
def test_kmeans_matches_oracle(seed):
    rng = np.random.default_rng(50 + seed)
    P, nb = int(rng.integers(2, 300)), int(rng.integers(1, 12))
    k = int(rng.integers(1, P))
    L = rng.integers(0, 50, (P, nb)).astype(float)
    C = rng.integers(0, 4, (P, nb)).astype(np.int32)
    W = rng.integers(1, 9, P).astype(np.int32)
    assert all(np.array_equal(a, b) for a, b in zip(F.kmeans(k, L, C, W), O.kmeans(k, L, C, W)))


def test_host_abi_rejects_bad_arguments():
    lib = F.lib()
    z = np.zeros(1, np.int32)
    assert lib.fbc_kmeans(0, 1, 1, np.zeros(1), z, z, 0.0, 100, z, z) == -1
    assert lib.fbc_louvain(1, z, z, np.ones(1), np.zeros(0, np.uint64), 0, 0.01, z) == -1
