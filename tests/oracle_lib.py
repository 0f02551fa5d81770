"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure.

The oracle restates the reference's algorithms (see oracle/fbc_oracle.h); the
tests use it as the checker for the HIP path.  The product never imports this
module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_DRIVER_FAST = os.path.join(ROOT, "oracle", "_ref", "ref_driver_fast")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_graph_from_file.restype = P
        L.orc_graph_from_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_graph_from_arcs.restype = P
        L.orc_graph_from_arcs.argtypes = [ctypes.c_uint64, u32p, u32p, f64p, ctypes.c_char_p, ctypes.c_int]
        L.orc_graph_free.argtypes = [P]
        L.orc_graph_n.argtypes = [P]
        L.orc_graph_edges.argtypes = [P]
        L.orc_graph_edges.restype = ctypes.c_longlong
        L.orc_graph_csr.argtypes = [P, u64p, u32p, f64p]
        L.orc_graph_bstar.argtypes = [P, ctypes.c_int, ctypes.c_int, i32p, f64p]
        L.orc_sssp.argtypes = [P, ctypes.c_int, f64p, f64p, i32p]
        L.orc_single_source_dep.argtypes = [P, ctypes.c_int, f64p]
        L.orc_exact_bc.argtypes = [P, ctypes.c_int, f64p]
        L.orc_exact_bc_range.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p]
        L.orc_subgraph.argtypes = [P, ctypes.c_int, i32p, ctypes.POINTER(ctypes.c_int), i32p,
                                   ctypes.POINTER(ctypes.c_longlong)]
        L.orc_cluster_eval.argtypes = [P, ctypes.c_int, i32p, f64p, ctypes.POINTER(ctypes.c_int),
                                       i32p, f64p, i32p]
        L.orc_select_pivots.argtypes = [ctypes.c_int, i32p, ctypes.c_int, f64p, i32p, f64p,
                                        ctypes.c_int, i32p, i32p, i32p]
        L.orc_kmeans.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p, i32p, i32p,
                                 ctypes.c_double, ctypes.c_int, i32p, i32p]
        L.orc_louvain.argtypes = [P, ctypes.c_int, u64p, ctypes.c_double, i32p]
        L.orc_clustered_bc.argtypes = [P, ctypes.c_int, u64p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, f64p, i32p, f64p, i32p, i32p, i32p,
                                       ctypes.POINTER(ctypes.c_double)]
        L.orc_clustered_pivots.argtypes = [P, ctypes.c_int, u64p, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, i32p, f64p, i32p, i32p, i32p]
        L.orc_clustered_bc_partition.argtypes = [P, i32p, ctypes.c_double, ctypes.c_int, f64p, f64p, i32p, i32p,
                                                 i32p]
        L.orc_kmeans_pivots.argtypes = [ctypes.c_int, ctypes.c_int, f64p, i32p, i32p, ctypes.c_double, i32p, i32p]
        L.orc_clustered_bc_sampled.argtypes = [P, ctypes.c_int, u64p, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, ctypes.c_int, f64p, f64p]
        _lib = L
    return _lib


class Graph:
    """Oracle graph (DirectedWeightedGraph semantics)."""

    def __init__(self, handle):
        self.h = handle
        self.n = lib().orc_graph_n(handle)
        self.edges = lib().orc_graph_edges(handle)

    @classmethod
    def from_file(cls, path):
        err = ctypes.create_string_buffer(256)
        h = lib().orc_graph_from_file(path.encode(), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    @classmethod
    def from_arcs(cls, src, dst, w):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.float64)
        err = ctypes.create_string_buffer(256)
        h = lib().orc_graph_from_arcs(len(src), src, dst, w, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_graph_free(self.h)
            self.h = None

    def csr(self):
        rp = np.zeros(self.n + 1, np.uint64)
        col = np.zeros(max(self.edges, 1), np.uint32)
        w = np.zeros(max(self.edges, 1), np.float64)
        lib().orc_graph_csr(self.h, rp, col, w)
        return rp, col[: self.edges], w[: self.edges]

    def bstar(self, v):
        k = lib().orc_graph_bstar(self.h, v, 0, np.zeros(1, np.int32), np.zeros(1))
        s = np.zeros(max(k, 1), np.int32)
        w = np.zeros(max(k, 1))
        lib().orc_graph_bstar(self.h, v, k, s, w)
        return dict(zip(s[:k].tolist(), w[:k].tolist()))

    def sssp(self, src):
        d = np.zeros(self.n)
        s = np.zeros(self.n)
        o = np.zeros(self.n, np.int32)
        k = lib().orc_sssp(self.h, src, d, s, o)
        return d, s, o[:k]

    def single_source_dep(self, src):
        dep = np.zeros(self.n)
        assert lib().orc_single_source_dep(self.h, src, dep) == 0
        return dep

    def exact_bc(self, threads=1):
        bc = np.zeros(self.n)
        lib().orc_exact_bc(self.h, threads, bc)
        return bc

    def exact_bc_range(self, s0, s1, threads=1):
        bc = np.zeros(self.n)
        lib().orc_exact_bc_range(self.h, s0, s1, threads, bc)
        return bc

    def subgraph(self, verts):
        v = np.ascontiguousarray(verts, np.int32)
        nb = ctypes.c_int()
        e = ctypes.c_longlong()
        b = np.zeros(max(len(v), 1), np.int32)
        assert lib().orc_subgraph(self.h, len(v), v, ctypes.byref(nb), b, ctypes.byref(e)) == 0
        return b[: nb.value], e.value

    def cluster_eval(self, verts, bc=None):
        v = np.ascontiguousarray(verts, np.int32)
        bc = np.zeros(self.n) if bc is None else np.ascontiguousarray(bc, np.float64).copy()
        nb = ctypes.c_int()
        b = np.zeros(max(len(v), 1), np.int32)
        L = np.zeros(max(len(v) * len(v), 1))
        C = np.zeros(max(len(v) * len(v), 1), np.int32)
        assert lib().orc_cluster_eval(self.h, len(v), v, bc, ctypes.byref(nb), b, L, C) == 0
        k = nb.value
        return bc, b[:k], L[: len(v) * k].reshape(len(v), k), C[: len(v) * k].reshape(len(v), k)

    def louvain(self, seeds, precision=0.01):
        s = np.ascontiguousarray(seeds, np.uint64)
        comm = np.zeros(self.n, np.int32)
        k = lib().orc_louvain(self.h, len(s), s, precision, comm)
        return comm, k

    def clustered_bc(self, seeds, precision=0.01, kfrac=0.0, threads=1):
        s = np.ascontiguousarray(seeds, np.uint64)
        n = self.n
        bc = np.zeros(n)
        comm = np.zeros(n, np.int32)
        intra = np.zeros(n)
        piv = np.zeros(n, np.int32)
        w = np.zeros(n, np.int32)
        pc = np.zeros(n, np.int32)
        te = ctypes.c_double()
        k = lib().orc_clustered_bc(self.h, len(s), s, precision, kfrac, threads, bc, comm, intra,
                                   piv, w, pc, ctypes.byref(te))
        assert k >= 0
        return dict(bc=bc, comm=comm, intra=intra, pivots=piv[:k], weights=w[:k],
                    pivot_cluster=pc[:k], te=te.value)


def clustered_bc_partition(g, comm, kfrac=0.0, threads=1):
    """Clustered BC with a caller's partition (community per vertex)."""
    n = g.n
    c = np.ascontiguousarray(comm, np.int32)
    assert len(c) == n
    bc, intra = np.zeros(n), np.zeros(n)
    piv, w, pc = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
    k = lib().orc_clustered_bc_partition(g.h, c, kfrac, threads, bc, intra, piv, w, pc)
    assert k >= 0
    return dict(bc=bc, intra=intra, pivots=piv[:k], weights=w[:k], pivot_cluster=pc[:k])


def kmeans_pivots(vi_len, vi_cnt, weights, kfrac):
    """KMeansPivotSelector after the exact selector: k-means over the P rows,
    then the first occurrence of each repeated centroid kept."""
    L = np.ascontiguousarray(vi_len, np.float64)
    C = np.ascontiguousarray(vi_cnt, np.int32)
    P, nb = L.shape
    w = np.ascontiguousarray(weights, np.int32)
    piv, pw = np.zeros(P, np.int32), np.zeros(P, np.int32)
    k = lib().orc_kmeans_pivots(P, nb, L.reshape(-1), C.reshape(-1), w, kfrac, piv, pw)
    assert k >= 0
    return piv[:k], pw[:k]


def clustered_pivots(g, seeds, precision=0.01, kfrac=0.0, threads=1):
    """Partition, intra-cluster BC and pivots of a clustered run (no pivot phase)."""
    s = np.ascontiguousarray(seeds, np.uint64)
    n = g.n
    comm = np.zeros(n, np.int32)
    intra = np.zeros(n)
    piv = np.zeros(n, np.int32)
    w = np.zeros(n, np.int32)
    pc = np.zeros(n, np.int32)
    k = lib().orc_clustered_pivots(g.h, len(s), s, precision, kfrac, threads, comm, intra, piv, w, pc)
    assert k >= 0
    return dict(comm=comm, intra=intra, pivots=piv[:k], weights=w[:k], pivot_cluster=pc[:k])


def select_pivots(verts, vi_len, vi_cnt, bc_of_vert, borders):
    verts = np.ascontiguousarray(verts, np.int32)
    L = np.ascontiguousarray(vi_len, np.float64).copy()
    C = np.ascontiguousarray(vi_cnt, np.int32)
    nb = L.shape[1] if L.ndim == 2 else 0
    bc = np.ascontiguousarray(bc_of_vert, np.float64)
    b = np.ascontiguousarray(borders, np.int32)
    piv = np.zeros(max(len(verts), 1), np.int32)
    w = np.zeros(max(len(verts), 1), np.int32)
    k = lib().orc_select_pivots(len(verts), verts, nb, L.reshape(-1), C.reshape(-1), bc, len(b),
                                b if len(b) else np.zeros(1, np.int32), piv, w)
    return piv[:k], w[:k], L


def kmeans(k, vi_len, vi_cnt, weights, min_variance=0.0, max_iteration=100):
    L = np.ascontiguousarray(vi_len, np.float64)
    C = np.ascontiguousarray(vi_cnt, np.int32)
    P, nb = L.shape
    w = np.ascontiguousarray(weights, np.int32)
    c = np.zeros(k, np.int32)
    cw = np.zeros(k, np.int32)
    r = lib().orc_kmeans(k, P, nb, L.reshape(-1), C.reshape(-1), w, min_variance, max_iteration, c, cw)
    assert r == k
    return c, cw
