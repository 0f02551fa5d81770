"""ctypes bindings of the MI355X betweenness library (include/fbc_gpu.h).

The library is built in-tree (fast-bc_amd/lib/libfastbc_amd.so, by
`make -C fast-bc_amd` or `__graft_entry__.build()`).  Nothing here falls back
to a CPU implementation: a missing library raises at load time and every
failing call raises FbcError with the library's message.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfastbc_amd.so")
GEN_PATH = os.path.join(PKG_DIR, "lib", "libfbcgen.so")
FBC_BIN = os.path.join(PKG_DIR, "bin", "fbc")

PREC_F64 = 0
PREC_F32C = 1

# Every symbol include/fbc_gpu.h declares.
EXPORTS = (
    "fbc_create", "fbc_destroy", "fbc_last_error", "fbc_set_tuning", "fbc_graph_upload",
    "fbc_graph_free", "fbc_sssp_sigma", "fbc_single_source_dep", "fbc_dep_batch",
    "fbc_exact_bc", "fbc_clusters_create", "fbc_clusters_borders", "fbc_clusters_edges",
    "fbc_clusters_free", "fbc_cluster_eval", "fbc_pivot_phase", "fbc_last_stats",
    "fbc_last_counters", "fbc_kmeans_batch", "fbc_pivot_phase_plain", "fbc_set_option",
    "fbc_cluster_eval_resident", "fbc_cluster_rows", "fbc_cluster_classes", "fbc_cluster_kmeans",
    "fbc_cluster_class_features", "fbc_dep_batch_noalloc",
)

# Every symbol include/fbc_host.h declares.
HOST_EXPORTS = ("fbc_louvain", "fbc_select_pivots", "fbc_kmeans", "fbc_clustered_bc", "fbc_clustered_bc_partition",
                "fbc_exact_bc_arcs", "fbc_host_release", "fbc_host_last_error", "fbc_host_last_tuning",
                "fbc_host_last_tuning_ex")

COUNTER_NAMES = ("rounds", "windows", "visits", "relax", "improve", "farpush", "levels",
                 "ns_dist", "ns_sigma", "ns_dep", "sigma_wide",
                 "plain_sources", "plain_fallback_sources", "plain_refusal",
                 "delta", "half_team_max", "tuned", "sigma_inexact", "lds_view")

ERRORS = {
    -1: "FBC_E_ARG", -2: "FBC_E_HIP", -3: "FBC_E_OOM", -4: "FBC_E_DEVICE",
    -5: "FBC_E_SIGMA_OVERFLOW", -6: "FBC_E_DIST_OVERFLOW", -7: "FBC_E_UNSUPPORTED",
}


class FbcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None
_gen = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def lib():
    """Load libfastbc_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C fast-bc_amd`")
        L = ctypes.CDLL(LIB_PATH)
        L.fbc_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.fbc_destroy.argtypes = [_vp]
        L.fbc_last_error.argtypes = [_vp]
        L.fbc_last_error.restype = ctypes.c_char_p
        L.fbc_set_tuning.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
        L.fbc_graph_upload.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint64, _u64p, _u32p, _u32p,
                                       ctypes.POINTER(_vp)]
        L.fbc_graph_free.argtypes = [_vp]
        L.fbc_sssp_sigma.argtypes = [_vp, _vp, _u32p, ctypes.c_uint32, _u64p, _u64p]
        L.fbc_single_source_dep.argtypes = [_vp, _vp, ctypes.c_uint32, ctypes.c_int, _f64p]
        L.fbc_dep_batch.argtypes = [_vp, _vp, _u32p, ctypes.c_uint32, ctypes.c_int, _f64p]
        L.fbc_dep_batch_noalloc.argtypes = [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_int, _vp]
        L.fbc_exact_bc.argtypes = [_vp, _vp, ctypes.c_int, _f64p]
        L.fbc_clusters_create.argtypes = [_vp, _vp, ctypes.c_uint32, _u32p, _u32p, ctypes.POINTER(_vp),
                                          _u32p]
        L.fbc_clusters_borders.argtypes = [_vp, _u32p]
        L.fbc_clusters_edges.argtypes = [_vp, ctypes.c_uint32]
        L.fbc_clusters_edges.restype = ctypes.c_int64
        L.fbc_clusters_free.argtypes = [_vp]
        L.fbc_cluster_eval.argtypes = [_vp, _vp, _f64p, _u32p, _i32p]
        L.fbc_cluster_eval_resident.argtypes = [_vp, _vp, _f64p]
        L.fbc_cluster_rows.argtypes = [_vp, _vp, _u32p, _i32p]
        L.fbc_cluster_classes.argtypes = [_vp, _vp, _f64p, _u32p, _u32p, _u32p]
        L.fbc_cluster_kmeans.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _u32p,
                                         _u32p, _i32p]
        L.fbc_cluster_class_features.argtypes = [_vp, _vp, ctypes.c_uint32, _i32p]
        L.fbc_pivot_phase.argtypes = [_vp, _vp, _u32p, _u32p, _u32p, ctypes.c_uint32, _f64p,
                                      ctypes.c_int, _f64p]
        L.fbc_pivot_phase_plain.argtypes = [_vp, _vp, _u32p, _u32p, _u32p, ctypes.c_uint32, _f64p, _f64p]
        L.fbc_set_option.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int64]
        L.fbc_last_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.fbc_last_counters.argtypes = [_vp, _u64p, ctypes.c_int]
        L.fbc_kmeans_batch.argtypes = [_vp, ctypes.c_uint32, _u32p, _u32p, _u32p,
                                       np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                       np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), ctypes.c_double,
                                       ctypes.c_uint32, _u32p, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        L.fbc_louvain.argtypes = [ctypes.c_uint64, i32p, i32p, _f64p, _u64p, ctypes.c_int, ctypes.c_double,
                                  i32p]
        L.fbc_select_pivots.argtypes = [ctypes.c_int, ctypes.c_int, _f64p, i32p, _f64p, u8p, i32p, i32p]
        L.fbc_kmeans.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, i32p, i32p, ctypes.c_double,
                                 ctypes.c_int, i32p, i32p]
        L.fbc_clustered_bc.argtypes = [ctypes.c_int, ctypes.c_uint64, i32p, i32p, _f64p, _u64p, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _f64p, i32p,
                                       _f64p, _u32p, _u32p, _u32p, _f64p]
        L.fbc_clustered_bc_partition.argtypes = [ctypes.c_int, ctypes.c_uint64, i32p, i32p, _f64p, i32p,
                                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, _f64p, _f64p, _u32p,
                                                 _u32p, _u32p, _f64p]
        L.fbc_exact_bc_arcs.argtypes = [ctypes.c_int, ctypes.c_uint64, i32p, i32p, _f64p, ctypes.c_int, _f64p]
        L.fbc_host_release.argtypes = []
        L.fbc_host_release.restype = None
        L.fbc_host_last_tuning.argtypes = [_u64p]
        L.fbc_host_last_tuning_ex.argtypes = [_u64p, ctypes.c_int]
        L.fbc_host_last_error.argtypes = []
        L.fbc_host_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Context:
    """A device context (fbc_create)."""

    def __init__(self, device=0):
        h = _vp()
        rc = lib().fbc_create(device, ctypes.byref(h))
        if rc:
            raise FbcError(rc, "fbc_create failed (a gfx950 device is required)")
        self.h = h

    def close(self):
        if self.h:
            lib().fbc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc:
            raise FbcError(rc, lib().fbc_last_error(self.h).decode())

    def set_tuning(self, bucket_delta=0, slots=0):
        self.check(lib().fbc_set_tuning(self.h, bucket_delta, slots))

    def set_option(self, name, value):
        """Named tuning option (fbc_set_option): "delta", "slots", "frontier_sort", "level_sort",
        "relabel", "scratch_mb", "half_team_max", "class_key_bits", "autotune", "ballot_append",
        "lds_view", "tune_log", "kmeans_budget_mb"."""
        self.check(lib().fbc_set_option(self.h, name.encode(), int(value)))

    def last_stats(self):
        ms, te = ctypes.c_double(), ctypes.c_double()
        lib().fbc_last_stats(self.h, ctypes.byref(ms), ctypes.byref(te))
        return ms.value, te.value

    def last_counters(self):
        c = np.zeros(len(COUNTER_NAMES), np.uint64)
        lib().fbc_last_counters(self.h, c, len(COUNTER_NAMES))
        return dict(zip(COUNTER_NAMES, (int(x) for x in c)))

    def graph(self, n, row_ptr, col, w):
        return Graph(self, n, row_ptr, col, w)

    def kmeans_batch(self, sets, min_variance=0.0, max_iteration=100):
        """PlusPlusKMeans on the GPU for a list of (k, features [P x D], weights).

        Features are integer VertexInfo rows (ValueError for other values)."""
        np_ = _u32([len(f) for _, f, _ in sets])
        nf = _u32([f.shape[1] for _, f, _ in sets])
        kk = _u32([k for k, _, _ in sets])
        flat = np.concatenate([np.asarray(f).reshape(-1) for _, f, _ in sets])
        feat = np.ascontiguousarray(flat, np.int32)
        if not np.array_equal(feat, flat):
            raise ValueError("k-means features must be integers within the int32 range")
        w = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for _, _, x in sets]), np.int32)
        oc = np.zeros(int(kk.sum()), np.uint32)
        ow = np.zeros(int(kk.sum()), np.int32)
        self.check(lib().fbc_kmeans_batch(self.h, len(sets), np_, nf, kk, feat, w, min_variance,
                                          max_iteration, oc, ow))
        out, o = [], 0
        for k in kk:
            out.append((oc[o:o + k].astype(np.int64), ow[o:o + k].astype(np.int64)))
            o += int(k)
        return out


class Graph:
    """A device-resident CSR graph (fbc_graph_upload)."""

    def __init__(self, ctx, n, row_ptr, col, w):
        self.ctx = ctx
        self.n = int(n)
        rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        col = _u32(col)
        w = _u32(w)
        self.m = int(rp[-1])
        h = _vp()
        ctx.check(lib().fbc_graph_upload(ctx.h, self.n, self.m, rp, col, w, ctypes.byref(h)))
        self.h = h

    @classmethod
    def from_arcs(cls, ctx, src, dst, w, n=None):
        """Arcs must be sorted by (src, dst) without repeats."""
        src = _u32(src)
        n = int(src.max()) + 1 if n is None else n
        rp = np.zeros(n + 1, np.uint64)
        np.add.at(rp, src.astype(np.int64) + 1, 1)
        rp = np.cumsum(rp).astype(np.uint64)
        return cls(ctx, n, rp, dst, w)

    def close(self):
        if self.h:
            lib().fbc_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sssp_sigma(self, sources):
        s = _u32(sources)
        d = np.zeros(len(s) * self.n, np.uint64)
        sg = np.zeros(len(s) * self.n, np.uint64)
        self.ctx.check(lib().fbc_sssp_sigma(self.ctx.h, self.h, s, len(s), d, sg))
        return d.reshape(len(s), self.n), sg.reshape(len(s), self.n)

    def single_source_dep(self, src, prec=PREC_F64):
        dep = np.zeros(self.n)
        self.ctx.check(lib().fbc_single_source_dep(self.ctx.h, self.h, int(src), prec, dep))
        return dep

    def dep_batch(self, sources, prec=PREC_F64):
        s = _u32(sources)
        dep = np.zeros(len(s) * self.n)
        self.ctx.check(lib().fbc_dep_batch(self.ctx.h, self.h, s, len(s), prec, dep))
        return dep.reshape(len(s), self.n)

    def exact_bc(self, prec=PREC_F64):
        bc = np.zeros(self.n)
        self.ctx.check(lib().fbc_exact_bc(self.ctx.h, self.h, prec, bc))
        return bc

    def clusters(self, comm_of=None, communities=None):
        return Clusters(self, comm_of=comm_of, communities=communities)


class Clusters:
    """Cluster views (fbc_clusters_create) of a vertex partition."""

    def __init__(self, graph, comm_of=None, communities=None):
        self.graph = graph
        ctx = graph.ctx
        if communities is None:
            comm_of = np.asarray(comm_of, np.int64)
            ncl = int(comm_of.max()) + 1 if len(comm_of) else 0
            order = np.argsort(comm_of, kind="stable")
            counts = np.bincount(comm_of, minlength=ncl)
            ptr = np.zeros(ncl + 1, np.uint32)
            ptr[1:] = np.cumsum(counts)
            verts = _u32(order)
        else:
            ncl = len(communities)
            ptr = np.zeros(ncl + 1, np.uint32)
            ptr[1:] = np.cumsum([len(c) for c in communities])
            verts = _u32(np.concatenate([np.asarray(c) for c in communities]) if ncl else [])
        self.ncl = ncl
        self.ptr = ptr
        self.verts = verts
        self.border_ptr = np.zeros(ncl + 1, np.uint32)
        h = _vp()
        ctx.check(lib().fbc_clusters_create(ctx.h, graph.h, ncl, ptr, verts, ctypes.byref(h),
                                             self.border_ptr))
        self.h = h
        self.borders_flat = np.zeros(max(int(self.border_ptr[-1]), 1), np.uint32)
        ctx.check(lib().fbc_clusters_borders(self.h, self.borders_flat))
        nb = np.diff(self.border_ptr.astype(np.int64))
        nc = np.diff(ptr.astype(np.int64))
        self.vi_off = np.zeros(ncl + 1, np.int64)
        self.vi_off[1:] = np.cumsum(nc * nb)

    def close(self):
        if self.h:
            lib().fbc_clusters_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def members(self, c):
        return self.verts[self.ptr[c]:self.ptr[c + 1]]

    def borders(self, c):
        return self.borders_flat[self.border_ptr[c]:self.border_ptr[c + 1]]

    def edges(self, c):
        return int(lib().fbc_clusters_edges(self.h, c))

    def evaluate(self, intra_bc=None):
        """Returns (intra_bc, vi_len, vi_cnt) with VertexInfo rows flattened."""
        n = self.graph.n
        bc = np.zeros(n) if intra_bc is None else np.ascontiguousarray(intra_bc, np.float64).copy()
        tot = max(int(self.vi_off[-1]), 1)
        L = np.zeros(tot, np.uint32)
        C = np.zeros(tot, np.int32)
        self.graph.ctx.check(lib().fbc_cluster_eval(self.graph.ctx.h, self.h, bc, L, C))
        return bc, L, C

    def evaluate_resident(self, intra_bc=None):
        """fbc_cluster_eval_resident: intra-cluster BC; the VertexInfo rows stay on the device."""
        n = self.graph.n
        bc = np.zeros(n) if intra_bc is None else np.ascontiguousarray(intra_bc, np.float64).copy()
        self.graph.ctx.check(lib().fbc_cluster_eval_resident(self.graph.ctx.h, self.h, bc))
        return bc

    def rows(self):
        """The resident VertexInfo rows (fbc_cluster_rows), flattened like evaluate()'s."""
        tot = max(int(self.vi_off[-1]), 1)
        L = np.zeros(tot, np.uint32)
        C = np.zeros(tot, np.int32)
        self.graph.ctx.check(lib().fbc_cluster_rows(self.graph.ctx.h, self.h, L, C))
        return L, C

    def classes(self, intra_bc):
        """fbc_cluster_classes: (cls_ptr, pivot row within its cluster, class size)."""
        cap = max(int(self.ptr[-1]), 1)
        ptr = np.zeros(self.ncl + 1, np.uint32)
        row = np.zeros(cap, np.uint32)
        w = np.zeros(cap, np.uint32)
        self.graph.ctx.check(lib().fbc_cluster_classes(self.graph.ctx.h, self.h,
                                                         np.ascontiguousarray(intra_bc, np.float64), ptr, row, w))
        t = int(ptr[-1])
        return ptr, row[:t].copy(), w[:t].copy()

    def kmeans(self, kfrac, min_variance=0.0, max_iteration=100):
        """fbc_cluster_kmeans over the last classes: (kptr, centroids, weights)."""
        cap = max(int(self.ptr[-1]), 1)
        kptr = np.zeros(self.ncl + 1, np.uint32)
        cent = np.zeros(cap, np.uint32)
        cw = np.zeros(cap, np.int32)
        self.graph.ctx.check(lib().fbc_cluster_kmeans(self.graph.ctx.h, self.h, kfrac, min_variance, max_iteration,
                                                        kptr, cent, cw))
        t = int(kptr[-1])
        return kptr, cent[:t].copy(), cw[:t].copy()

    def class_features(self, c, nclasses):
        """fbc_cluster_class_features: [nclasses x 2 * borders] int32 k-means points of cluster c."""
        nb = int(self.border_ptr[c + 1] - self.border_ptr[c])
        f = np.zeros(max(nclasses * 2 * nb, 1), np.int32)
        self.graph.ctx.check(lib().fbc_cluster_class_features(self.graph.ctx.h, self.h, c, f))
        return f[:nclasses * 2 * nb].reshape(nclasses, 2 * nb)

    def vi_rows(self, flat, c):
        nb = int(self.border_ptr[c + 1] - self.border_ptr[c])
        nc = int(self.ptr[c + 1] - self.ptr[c])
        return flat[self.vi_off[c]:self.vi_off[c] + nc * nb].reshape(nc, nb)

    def pivot_phase(self, pivots, weights, pivot_cluster, intra_bc, bc, prec=PREC_F64, plain=False):
        """bc + the pivot loop's terms.  plain=True: one cooperative launch per
        pivot, FP64 (the baseline form); otherwise all pivots in one launch."""
        out = np.ascontiguousarray(bc, np.float64).copy()
        p = _u32(pivots)
        args = (self.graph.ctx.h, self.h, p, _u32(weights), _u32(pivot_cluster), len(p),
                np.ascontiguousarray(intra_bc, np.float64))
        if plain:
            if prec != PREC_F64:
                raise ValueError("the plain form computes in FP64")
            self.graph.ctx.check(lib().fbc_pivot_phase_plain(*args, out))
        else:
            self.graph.ctx.check(lib().fbc_pivot_phase(*args, prec, out))
        return out


# ---- host engine (include/fbc_host.h) ---------------------------------------------

def _arcs_i32(src, dst, w):
    return (np.ascontiguousarray(src, np.int32), np.ascontiguousarray(dst, np.int32),
            np.ascontiguousarray(w, np.float64))


def _host_check(rc, what):
    if rc < 0:
        raise FbcError(rc, f"{what}: {lib().fbc_host_last_error().decode()}")
    return rc


def louvain(src, dst, w, seeds, precision=0.01):
    """Community of every vertex (LouvainGraphPartition::partitionGraph)."""
    s, d, ww = _arcs_i32(src, dst, w)
    n = int(s.max()) + 1 if len(s) else 0
    comm = np.zeros(max(n, 1), np.int32)
    k = _host_check(lib().fbc_louvain(len(s), s, d, ww, np.ascontiguousarray(seeds, np.uint64), len(seeds),
                                      precision, comm), "fbc_louvain")
    return comm[:n], k


def select_pivots(vi_len, vi_cnt, bc_row, is_border):
    """VertexInfoPivotSelector on one cluster's rows; returns (rows, weights, normalized len)."""
    L = np.ascontiguousarray(vi_len, np.float64).copy()
    nv, nb = L.shape
    C = np.ascontiguousarray(vi_cnt, np.int32)
    rows = np.zeros(max(nv, 1), np.int32)
    wts = np.zeros(max(nv, 1), np.int32)
    k = _host_check(lib().fbc_select_pivots(nv, nb, L.reshape(-1), C.reshape(-1),
                                            np.ascontiguousarray(bc_row, np.float64),
                                            np.ascontiguousarray(is_border, np.uint8), rows, wts),
                    "fbc_select_pivots")
    return rows[:k], wts[:k], L


def kmeans(k, vi_len, vi_cnt, weights, min_variance=0.0, max_iteration=100):
    L = np.ascontiguousarray(vi_len, np.float64)
    P, nb = L.shape
    c = np.zeros(k, np.int32)
    cw = np.zeros(k, np.int32)
    _host_check(lib().fbc_kmeans(k, P, nb, L.reshape(-1), np.ascontiguousarray(vi_cnt, np.int32).reshape(-1),
                                 np.ascontiguousarray(weights, np.int32), min_variance, max_iteration, c, cw),
                "fbc_kmeans")
    return c, cw


PHASES = ("louvain_s", "clusters_s", "pivot_selection_s", "pivot_phase_s", "total_s",
          "cluster_te", "pivot_te", "pivot_kernel_ms")


FORM_BATCHED, FORM_PLAIN, FORM_SELECT_ONLY = 0, 1, 2


def clustered_bc(src, dst, w, seeds, precision=0.01, kfrac=0.0, prec=PREC_F64, device=0,
                 form=FORM_BATCHED, partition=None):
    """End-to-end ClusteredBrandeBC::computeBC on the GPU; returns a dict.
    form: FORM_BATCHED (all pivots in one launch), FORM_PLAIN (one cooperative
    launch per pivot, FP64) or FORM_SELECT_ONLY (stop after pivot selection;
    bc is then the intra-cluster BC).  partition: a community index per
    vertex in place of Louvain (negative: in no community; seeds and
    precision are then unused)."""
    s, d, ww = _arcs_i32(src, dst, w)
    n = int(s.max()) + 1
    bc = np.zeros(n)
    comm = np.zeros(n, np.int32)
    intra = np.zeros(n)
    piv = np.zeros(n, np.uint32)
    pw = np.zeros(n, np.uint32)
    pc = np.zeros(n, np.uint32)
    ph = np.zeros(8)
    if partition is not None:
        part = np.ascontiguousarray(partition, np.int32)
        if part.shape != (n,):
            raise ValueError(f"partition needs one entry per vertex ({n})")
        k = _host_check(lib().fbc_clustered_bc_partition(device, len(s), s, d, ww, part, kfrac, prec, form, bc,
                                                         intra, piv, pw, pc, ph), "fbc_clustered_bc_partition")
        # communities as the library numbers them: non-empty indices, ascending
        u = np.unique(part[part >= 0])
        comm = np.where(part >= 0, np.searchsorted(u, part), -1).astype(np.int32)
    else:
        k = _host_check(lib().fbc_clustered_bc(device, len(s), s, d, ww, np.ascontiguousarray(seeds, np.uint64),
                                               len(seeds), precision, kfrac, prec, form, bc, comm, intra, piv, pw,
                                               pc, ph),
                        "fbc_clustered_bc")
    tun = np.zeros(8, np.uint64)
    lib().fbc_host_last_tuning_ex(tun, len(tun))
    t = [int(x) for x in tun]
    return dict(bc=bc, comm=comm, intra=intra, pivots=piv[:k], weights=pw[:k], pivot_cluster=pc[:k],
                phases=dict(zip(PHASES, ph.tolist())),
                tuning={"cluster": dict(zip(("delta", "half_team_max", "how", "lds_view"), t[:4])),
                        "pivot": dict(zip(("delta", "half_team_max", "how"), t[4:7]))})


def host_release():
    """Release the device contexts and the resident graph the host engine keeps."""
    lib().fbc_host_release()


def exact_bc_arcs(src, dst, w, prec=PREC_F64, device=0):
    s, d, ww = _arcs_i32(src, dst, w)
    bc = np.zeros(int(s.max()) + 1)
    _host_check(lib().fbc_exact_bc_arcs(device, len(s), s, d, ww, prec, bc), "fbc_exact_bc_arcs")
    return bc


# ---- generators ---------------------------------------------------------------

class _Arcs(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint64),
                ("src", ctypes.POINTER(ctypes.c_uint32)), ("dst", ctypes.POINTER(ctypes.c_uint32)),
                ("w", ctypes.POINTER(ctypes.c_uint32))]


def _genlib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build it with `make -C fast-bc_amd`")
        G = ctypes.CDLL(GEN_PATH)
        G.fbcgen_config.argtypes = [ctypes.c_int, ctypes.POINTER(_Arcs)]
        G.fbcgen_grid.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_int, ctypes.POINTER(_Arcs)]
        G.fbcgen_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_Arcs)]
        G.fbcgen_write.argtypes = [ctypes.POINTER(_Arcs), ctypes.c_char_p]
        G.fbcgen_free.argtypes = [ctypes.POINTER(_Arcs)]
        _gen = G
    return _gen


def _take(a):
    m = a.m
    out = (np.ctypeslib.as_array(a.src, (m,)).copy(), np.ctypeslib.as_array(a.dst, (m,)).copy(),
           np.ctypeslib.as_array(a.w, (m,)).copy())
    n = a.n
    _genlib().fbcgen_free(ctypes.byref(a))
    return n, out


def generate_config(cfg):
    """(n, (src, dst, w)) of benchmark workload cfg 0..4 (see gen/fbc_gen.h)."""
    a = _Arcs()
    rc = _genlib().fbcgen_config(cfg, ctypes.byref(a))
    if rc:
        raise ValueError(f"fbcgen_config({cfg}) failed: {rc}")
    return _take(a)


def generate_grid(rows, cols, wmax, seed, road=True):
    a = _Arcs()
    rc = _genlib().fbcgen_grid(rows, cols, wmax, seed, int(road), ctypes.byref(a))
    if rc:
        raise ValueError(f"fbcgen_grid failed: {rc}")
    return _take(a)


def generate_rmat(scale, edge_factor, a=0.57, b=0.19, c=0.19, wmax=255, seed=4):
    s = _Arcs()
    rc = _genlib().fbcgen_rmat(scale, edge_factor, a, b, c, wmax, seed, ctypes.byref(s))
    if rc:
        raise ValueError(f"fbcgen_rmat failed: {rc}")
    return _take(s)


def write_edge_list(n, arcs, path):
    """'<src> <dst> <w>' lines (the reference's input format), written in C."""
    src, dst, w = (np.ascontiguousarray(x, np.uint32) for x in arcs)
    P = ctypes.POINTER(ctypes.c_uint32)
    a = _Arcs(n, len(src), src.ctypes.data_as(P), dst.ctypes.data_as(P), w.ctypes.data_as(P))
    rc = _genlib().fbcgen_write(ctypes.byref(a), path.encode())
    if rc:
        raise OSError(-rc, f"writing {path}")


def csr_from_arcs(n, src, dst, w):
    """Row pointers for arcs sorted by (src, dst)."""
    rp = np.zeros(n + 1, np.uint64)
    np.add.at(rp, np.asarray(src, np.int64) + 1, 1)
    return np.cumsum(rp).astype(np.uint64), _u32(dst), _u32(w)
