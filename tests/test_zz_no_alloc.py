"""fbc_dep_batch_noalloc (the entry point that allocates nothing) against
fbc_dep_batch, through the C ABI: two calls in a row, the second on a
smaller input (the shared static scratch keeps the first call's contents),
bit for bit in FP64 and in FP32 two-sum; a 1M-vertex graph (one source per
workgroup, the largest team); the size limit and bad arguments reported as
errors."""
import ctypes

import numpy as np
import pytest

import fastbc_amd as F

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


class DeviceBuffer:
    """Device memory owned by the caller (the entry point takes device
    pointers and allocates nothing itself)."""

    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0   # device to host
        return out

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


def noalloc(ctx, g, src, prec, n_rows=None):
    L = F.lib()
    src = np.ascontiguousarray(src, np.uint32)
    ds = DeviceBuffer(src.nbytes)
    ds.upload(src)
    rows = len(src) if n_rows is None else n_rows
    dd = DeviceBuffer(rows * g.n * 8)
    rc = L.fbc_dep_batch_noalloc(ctx.h, g.h, ds.p, len(src), prec, dd.p)
    ctx.check(rc)
    return dd.download(np.float64, rows * g.n).reshape(rows, g.n)


def test_noalloc_matches_dep_batch(gpu_ctx):
    n0, arcs0 = F.generate_config(0)
    g0 = F.Graph(gpu_ctx, n0, *F.csr_from_arcs(n0, *arcs0))
    n1, arcs1 = F.generate_grid(30, 30, 100, 9, road=True)
    g1 = F.Graph(gpu_ctx, n1, *F.csr_from_arcs(n1, *arcs1))
    src0 = np.random.default_rng(3).choice(n0, 96, replace=False)
    src1 = np.array([0, 17, n1 // 2, n1 - 1], np.uint32)
    for prec in (F.PREC_F64, F.PREC_F32C):
        print(f"no-alloc prec {prec}: {len(src0)} sources on {n0} vertices, then {len(src1)} on {n1}", flush=True)
        a = noalloc(gpu_ctx, g0, src0, prec)
        assert np.array_equal(a, g0.dep_batch(src0, prec))
        b = noalloc(gpu_ctx, g1, src1, prec)          # smaller input, the same static scratch
        assert np.array_equal(b, g1.dep_batch(src1, prec))
        _, te = gpu_ctx.last_stats()
        assert te > 0


def test_noalloc_workgroup_team_and_limits(gpu_ctx):
    n, arcs = F.generate_config(2)        # 1,048,576 vertices: one source per workgroup
    g = F.Graph(gpu_ctx, n, *F.csr_from_arcs(n, *arcs))
    src = np.array([5, n - 3], np.uint32)
    print(f"no-alloc: {len(src)} sources on {n} vertices", flush=True)
    assert np.array_equal(noalloc(gpu_ctx, g, src, F.PREC_F64), g.dep_batch(src))
    with pytest.raises(F.FbcError) as e:         # a source that is not a vertex
        noalloc(gpu_ctx, g, [1, n], F.PREC_F64)
    assert e.value.code == -1
    del g
    # a graph whose single source needs more than the static array holds
    big = 1_800_000
    s = np.arange(big, dtype=np.int64)
    gbig = F.Graph(gpu_ctx, big, *F.csr_from_arcs(big, s, (s + 1) % big, np.ones(big, np.int64)))
    with pytest.raises(F.FbcError) as e:
        noalloc(gpu_ctx, gbig, [0], F.PREC_F64)
    assert e.value.code == -7 and "static array" in str(e.value)
