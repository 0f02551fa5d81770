"""Benchmark: clustered betweenness centrality on the 1,048,576-vertex road grid.

Workload: the 1024x1024 road-like grid (seed 3: 10% of links deleted under a
spanning-tree guard, 2% diagonal arcs, weights 1..1000; 3,846,612 arcs), run
as main.cpp composes the clustered computer with -s 1,2,3,4 -e 4 -p 0.01
-k 0.5 (ClusteredBrandeBC::computeBC, ClusteredBrandesBC.h:63-169).

One step = one computeBC: Louvain on the host, every (cluster, source)
Brandes pass in one persistent launch, the VertexInfo pivot classes and the
k-means super-classes (GPU), then every pivot in one persistent launch -- the
batched form, with FP32 two-sum dependency sums (the headline).  Timed like
the reference's timer (main.cpp:269-277): computeBC only, graph building
excluded; the library call is synchronous.

Beside the headline, in the same JSON line:
  fp64_form        the batched form with FP64 dependency sums (one run) and
                   the largest relative BC difference between the two forms;
  plain_fp64_form  the pivot phase as the reference runs it, one
                   singleSourceBrandes per pivot -- here one cooperative
                   launch of 16 workgroups per pivot, FP64 -- timed over a sample of
                   the run's pivots and scaled to all of them; the other
                   phases are the batched form's (a conservative baseline);
  roofline         the pivot-phase launch (the dominant kernel): 48
                   algorithmic bytes per traversed edge over its HIP-event
                   time, against 8 TB/s; traffic from the rocprofv3 PMC
                   summary in profiles/ when present;
  cpu_baseline     the reference's DijkstraSSBrandesBC (oracle/_ref build) on
                   the host cores over a bounded sample of the run's pivots.
value = traversed edges of the step (cluster phase + pivot phase) per second
of computeBC wall time.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-bc_amd"))

METRIC = "GTEPS + BC wall-time, 1M-vertex grid: batched FP32-comp vs plain FP64, 1 GPU"
BYTES_PER_TE = 48.0
HBM_PEAK_GBS = 8000.0
SEEDS = [1, 2, 3, 4]
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_brandes_pivot_cfg2_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def big_stack():
    """The reference keeps N-double reduction arrays on thread stacks
    (ClusteredBrandesBC.h:143, DijkstraClusterEvaluator.h:62)."""
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def cpu_phases(ref, gpath, n, arcs, comm, cores, tmp):
    """The reference's other computeBC phases on the host: Louvain in full
    (the reference's Partition code, 4 instances), and DijkstraClusterEvaluator
    on the smallest bordered community, scaled to every bordered community by
    traversed edges (n_c * m_c).  The evaluator's cost grows faster than that
    (1,159 / 1,778 / 7,751 vertices: 7.9 / 20.9 / 602 s on one core), so the
    scaled figure is a lower bound.  Both run concurrently (5 threads)."""
    import numpy as np
    env = dict(os.environ, OMP_STACKSIZE="256M")
    part = os.path.join(tmp, "part.txt")
    np.savetxt(part, comm, fmt="%d")
    s, d = arcs[0].astype(np.int64), arcs[1].astype(np.int64)
    k = int(comm.max()) + 1
    inside = comm[s] == comm[d]
    m_c = np.bincount(comm[s[inside]], minlength=k).astype(np.float64)
    n_c = np.bincount(comm, minlength=k).astype(np.float64)
    isb = np.zeros(n, bool)
    isb[s[~inside]] = True
    isb[d[~inside]] = True
    bordered = np.bincount(comm[isb], minlength=k) > 0
    small = int(np.argmin(np.where(bordered, n_c, np.inf)))
    lv = subprocess.Popen([ref, "louvain", gpath, "0.01"] + [str(x) for x in SEEDS],
                          env=dict(env, OMP_NUM_THREADS="4"), stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                          text=True, preexec_fn=big_stack)
    cl = subprocess.run([ref, "clusters", gpath, part, str(small)], env=dict(env, OMP_NUM_THREADS="1"),
                        capture_output=True, text=True, timeout=900, preexec_fn=big_stack)
    lv_err = lv.communicate(timeout=900)[1]
    if cl.returncode or lv.returncode:
        raise RuntimeError(cl.stderr + lv_err)
    louvain_s = float(lv_err.strip().splitlines()[-1].split("=")[1])
    t_small = float(cl.stdout.split("seconds")[1].split()[0])
    te_small = n_c[small] * m_c[small]
    te_all = float((n_c * m_c)[bordered].sum())
    return {"louvain_s": round(louvain_s, 2), "cluster_phase_s_lower_bound": round(t_small * te_all / te_small / cores, 1),
            "cluster_sample": f"community {small} ({int(n_c[small])} vertices, {int(m_c[small])} internal arcs): "
                              f"{t_small:.1f} s on one core; all {int(bordered.sum())} bordered communities scaled by "
                              f"n_c*m_c over {cores} cores"}


def cpu_model():
    """The host CPU's model name and logical CPU count, from /proc/cpuinfo."""
    try:
        names = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else None
    except OSError:
        return None


def cpu_baseline(n, arcs, sources, seconds_target, comm):
    """Time the reference's single-source Brandes on the host (bounded sample)."""
    import numpy as np
    import fastbc_amd as F
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver_fast")
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))   # the host-core share of one GPU
    tmp = tempfile.mkdtemp(prefix="fbc_bench_")
    gpath = os.path.join(tmp, "graph.txt")
    F.write_edge_list(n, arcs, gpath)
    m = len(arcs[0])
    if os.path.exists(ref):
        # one pivot per core first, to size the sample
        k = cores
        ppath = os.path.join(tmp, "piv.txt")
        for attempt in range(2):
            with open(ppath, "w") as f:
                for s in sources[::max(1, len(sources) // k)][:k]:   # spread over the clusters
                    f.write(f"{int(s)} 1\n")
            out = os.path.join(tmp, "out.txt")
            # the reference's OpenMP array reduction over N doubles
            # (ClusteredBrandesBC.h:143) keeps a private N-array on each
            # thread's stack: 8 MB here, beyond default stack limits
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close",
                       OMP_STACKSIZE="256M")
            r = subprocess.run([ref, "ssb", gpath, ppath, out], env=env, capture_output=True, text=True,
                               timeout=900, preexec_fn=big_stack)
            if r.returncode:
                raise RuntimeError(r.stderr)
            secs = float(open(out).readline().split()[2])
            if attempt == 0:
                k = int(min(len(sources), max(k, k * seconds_target / max(secs, 1e-3))))
                k = max(cores, (k // cores) * cores)
        kind = "reference"
        sample = (f"{k} of the run's pivots through DijkstraSSBrandesBC::singleSourceBrandes "
                  f"(reference headers, g++ -O3, OpenMP reduction as ClusteredBrandesBC.h:141-155), "
                  f"{secs:.1f} s, graph load excluded")
        pivot_s = secs * len(sources) / k
        est = {"pivot_phase_s": round(pivot_s, 1)}
        try:
            est.update(cpu_phases(ref, gpath, n, arcs, comm, cores, tmp))
            est["computeBC_s_lower_bound"] = round(est["louvain_s"] + est["cluster_phase_s_lower_bound"] + pivot_s, 1)
        except Exception as e:  # the pivot-phase figure stands without it
            log(f"cpu phases failed: {e}")
        est["note"] = ("extrapolated: Louvain measured in full; pivot phase from the pivot sample, scaled by the "
                       "run's pivot count (every pivot traverses the whole graph); cluster phase a lower bound; "
                       "pivot selection not included")
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        og = O.Graph.from_arcs(arcs[0], arcs[1], arcs[2].astype(np.float64))
        k = cores
        t = time.time()
        for s in sources[:k]:
            og.single_source_dep(int(s))
        secs = time.time() - t
        cores = 1
        kind = "port"
        sample = f"{k} of the run's pivots through the oracle's single-source Brandes, one thread, {secs:.1f} s"
    out = {"value": k * m / secs / 1e9, "unit": "GTEPS", "cores": cores, "kind": kind, "sample": sample,
           "cpu_model": cpu_model()}
    if kind == "reference":
        out["end_to_end"] = est
    return out


def ref_exact(n, arcs, cores, sample_sources=None):
    """The reference's ExactBrandesBC on the host: every source when
    sample_sources is None (returns its seconds and BC), else
    DijkstraSSBrandesBC over that many sources spread over the vertices, the
    seconds scaled to all n sources (returns the estimate and no BC)."""
    import numpy as np
    import fastbc_amd as F
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver_fast")
    if not os.path.exists(ref):
        return None, None
    tmp = tempfile.mkdtemp(prefix="fbc_table_")
    gpath = os.path.join(tmp, "graph.txt")
    out = os.path.join(tmp, "out.txt")
    F.write_edge_list(n, arcs, gpath)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_STACKSIZE="256M")
    if sample_sources is None:
        r = subprocess.run([ref, "exact", gpath, out], env=env, capture_output=True, text=True, timeout=900,
                           preexec_fn=big_stack)
        if r.returncode:
            raise RuntimeError(r.stderr)
        secs = float(r.stderr.split("seconds=")[1].split()[0])
        return secs, np.loadtxt(out)
    ppath = os.path.join(tmp, "piv.txt")
    with open(ppath, "w") as f:
        for v in range(0, n, max(1, n // sample_sources))[:sample_sources]:
            f.write(f"{v} 1\n")
    r = subprocess.run([ref, "ssb", gpath, ppath, out], env=env, capture_output=True, text=True, timeout=900,
                       preexec_fn=big_stack)
    if r.returncode:
        raise RuntimeError(r.stderr)
    return float(open(out).readline().split()[2]) * n / sample_sources, None


def table(args):
    """One JSON line per (configuration, form): computeBC wall time (steady
    state: each form's first call on a graph is a warm-up) and GTEPS of the
    batched FP32C and FP64 forms, FP32C's largest relative difference from
    FP64, the plain FP64 form (every pivot when there are at most 3,000, else
    a strided sample of 32 scaled to all), and the reference's own code on
    the host cores (every source of cfg0's exact BC; a sample for the rest)."""
    import numpy as np
    import fastbc_amd as F
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))

    def emit(**kw):
        print(json.dumps(kw), flush=True)

    def rel(a, b):
        nz = b != 0
        return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]))) if nz.any() else 0.0

    for cfg, mode, kfrac in ((0, "exact", 0), (0, "clustered", 0.5), (1, "clustered", 0.5), (3, "exact", 0),
                             (2, "clustered", 0.5), (4, "clustered", 0.25)):
        n, arcs = F.generate_config(cfg)
        m = len(arcs[0])
        log(f"cfg{cfg} {mode}: n={n} m={m}")
        if mode == "exact":
            res = {}
            for name, prec in (("fp64", F.PREC_F64), ("fp32c", F.PREC_F32C)):
                F.exact_bc_arcs(*arcs, prec=prec)
                t = time.perf_counter()
                res[name] = F.exact_bc_arcs(*arcs, prec=prec)
                secs = time.perf_counter() - t
                emit(config=f"cfg{cfg}", mode=mode, form=f"GPU batched {name}", wall_s=round(secs, 4),
                     gteps=round(n * m / secs / 1e9, 3), traversed_edges=n * m,
                     fp32c_max_rel_diff_vs_fp64=rel(res["fp32c"], res["fp64"]) if name == "fp32c" else None)
            F.host_release()
            secs, ref_bc = ref_exact(n, arcs, cores, None if cfg == 0 else 16 * cores)
            if secs is not None:
                emit(config=f"cfg{cfg}", mode=mode, form=f"reference CPU ({cores} cores)", wall_s=round(secs, 2),
                     gteps=round(n * m / secs / 1e9, 4), cores=cores, cpu_model=cpu_model(),
                     sample="every source" if ref_bc is not None else f"{16 * cores} sources, scaled to {n}",
                     gpu_fp64_max_rel_err_vs_reference=rel(res["fp64"], ref_bc) if ref_bc is not None else None)
            continue
        runs = {}
        for name, prec in (("fp32c", F.PREC_F32C), ("fp64", F.PREC_F64)):
            if name == "fp32c":
                t = time.perf_counter()
                F.clustered_bc(*arcs, SEEDS, 0.01, kfrac=kfrac, prec=prec)
                log(f"  cold first call {time.perf_counter() - t:.1f} s")
            r = F.clustered_bc(*arcs, SEEDS, 0.01, kfrac=kfrac, prec=prec)
            runs[name] = r
            ph = r["phases"]
            te = ph["cluster_te"] + ph["pivot_te"]
            emit(config=f"cfg{cfg}", mode=f"{mode} -k {kfrac}", form=f"GPU batched {name}",
                 wall_s=round(ph["total_s"], 3), gteps=round(te / ph["total_s"] / 1e9, 3),
                 pivot_kernel_gteps=round(ph["pivot_te"] / (ph["pivot_kernel_ms"] * 1e-3) / 1e9, 3),
                 roofline_frac=round(ph["pivot_te"] * BYTES_PER_TE / (ph["pivot_kernel_ms"] * 1e-3) / 1e9
                                     / HBM_PEAK_GBS, 4),
                 phases_s={k: round(v, 3) for k, v in ph.items() if k.endswith("_s")},
                 clusters=int(np.max(r["comm"])) + 1, pivots=len(r["pivots"]), traversed_edges=te,
                 fp32c_max_rel_diff_vs_fp64=rel(runs["fp32c"]["bc"], r["bc"]) if name == "fp64" else None)
        r64 = runs["fp64"]
        p64 = r64["phases"]
        F.host_release()
        ctx = F.Context(0)
        g = F.Graph(ctx, n, *F.csr_from_arcs(n, *arcs))
        cl = g.clusters(comm_of=r64["comm"])
        npiv = len(r64["pivots"])
        k = npiv if npiv <= 3000 else 32
        stride = max(1, npiv // k)
        sp, sw, sc = (a[::stride][:k] for a in (r64["pivots"], r64["weights"], r64["pivot_cluster"]))
        zero = np.zeros(n)
        plain = cl.pivot_phase(sp, sw, sc, r64["intra"], zero, plain=True)
        plain_ms, _ = ctx.last_stats()
        pc = ctx.last_counters()
        batched = cl.pivot_phase(sp, sw, sc, r64["intra"], zero, prec=F.PREC_F64)
        agree = float(np.max(np.abs(plain - batched)) / max(np.abs(batched).max(), 1e-300))
        plain_pivot_s = plain_ms * 1e-3 / k * npiv
        plain_total = p64["total_s"] - p64["pivot_phase_s"] + plain_pivot_s
        emit(config=f"cfg{cfg}", mode=f"{mode} -k {kfrac}", form="GPU plain fp64 (one cooperative launch per pivot)",
             wall_s=round(plain_total, 2), gteps=round((p64["cluster_te"] + p64["pivot_te"]) / plain_total / 1e9, 4),
             ms_per_pivot=round(plain_ms / k, 3),
             sample="every pivot" if k == npiv else f"{k} of {npiv} pivots (every {stride}th), scaled",
             cooperative_launches=pc["plain_sources"], batched_fallback_sources=pc["plain_fallback_sources"],
             max_diff_vs_batched_fp64=agree)
        del cl, g, ctx
        if cfg != 2:   # the default bench line carries cfg2's
            try:
                cb = cpu_baseline(n, arcs, r64["pivots"], 10.0, r64["comm"])
                emit(config=f"cfg{cfg}", mode=f"{mode} -k {kfrac}", form=f"reference CPU ({cb['cores']} cores)",
                     **{k2: v for k2, v in cb.items() if k2 != "unit"})
            except Exception as e:
                log(f"cpu baseline failed: {e}")


def traffic_per_launch(npiv):
    """HBM bytes of the pivot-phase launch from the committed PMC summary,
    scaled from its pivot count to this run's."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            t = json.load(f)
        return t["bytes_per_pivot"] * npiv, t
    except (OSError, ValueError, KeyError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kfrac", type=float, default=0.5)
    ap.add_argument("--plain-sample", type=int, default=32, help="pivots timed in the plain form")
    ap.add_argument("--plain-blocks", type=int, default=16, help="workgroups of the plain form's grid")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 and plain forms")
    ap.add_argument("--table", action="store_true",
                    help="measure every benchmark configuration (one JSON line per configuration and form)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): it measures a single "
              f"MI355X; the pivot and (cluster, source) queues would be split across GPUs", file=sys.stderr)
        sys.exit(2)

    if args.table:
        table(args)
        return

    import numpy as np
    import fastbc_amd as F

    t0 = time.time()
    n, arcs = F.generate_config(2)
    m = len(arcs[0])
    log(f"graph: n={n} m={m} ({time.time() - t0:.1f}s)")

    def step(prec, form=F.FORM_BATCHED):
        t = time.perf_counter()
        r = F.clustered_bc(*arcs, SEEDS, 0.01, kfrac=args.kfrac, prec=prec, form=form)
        r["call_s"] = time.perf_counter() - t
        return r

    cold = cold_phases = None
    for i in range(args.warmup):
        r = step(F.PREC_F32C)
        if cold is None:
            cold = r["phases"]["total_s"]
            cold_phases = {k: round(v, 3) for k, v in r["phases"].items() if k.endswith("_s")}
        log(f"warmup {i}: computeBC {r['phases']['total_s']:.2f}s phases {r['phases']} tuning {r['tuning']}")
    runs = []
    for i in range(args.steps):
        r = step(F.PREC_F32C)
        runs.append(r)
        log(f"step {i}: computeBC {r['phases']['total_s']:.2f}s phases {r['phases']}")
    last = runs[-1]
    ph = last["phases"]
    total_s = sum(r["phases"]["total_s"] for r in runs) / len(runs)
    te = ph["cluster_te"] + ph["pivot_te"]
    npiv = len(last["pivots"])
    gteps = te / total_s / 1e9
    piv_ms = ph["pivot_kernel_ms"]
    achieved = ph["pivot_te"] * BYTES_PER_TE / (piv_ms * 1e-3) / 1e9
    traffic, tsum = traffic_per_launch(npiv)

    result = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32x2 (two-sum) dependencies, uint32 paths, uint64 sigma, fp64 reduction",
        "data": "synthetic",
        "config": {
            "workload": "1024x1024 road-like grid (seed 3), clustered BC -s 1,2,3,4 -e 4 -p 0.01 -k 0.5, "
                        "batched form, one computeBC per step",
            "vertices": n, "arcs": m, "clusters": int(np.max(last["comm"])) + 1, "pivots": npiv,
            "traversed_edges_per_step": te,
        },
        "bc_wall_s": round(total_s, 3),
        # the process's first computeBC on the graph: device upload, the
        # relabelled copy, slot memory and the autotuner's trial launches
        "cold_first_call_s": round(cold, 3) if cold is not None else None,
        "cold_first_call_phases_s": cold_phases,
        "tuning": last["tuning"],
        "phases_s": {k: round(v, 3) for k, v in ph.items() if k.endswith("_s")},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": round(traffic) if traffic else None,
            "kernel": "k_brandes<MODE_PIVOT, F32C> (pivot phase, one launch)",
            "kernel_ms": round(piv_ms, 2),
        },
    }
    if tsum:
        result["roofline"]["traffic_source"] = os.path.relpath(TRAFFIC_SUMMARY, ROOT)

    if not args.no_compare:
        r64 = step(F.PREC_F64)
        p64 = r64["phases"]
        bc32, bc64 = last["bc"], r64["bc"]
        nz = bc64 != 0
        rel = float(np.max(np.abs(bc32[nz] - bc64[nz]) / np.abs(bc64[nz]))) if nz.any() else 0.0
        log(f"fp64 batched: computeBC {p64['total_s']:.2f}s; max rel diff fp32c vs fp64 {rel:.3e}")
        result["fp64_form"] = {
            "bc_wall_s": round(p64["total_s"], 3),
            "value": round((p64["cluster_te"] + p64["pivot_te"]) / p64["total_s"] / 1e9, 4), "unit": "GTEPS",
            "pivot_kernel_ms": round(p64["pivot_kernel_ms"], 2),
            "fp32c_max_rel_diff": rel,
            "same_pivots": bool(np.array_equal(r64["pivots"], last["pivots"])),
        }
        # the plain form over a sample of the same pivots
        ctx = F.Context(0)
        g = F.Graph(ctx, n, *F.csr_from_arcs(n, *arcs))
        cl = g.clusters(comm_of=r64["comm"])
        k = min(args.plain_sample, npiv)
        stride = max(1, npiv // k)   # pivots spread over the clusters
        sp, sw, sc = (a[::stride][:k] for a in (r64["pivots"], r64["weights"], r64["pivot_cluster"]))
        zero = np.zeros(n)
        ctx.set_tuning(0, args.plain_blocks)
        t = time.perf_counter()
        plain = cl.pivot_phase(sp, sw, sc, r64["intra"], zero, plain=True)
        plain_wall = time.perf_counter() - t
        plain_ms, plain_te = ctx.last_stats()
        pc = ctx.last_counters()
        ctx.set_tuning(0, 0)
        batched = cl.pivot_phase(sp, sw, sc, r64["intra"], zero, prec=F.PREC_F64)
        scale = max(np.abs(batched).max(), 1e-300)
        agree = float(np.max(np.abs(plain - batched)) / scale)
        plain_pivot_s = plain_ms * 1e-3 / k * npiv
        plain_total = p64["total_s"] - p64["pivot_phase_s"] + plain_pivot_s
        log(f"plain: {k} pivots {plain_ms:.1f} ms ({plain_ms / k:.2f} ms/pivot); pivot phase for {npiv} "
            f"pivots {plain_pivot_s:.1f}s; computeBC {plain_total:.1f}s; agreement {agree:.2e}")
        result["plain_fp64_form"] = {
            "bc_wall_s": round(plain_total, 2),
            "value": round(te / plain_total / 1e9, 4), "unit": "GTEPS",
            "pivot_phase_s": round(plain_pivot_s, 2),
            "ms_per_pivot": round(plain_ms / k, 3),
            "pivot_gteps": round(plain_te / (plain_ms * 1e-3) / 1e9, 4),
            "sample": f"{k} of the run's {npiv} pivots (every {stride}th), one cooperative launch of "
                      f"{args.plain_blocks} workgroups each; pivot phase scaled by {npiv}/{k}",
            "max_diff_vs_batched_fp64": agree,
            "cooperative_launches": pc["plain_sources"],
            "batched_fallback_sources": pc["plain_fallback_sources"],
        }
        if pc["plain_fallback_sources"]:
            log(f"plain form: {pc['plain_fallback_sources']} pivots ran on the batched kernel "
                f"(cooperative launch refused, HIP status {pc['plain_refusal']}); no plain figure")
            result["plain_fp64_form"]["value"] = None
        result["speedup_batched_fp32c_vs_plain_fp64"] = round(plain_total / total_s, 2)
        result["pivot_phase_speedup"] = round(plain_pivot_s / ph["pivot_phase_s"], 1)
        del g, cl, ctx
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(n, arcs, last["pivots"], args.cpu_seconds, last["comm"])
        except Exception as e:  # the GPU figure stands without it
            log(f"cpu baseline failed: {e}")
            result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
