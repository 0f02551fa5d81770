"""Benchmark: clustered-betweenness pivot phase on the 1,048,576-vertex road grid.

Workload: the 1024x1024 road-like grid (seed 3: 10% of links deleted under a
spanning-tree guard, 2% diagonal arcs, weights 1..1000; 3,846,612 arcs).
One step = one launch of the global pivot phase
(ClusteredBrandesBC.h:141-166) over a batch of pivot sources: a full-graph
SSSP with exact uint64 path counts and the dependency sums per source,
accumulated into the FP64 betweenness vector in HBM.  The headline form is the
batched persistent kernel with FP32 two-sum dependency sums; the FP64 form is
timed beside it.

Output: one JSON line (see the repository's measurement notes) with
  value     = traversed edges per second (GTEPS) of the headline form,
  roofline  = 48 algorithmic bytes per traversed edge over the kernel's
              HIP-event time, against 8 TB/s,
  cpu_baseline = the reference's DijkstraSSBrandesBC (oracle/_ref build) on
              the host cores over a bounded sample of the same sources.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, arcs, sources, seconds_target):
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
                for s in sources[:k]:
                    f.write(f"{int(s)} 1\n")
            out = os.path.join(tmp, "out.txt")
            # the reference's OpenMP array reduction over N doubles
            # (ClusteredBrandesBC.h:143) keeps a private N-array on each
            # thread's stack: 8 MB here, beyond default stack limits
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close",
                       OMP_STACKSIZE="256M")
            import resource

            def big_stack():
                resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

            r = subprocess.run([ref, "ssb", gpath, ppath, out], env=env, capture_output=True, text=True,
                               timeout=900, preexec_fn=big_stack)
            if r.returncode:
                raise RuntimeError(r.stderr)
            secs = float(open(out).readline().split()[2])
            if attempt == 0:
                k = int(min(len(sources), max(k, k * seconds_target / max(secs, 1e-3))))
                k = max(cores, (k // cores) * cores)
        kind = "reference"
        sample = (f"{k} pivot sources of the same batch through DijkstraSSBrandesBC::singleSourceBrandes "
                  f"(reference headers, g++ -O3, OpenMP reduction as ClusteredBrandesBC.h:141-155), "
                  f"{secs:.1f} s, graph load excluded")
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
        sample = f"{k} pivot sources through the oracle's single-source Brandes, one thread, {secs:.1f} s"
    return {"value": k * m / secs / 1e9, "unit": "GTEPS", "cores": cores, "kind": kind, "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=2048, help="pivot sources per step")
    ap.add_argument("--delta", type=int, default=0, help="near-far window (0 = automatic)")
    ap.add_argument("--slots", type=int, default=0, help="sources in flight (0 = automatic)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
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
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the pivot queue would shard "
              f"across GPUs as described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import fastbc_amd as F

    t0 = time.time()
    n, arcs = F.generate_config(2)
    m = len(arcs[0])
    ctx = F.Context(0)
    ctx.set_tuning(args.delta, args.slots)
    g = F.Graph(ctx, n, *F.csr_from_arcs(n, *arcs))
    # pivot batch: a seeded sample of vertices, every one its own weight-1 class
    # of a single all-vertex cluster (the per-source work does not depend on it)
    cl = g.clusters(comm_of=np.zeros(n, np.int64))
    rng = np.random.default_rng(12345)
    piv = rng.choice(n, args.batch, replace=False).astype(np.uint32)
    wts = np.ones(args.batch, np.uint32)
    pcl = np.zeros(args.batch, np.uint32)
    intra = np.zeros(n)
    log(f"setup {time.time() - t0:.1f}s: n={n} m={m} batch={args.batch}")

    def run_form(prec, steps, warmup):
        bc = np.zeros(n)
        for _ in range(warmup):
            cl.pivot_phase(piv, wts, pcl, intra, bc, prec=prec)
        kms, te = 0.0, 0.0
        t = time.perf_counter()
        for _ in range(steps):
            bc = cl.pivot_phase(piv, wts, pcl, intra, bc, prec=prec)
            ms, e = ctx.last_stats()
            kms += ms
            te += e
        wall = time.perf_counter() - t
        return wall, kms, te, ctx.last_counters()

    wall, kms, te, counters = run_form(F.PREC_F32C, args.steps, args.warmup)
    log(f"f32c: {args.steps} steps wall {wall:.3f}s kernel {kms:.1f}ms TE {te:.3e} counters {counters}")
    wall64, kms64, te64, _ = run_form(F.PREC_F64, max(1, args.steps // 2), 0)
    log(f"f64:  wall {wall64:.3f}s kernel {kms64:.1f}ms")

    gteps = te / wall / 1e9
    achieved = te / args.steps * BYTES_PER_TE / (kms / args.steps * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32x2 (two-sum) dependencies, uint32 paths, uint64 sigma, fp64 reduction",
        "data": "synthetic",
        "config": {
            "workload": "1024x1024 road-like grid (seed 3), global pivot phase, one batch of pivot sources per step",
            "vertices": n, "arcs": m, "batch_sources": args.batch,
            "delta": args.delta or "auto", "slots": args.slots or "auto",
            "te_per_step": te / args.steps,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
        },
        "fp64_form": {"value": round(te64 / wall64 / 1e9, 4), "unit": "GTEPS",
                      "kernel_ms_per_step": round(kms64 / max(1, args.steps // 2), 3)},
        "kernel_counters": counters,
        "pivot_phase_s_for_524288_pivots": round(524288 * m / (gteps * 1e9), 1),
    }
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(n, arcs, piv, args.cpu_seconds)
        except Exception as e:  # the GPU figure stands without it
            log(f"cpu baseline failed: {e}")
            result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
