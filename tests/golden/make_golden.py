"""Writes the golden fixtures of tests/golden/ by running the reference's own
code (oracle/_ref/ref_driver, built from the headers under /root/reference by
`make -C oracle ref`).  Runs only where the reference sources are present;
the fixtures are committed, so the tests never need the reference.

Fixtures (numpy .npz, no pickles):
  cfg0_exact.npz        ExactBrandesBC::computeBC on the cfg0 grid (1 thread)
  louvain.npz           Partition-based Louvain (seeds 1,2,3,4, precision 0.01)
                        on cfg0 and cfg1
  cfg0_clusters.npz     DijkstraClusterEvaluator on three cfg0 communities:
                        borders, intra BC, VertexInfo rows
  kmeans_cases.npz      PlusPlusKMeans on seeded random VertexInfo rows
  cfg0_ssb.npz          DijkstraSSBrandesBC summed over 16 weighted pivots
  graph_digests.json    sha256 of the generated cfg0..cfg3 edge lists

Large-configuration fixtures (`--part cfg3`, `--part cfg2`, `--part cfg4`;
the default `--part base` writes the ones above):
  cfg3_exact.npz        ExactBrandesBC::computeBC on the R-MAT scale-15 graph
                        (ref_driver_fast, OpenMP; thread count recorded)
  cfg2_reference.npz    on the 1024x1024 road grid: the reference Louvain
                        partition (seeds 1,2,3,4, precision 0.01); for its
                        three smallest communities, DijkstraClusterEvaluator's
                        borders and intra-cluster BC, and sha256 digests of
                        the VertexInfo rows (uint32 lengths, int32 counts,
                        row-major); PlusPlusKMeans run by the reference on
                        those communities' exact classes (k = P/2) with the
                        resulting pivots and weights; DijkstraSSBrandesBC
                        dependency rows of four of the run's pivots as sha256
                        digests of their FP64 bytes (plus their sums)
  cfg2_oracle_pivots.npz  the oracle's pivots, weights and pivot clusters of
                        the whole cfg2 selection (all 211 communities), and a
                        digest of its intra-cluster BC vector
  cfg4_reference.npz    on the 2048x2048 road grid: digest and community
                        count of the reference Louvain partition; dependency
                        digests of two sources
  degenerate.npz        (`--part degenerate`) five small graphs of unusual
                        shape: partitions, exact BC, intra-cluster BC and
                        weighted dependency sums from the reference's code
"""
import hashlib
import json
import os
import resource
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "fast-bc_amd"))
import fastbc_amd as F  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def big_stack():
    """The reference keeps N-double OpenMP reduction arrays on thread stacks
    (ClusteredBrandesBC.h:143, DijkstraClusterEvaluator.h:62): 8 MB at 1M
    vertices, beyond the default stack limit."""
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run(args, env=None):
    e = dict(os.environ, OMP_NUM_THREADS="1", OMP_STACKSIZE="256M")
    if env:
        e.update(env)
    r = subprocess.run(args, env=e, capture_output=True, text=True, check=True, preexec_fn=big_stack)
    return r.stdout


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/ref_driver is not built (make -C oracle ref, needs /root/reference)")
    tmp = tempfile.mkdtemp(prefix="fbc_golden_")
    digests = {}
    paths = {}
    for cfg in (0, 1, 2, 3):
        n, arcs = F.generate_config(cfg)
        p = os.path.join(tmp, f"cfg{cfg}.txt")
        F.write_edge_list(n, arcs, p)
        digests[f"cfg{cfg}"] = {"n": int(n), "m": int(len(arcs[0])),
                                "sha256": hashlib.sha256(open(p, "rb").read()).hexdigest()}
        paths[cfg] = p
    with open(os.path.join(HERE, "graph_digests.json"), "w") as f:
        json.dump(digests, f, indent=1)

    # ExactBrandesBC on cfg0, one thread (source order = the sum order)
    out = os.path.join(tmp, "exact.txt")
    run([REF, "exact", paths[0], out])
    np.savez_compressed(os.path.join(HERE, "cfg0_exact.npz"), bc=np.loadtxt(out))

    # Louvain partitions
    parts = {}
    for cfg in (0, 1):
        txt = run([REF, "louvain", paths[cfg], "0.01", "1", "2", "3", "4"], env={"OMP_NUM_THREADS": "4"})
        parts[f"cfg{cfg}"] = np.array(txt.split(), dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "louvain.npz"), **parts)

    # cluster evaluation of three cfg0 communities
    comm = parts["cfg0"]
    res = {}
    for i, c in enumerate((0, 5, 17)):
        verts = np.where(comm == c)[0]
        vp = os.path.join(tmp, f"c{c}.txt")
        np.savetxt(vp, verts, fmt="%d")
        op = os.path.join(tmp, f"c{c}_out.txt")
        run([REF, "cluster", paths[0], vp, op])
        lines = open(op).read().split("\n")
        borders = np.array(lines[0].split()[1:], dtype=np.int32)
        rows = [ln.split() for ln in lines[1:] if ln]
        res[f"verts{i}"] = verts.astype(np.int32)
        res[f"borders{i}"] = borders
        res[f"bc{i}"] = np.array([float(r[1]) for r in rows])
        res[f"len{i}"] = np.array([[float(x) for x in r[2::2]] for r in rows]).reshape(len(rows), -1)
        res[f"cnt{i}"] = np.array([[int(x) for x in r[3::2]] for r in rows], dtype=np.int32).reshape(len(rows), -1)
    np.savez_compressed(os.path.join(HERE, "cfg0_clusters.npz"), **res)

    # k-means cases
    rng = np.random.default_rng(2024)
    cases = {}
    for t in range(24):
        P = int(rng.integers(1, 80))
        nb = int(rng.integers(1, 7))
        k = int(rng.integers(1, P + 1))
        L = rng.integers(0, 7, (P, nb)).astype(float) * (1.0 if t % 3 else 41.0)
        C = rng.integers(0, 3, (P, nb)).astype(np.int32)
        W = rng.integers(1, 6, P).astype(np.int32)
        kp = os.path.join(tmp, "km.txt")
        with open(kp, "w") as f:
            f.write(f"{P} {nb} {k}\n")
            for p in range(P):
                f.write(str(W[p]) + " " + " ".join(f"{L[p, j]:.17g} {C[p, j]}" for j in range(nb)) + "\n")
        ko = os.path.join(tmp, "km_out.txt")
        run([REF, "kmeans", kp, ko])
        r = np.loadtxt(ko, dtype=np.int64).reshape(-1, 2)
        cases[f"L{t}"], cases[f"C{t}"], cases[f"W{t}"] = L, C, W
        cases[f"k{t}"] = np.array([k])
        cases[f"cent{t}"], cases[f"cw{t}"] = r[:, 0].astype(np.int32), r[:, 1].astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "kmeans_cases.npz"), **cases)

    # weighted single-source Brandes sum (the pivot loop body)
    piv = rng.choice(4096, 16, replace=False)
    wts = rng.integers(1, 5, 16)
    pp = os.path.join(tmp, "piv.txt")
    np.savetxt(pp, np.stack([piv, wts], 1), fmt="%d")
    so = os.path.join(tmp, "ssb.txt")
    run([REF + "_fast", "ssb", paths[0], pp, so])
    np.savez_compressed(os.path.join(HERE, "cfg0_ssb.npz"), pivots=piv.astype(np.int32),
                        weights=wts.astype(np.int32), bc=np.loadtxt(so))
    print("golden fixtures written to", HERE)


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8).copy()


def ref_dep_rows(ref, graph, sources, tmp):
    """DijkstraSSBrandesBC::singleSourceBrandes rows (ssb mode, one pivot of
    weight 1 per run: the summed row is the dependency row itself)."""
    procs, outs = [], []
    for s in sources:
        pp = os.path.join(tmp, f"src{s}.txt")
        with open(pp, "w") as f:
            f.write(f"{int(s)} 1\n")
        op = os.path.join(tmp, f"dep{s}.txt")
        outs.append(op)
        procs.append(subprocess.Popen([ref, "ssb", graph, pp, op],
                                      env=dict(os.environ, OMP_NUM_THREADS="1", OMP_STACKSIZE="256M"),
                                      preexec_fn=big_stack))
    for p in procs:
        assert p.wait() == 0
    return [np.loadtxt(op, comments="#") for op in outs]


def degenerate_cases():
    """Small graphs with the shapes the benchmark grids lack: a 2-cycle, ids
    with no arcs, two components, a star, and a random graph with self-loops,
    repeated arcs (the loader sums them) and vertices no arc reaches."""
    yield "two_cycle", [0, 1], [1, 0], [3, 5]
    yield "isolated_ids", [0, 5, 9], [5, 9, 0], [1, 1, 1]
    yield "two_components", [0, 1, 2, 3, 0, 4, 5, 6, 7, 4], [1, 2, 3, 0, 2, 5, 6, 7, 4, 6], [2, 2, 2, 2, 3, 1, 4, 1, 4, 5]
    hub = 20
    yield "star", [0] * hub + list(range(1, hub + 1)), list(range(1, hub + 1)) + [0] * hub, [1] * (2 * hub)
    rng = np.random.default_rng(7)
    n = 60
    s = [int(x) for x in rng.integers(0, n, 220)]
    d = [int(x) for x in rng.integers(0, n, 220)]
    w = [int(x) for x in rng.integers(1, 8, 220)]
    s.append(n - 1), d.append(0), w.append(3)
    for v in rng.choice(n, 8, replace=False):
        s.append(int(v)), d.append(int(v)), w.append(int(rng.integers(1, 8)))
    for i in rng.choice(len(s), 20, replace=False):
        s.append(s[i]), d.append(d[i]), w.append(int(rng.integers(1, 8)))
    yield "loops_duplicates", s, d, w


def part_degenerate(tmp):
    """degenerate.npz: per case, the arcs; the reference Louvain partition
    (seeds 1,2,3,4, precision 0.01); ExactBrandesBC (1 thread); the intra-
    cluster BC of DijkstraClusterEvaluator over every bordered community (the
    reference build skips borderless ones, main.cpp:6); and, for kfrac 0 and
    0.5, the oracle's pivots and weights (the pivot selectors include spdlog
    and do not build here) with the reference's DijkstraSSBrandesBC sum over
    them (ssb mode, 1 thread)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    res = {}
    for name, s, d, w in degenerate_cases():
        s, d, w = (np.asarray(a, np.int32) for a in (s, d, w))
        gp = os.path.join(tmp, f"{name}.txt")
        with open(gp, "w") as f:
            f.writelines(f"{a} {b} {c}\n" for a, b, c in zip(s, d, w))
        res[f"{name}_src"], res[f"{name}_dst"], res[f"{name}_w"] = s, d, w
        comm = np.array(run([REF, "louvain", gp, "0.01", "1", "2", "3", "4"], env={"OMP_NUM_THREADS": "4"}).split(),
                        dtype=np.int32)
        res[f"{name}_comm"] = comm
        ep = os.path.join(tmp, f"{name}_exact.txt")
        run([REF, "exact", gp, ep])
        res[f"{name}_exact"] = np.loadtxt(ep, ndmin=1)
        intra = np.zeros(len(comm))
        for c in np.unique(comm):
            verts = np.where(comm == c)[0]
            vp = os.path.join(tmp, f"{name}_c{c}.txt")
            np.savetxt(vp, verts, fmt="%d")
            op = os.path.join(tmp, f"{name}_c{c}_out.txt")
            run([REF, "cluster", gp, vp, op])
            lines = open(op).read().split("\n")
            if len(lines[0].split()) > 1:   # bordered
                rows = [ln.split() for ln in lines[1:] if ln]
                intra[verts] = [float(r[1]) for r in rows]
        res[f"{name}_intra"] = intra
        og = O.Graph.from_file(gp)
        for tag, kf in (("k0", 0.0), ("k5", 0.5)):
            r = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=kf)
            assert np.array_equal(r["comm"], comm)
            piv, wts = r["pivots"].astype(np.int32), r["weights"].astype(np.int32)
            res[f"{name}_{tag}_pivots"], res[f"{name}_{tag}_weights"] = piv, wts
            res[f"{name}_{tag}_pcl"] = r["pivot_cluster"].astype(np.int32)
            ssb = np.zeros(len(comm))
            if len(piv):
                pp = os.path.join(tmp, f"{name}_{tag}_piv.txt")
                np.savetxt(pp, np.stack([piv, wts], 1), fmt="%d")
                so = os.path.join(tmp, f"{name}_{tag}_ssb.txt")
                run([REF, "ssb", gp, pp, so])
                ssb = np.loadtxt(so, comments="#", ndmin=1)
            res[f"{name}_{tag}_ssb"] = ssb
    # VertexInfo counts beyond the int range: a 20x20 unit-weight grid (each
    # vertex to its four neighbours) as one community, a 4-cycle tied to its
    # far corner as another.  sigma from a corner to the border (the far
    # corner) is C(38,19) > INT_MAX, which DijkstraClusterEvaluator's (int)
    # conversion (DijkstraClusterEvaluator.h:172-175) turns into INT_MIN on
    # x86-64.
    R = 20
    s, d = [], []
    for r in range(R):
        for c in range(R):
            for dr, dc in ((0, 1), (1, 0), (0, -1), (-1, 0)):
                if 0 <= r + dr < R and 0 <= c + dc < R:
                    s.append(r * R + c), d.append((r + dr) * R + c + dc)
    far, n0 = R * R - 1, R * R
    for a, b in ((far, n0), (n0, far), (n0, n0 + 1), (n0 + 1, n0 + 2), (n0 + 2, n0 + 3), (n0 + 3, n0),
                 (n0 + 1, n0), (n0 + 2, n0 + 1), (n0 + 3, n0 + 2), (n0, n0 + 3)):
        s.append(a), d.append(b)
    gp = os.path.join(tmp, "bigsig.txt")
    with open(gp, "w") as f:
        f.writelines(f"{a} {b} 1\n" for a, b in zip(s, d))
    vp = os.path.join(tmp, "bigsig_verts.txt")
    np.savetxt(vp, np.arange(n0), fmt="%d")
    op = os.path.join(tmp, "bigsig_out.txt")
    run([REF, "cluster", gp, vp, op])
    lines = open(op).read().split("\n")
    rows = [ln.split() for ln in lines[1:] if ln]
    res["bigsig_src"], res["bigsig_dst"] = np.asarray(s, np.int32), np.asarray(d, np.int32)
    res["bigsig_part"] = np.repeat(np.int32([0, 1]), [n0, 4])
    res["bigsig_borders"] = np.array(lines[0].split()[1:], dtype=np.int32)
    res["bigsig_intra"] = np.array([float(r[1]) for r in rows])
    res["bigsig_len"] = np.array([[float(x) for x in r[2::2]] for r in rows])
    res["bigsig_cnt"] = np.array([[int(x) for x in r[3::2]] for r in rows], dtype=np.int32)
    assert (res["bigsig_cnt"] == np.iinfo(np.int32).min).any()
    np.savez_compressed(os.path.join(HERE, "degenerate.npz"), **res)


def part_cfg3(tmp):
    n, arcs = F.generate_config(3)
    p = os.path.join(tmp, "cfg3.txt")
    F.write_edge_list(n, arcs, p)
    out = os.path.join(tmp, "exact3.txt")
    thr = min(os.cpu_count() or 1, 8)
    r = subprocess.run([REF + "_fast", "exact", p, out], env=dict(os.environ, OMP_NUM_THREADS=str(thr)),
                       capture_output=True, text=True, check=True)
    np.savez_compressed(os.path.join(HERE, "cfg3_exact.npz"), bc=np.loadtxt(out), threads=np.array([thr]))
    print(r.stderr.strip())


def part_cfg2(tmp):
    n, arcs = F.generate_config(2)
    gp = os.path.join(tmp, "cfg2.txt")
    F.write_edge_list(n, arcs, gp)
    txt = run([REF + "_fast", "louvain", gp, "0.01", "1", "2", "3", "4"], env={"OMP_NUM_THREADS": "4"})
    comm = np.array(txt.split(), dtype=np.int64)
    assert len(comm) == n and comm.max() < 65536
    res = {"comm": comm.astype(np.uint16)}
    k = int(comm.max()) + 1
    s, d = arcs[0].astype(np.int64), arcs[1].astype(np.int64)
    cross = comm[s] != comm[d]
    isb = np.zeros(n, bool)
    isb[s[cross]] = True
    isb[d[cross]] = True
    sizes = np.bincount(comm, minlength=k)
    nb = np.bincount(comm[isb], minlength=k)
    small = [int(c) for c in np.argsort(sizes, kind="stable") if nb[c] > 0][:3]
    res["clusters"] = np.array(small, np.int32)
    for i, c in enumerate(small):
        verts = np.where(comm == c)[0]
        vp = os.path.join(tmp, f"c{c}.txt")
        np.savetxt(vp, verts, fmt="%d")
        op = os.path.join(tmp, f"c{c}_out.txt")
        run([REF + "_fast", "cluster", gp, vp, op])
        lines = open(op).read().split("\n")
        borders = np.array(lines[0].split()[1:], dtype=np.int32)
        rows = [ln.split() for ln in lines[1:] if ln]
        assert [int(r[0]) for r in rows] == verts.tolist()
        L = np.array([[float(x) for x in r[2::2]] for r in rows])
        C = np.array([[int(x) for x in r[3::2]] for r in rows], dtype=np.int32)
        assert (L == np.floor(L)).all() and L.max() < 2 ** 32
        res[f"borders{i}"] = borders
        res[f"intra{i}"] = np.array([float(r[1]) for r in rows])
        res[f"shape{i}"] = np.array(L.shape, np.int64)
        res[f"len_sha{i}"] = sha(L.astype(np.uint32))
        res[f"cnt_sha{i}"] = sha(C)
        # exact classes (normalized rows, first appearance) as the selector
        # forms them, then the reference's PlusPlusKMeans on them, k = P/2
        Ln = L - L.min(axis=1, keepdims=True)
        key = {}
        reps, card = [], []
        for r in range(len(verts)):
            t = (tuple(C[r]), tuple(Ln[r]))
            if t not in key:
                key[t] = len(reps)
                reps.append(r)
                card.append(0)
            card[key[t]] += 1
        P = len(reps)
        kk = max(int(P * 0.5), 1)
        kp = os.path.join(tmp, f"km{c}.txt")
        with open(kp, "w") as f:
            f.write(f"{P} {L.shape[1]} {kk}\n")
            for j, r in enumerate(reps):
                f.write(str(card[j]) + " " + " ".join(f"{Ln[r, b]:.17g} {C[r, b]}" for b in range(L.shape[1])) + "\n")
        ko = os.path.join(tmp, f"km{c}_out.txt")
        run([REF + "_fast", "kmeans", kp, ko])
        km = np.loadtxt(ko, dtype=np.int64).reshape(-1, 2)
        res[f"classes{i}"] = np.array([P], np.int64)
        res[f"km_cent{i}"] = km[:, 0].astype(np.int32)     # class indices, before the dedupe
        res[f"km_w{i}"] = km[:, 1].astype(np.int32)
    op = os.path.join(HERE, "cfg2_oracle_pivots.npz")
    if os.path.exists(op):
        piv = np.load(op)
        srcs = []
        for c in np.unique(piv["pivot_cluster"])[[0, 50, 100, 150]]:
            srcs.append(int(piv["pivots"][piv["pivot_cluster"] == c][0]))
        rows = ref_dep_rows(REF + "_fast", gp, srcs, tmp)
        res["dep_src"] = np.array(srcs, np.uint32)
        res["dep_sha"] = np.stack([sha(np.asarray(r, np.float64)) for r in rows])
        res["dep_sum"] = np.array([float(np.sum(r)) for r in rows])
    np.savez_compressed(os.path.join(HERE, "cfg2_reference.npz"), **res)


def part_cfg2_oracle(tmp):
    """The oracle's whole cfg2 selection (slow: every cluster source)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n, (s, d, w) = F.generate_config(2)
    og = O.Graph.from_arcs(s, d, w.astype(np.float64))
    r = O.clustered_pivots(og, [1, 2, 3, 4], 0.01, kfrac=0.5, threads=min(os.cpu_count() or 1, 8))
    np.savez_compressed(os.path.join(HERE, "cfg2_oracle_pivots.npz"), pivots=r["pivots"].astype(np.uint32),
                        weights=r["weights"].astype(np.uint32), pivot_cluster=r["pivot_cluster"].astype(np.uint32),
                        intra_sha=sha(r["intra"]), comm_sha=sha(r["comm"].astype(np.int32)))


def part_cfg4(tmp):
    n, arcs = F.generate_config(4)
    gp = os.path.join(tmp, "cfg4.txt")
    F.write_edge_list(n, arcs, gp)
    txt = run([REF + "_fast", "louvain", gp, "0.01", "1", "2", "3", "4"], env={"OMP_NUM_THREADS": "4"})
    comm = np.array(txt.split(), dtype=np.int32)
    assert len(comm) == n
    srcs = [0, 2 * 1048576 + 1024, n - 1]
    rows = ref_dep_rows(REF + "_fast", gp, srcs, tmp)
    np.savez_compressed(os.path.join(HERE, "cfg4_reference.npz"), comm_sha=sha(comm),
                        ncomm=np.array([int(comm.max()) + 1]), dep_src=np.array(srcs, np.uint32),
                        dep_sha=np.stack([sha(np.asarray(r, np.float64)) for r in rows]),
                        dep_sum=np.array([float(np.sum(r)) for r in rows]))


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", default="base", choices=["base", "cfg3", "cfg2", "cfg2_oracle", "cfg4", "degenerate"])
    a = ap.parse_args()
    if a.part == "base":
        main()
    else:
        if not os.path.exists(REF):
            sys.exit("oracle/_ref/ref_driver is not built (make -C oracle ref, needs /root/reference)")
        t = tempfile.mkdtemp(prefix="fbc_golden_")
        {"cfg3": part_cfg3, "cfg2": part_cfg2, "cfg2_oracle": part_cfg2_oracle, "cfg4": part_cfg4,
         "degenerate": part_degenerate}[a.part](t)
        print("written", a.part)
