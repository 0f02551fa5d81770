"""The drop-in surfaces: the reference-named C++ API (tests/cpp/test_api.cpp)
and the fbc command line (grammar, exit codes, bc.txt format of main.cpp)."""
import fcntl
import os
import subprocess

import numpy as np
import pytest

import fastbc_amd as F
import oracle_lib as O
from conftest import GOLDEN, ROOT

API = os.path.join(ROOT, "tests", "cpp", "test_api")
DWG9 = os.path.join(GOLDEN, "dwg9.txt")


@pytest.fixture(scope="module")
def cfg0_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cfg0")
    n, arcs = F.generate_config(0)
    g = str(d / "cfg0.txt")
    F.write_edge_list(n, arcs, g)
    lv = str(d / "louvain0.txt")
    np.savetxt(lv, np.load(os.path.join(GOLDEN, "louvain.npz"))["cfg0"], fmt="%d")
    return g, lv


def build_api():
    """make is incremental and the program depends on every host header and on
    the library, so a binary built against an older struct layout is rebuilt
    rather than run.  A file lock keeps parallel test workers from running the
    build at the same time (one could otherwise run a half-written binary)."""
    with open(os.path.join(os.path.dirname(API), ".build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        try:
            subprocess.run(["make", "-s", "-C", os.path.dirname(API)], check=True)
        finally:
            fcntl.flock(lk, fcntl.LOCK_UN)


def test_cpp_api_host_side(cfg0_files):
    build_api()
    r = subprocess.run([API, "cpu", DWG9, *cfg0_files], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_api_device_side(cfg0_files):
    build_api()
    r = subprocess.run([API, "gpu", DWG9, cfg0_files[0]], capture_output=True, text=True, timeout=600)
    # the program reports whether its plain-form call ran on a cooperative
    # launch or was refused one (another process holding the device's
    # cooperative-launch resources) and finished on the batched kernel
    print(r.stdout)
    print([ln for ln in r.stdout.splitlines() if ln.startswith("plain form:")])
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


def fbc(*args, cwd=None):
    return subprocess.run([F.FBC_BIN, *args], capture_output=True, text=True, cwd=cwd, timeout=600)


def test_cli_option_errors(tmp_path):
    """main.cpp:93-104,123-181: -1 (255) for option errors, -2 (254) for an
    existing output file."""
    out = str(tmp_path / "bc.txt")
    r = fbc("-e", "abc", DWG9)
    assert r.returncode == 255 and r.stdout.startswith("invalid argument for -e: 'abc'")
    assert fbc("-o", out).returncode == 255 and "Missing input file path" in fbc("-o", out).stdout
    assert fbc("-s", "1,2", DWG9, "-o", out).returncode == 255          # -s without -e
    r = fbc("-s", "1,1", "-e", "2", DWG9, "-o", out)
    assert r.returncode == 255 and "Duplicate value" in r.stdout
    r = fbc("-s", "1,2,3", "-e", "4", DWG9, "-o", out)
    assert r.returncode == 255 and "different from louvain executors" in r.stdout
    r = fbc("-k", "1.5", DWG9, "-o", out)
    assert r.returncode == 255 and "Kfrac value must be in range 0-1." in r.stdout
    open(out, "w").close()
    r = fbc("-o", out, DWG9)
    assert r.returncode == 254 and f'File "{out}" already existing' in r.stdout
    r = fbc("--bogus", "-o", str(tmp_path / "x.txt"), str(tmp_path / "missing.txt"))
    assert r.returncode == 255 and "error opening" in r.stdout     # unknown options are ignored


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["exact", "clustered", "2-clustered"])
def test_cli_end_to_end(tmp_path, cfg0_files, mode):
    """bc.txt: default ostream formatting, negatives written as 0
    (main.cpp:282-294); values against the oracle."""
    g = cfg0_files[0]
    out, full = str(tmp_path / "bc.txt"), str(tmp_path / "bc_full.txt")
    args = {"exact": ["--exact"], "clustered": ["-s", "1,2,3,4", "-e", "4"],
            "2-clustered": ["-s", "1,2,3,4", "-e", "4", "-p", "0.01", "-k", "0.5"]}[mode]
    r = fbc(*args, "-o", out, "--full-output", full, g)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[  info   ] Loaded graph contains 4096 vertices and 16128 edges" in r.stdout
    assert "Total computation time:" in r.stdout
    vals = np.loadtxt(full)
    og = O.Graph.from_file(g)
    if mode == "exact":
        ref = og.exact_bc(8)
    else:
        ref = og.clustered_bc([1, 2, 3, 4], 0.01, kfrac=0.5 if mode == "2-clustered" else 0.0, threads=1)["bc"]
    scale = np.abs(ref).max()
    assert np.abs(vals - ref).max() <= 1e-9 * scale
    expect = "".join(("%g" % x if x >= 0 else "0") + "\n" for x in vals)
    # %g and the default ostream format agree (6 significant digits)
    assert open(out).read() == expect


def _dump_oracle(path):
    try:
        g = O.Graph.from_file(path)
    except ValueError as e:
        return ("THROW", str(e))
    rp, col, w = g.csr()
    arcs = [(u, int(col[e]), float(w[e])) for u in range(g.n) for e in range(int(rp[u]), int(rp[u + 1]))]
    return (g.n, arcs)


def _dump_product(path, out):
    build_api()
    r = subprocess.run([API, "load", path, out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = open(out).read().splitlines()
    if lines[0].startswith("THROW"):
        return ("THROW", lines[0][6:])
    n = int(lines[0].split()[1])
    arcs = [(int(a), int(b), float(c)) for a, b, c in (x.split() for x in lines[1:])]
    return (n, arcs)


@pytest.mark.parametrize("case", ["plain", "split_triple", "late_bad_weight", "malformed_id", "partial_tail",
                                  "decimal_weights"])
def test_edge_list_loader_matches_oracle(tmp_path, case):
    """The loader parses inputs of a megabyte and more in parallel chunks cut
    at line ends; its stop rules (a malformed id ends the input, a weight <= 0
    or unreadable throws, a triple cut by the end is dropped) and repeated-arc
    sums must match the oracle's sequential DirectedWeightedGraph loader."""
    rng = np.random.default_rng(11)
    n, m = 3000, 120000
    s, d = rng.integers(0, n, m), rng.integers(0, n, m)
    s[:n] = np.arange(n)   # every vertex has an out-arc
    w = rng.integers(1, 1000, m).astype(float)
    if case == "decimal_weights":
        w = w + 0.375
    lines = [f"{a} {b} {c:g}" for a, b, c in zip(s, d, w)]
    mid = m // 2
    if case == "split_triple":
        lines[mid] = f"{s[mid]} {d[mid]}\n{w[mid]:g}"
    elif case == "late_bad_weight":
        lines[int(m * 0.9)] = "1 2 0"
    elif case == "malformed_id":
        lines[mid] = "1 x 3"
    text = "\n".join(lines) + "\n"
    if case == "partial_tail":
        text += "5 6"
    p = tmp_path / "g.txt"
    p.write_text(text)
    assert len(text) > (1 << 20)
    # the reference's stream loop (DirectedWeightedGraph.h:71-90) stops at a
    # malformed id and at a triple cut by the end (its stale weight passes the
    # check); the oracle reports those as errors, so it reads the text up to there
    expect_text = text
    if case == "malformed_id":
        expect_text = "\n".join(lines[:mid]) + "\n"
    elif case == "partial_tail":
        expect_text = text[: -len("5 6")]
    q = tmp_path / "expect.txt"
    q.write_text(expect_text)
    exp = _dump_oracle(str(q))
    got = _dump_product(str(p), str(tmp_path / "dump.txt"))
    if exp[0] == "THROW":
        assert got[0] == "THROW" and "greater than zero" in got[1] and "greater than zero" in exp[1]
    else:
        assert got[0] == exp[0] and got[1] == exp[1]


def test_csr_cache(tmp_path, cfg0_files):
    """The binary CSR cache (fbc --csr-cache): a second load reads the cache
    and gives the same graph, file-order sums included; a changed or damaged
    cache is not used."""
    build_api()
    text = open(cfg0_files[0]).read()
    extra = "0 1 7\n4095 4094 3\n0 1 2\n"          # repeated arcs: summed in file order
    src = tmp_path / "g.txt"
    src.write_text(text + extra)
    cache = str(tmp_path / "g.csr")

    def dump(name):
        out = str(tmp_path / name)
        r = subprocess.run([API, "cache", str(src), cache, out], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        lines = open(out).read().splitlines()
        return lines[0], lines[1:]

    h1, g1 = dump("a.txt")
    h2, g2 = dump("b.txt")
    assert (h1, h2) == ("hit 0", "hit 1") and g1 == g2
    ref = str(tmp_path / "ref.txt")
    subprocess.run([API, "load", str(src), ref], check=True)
    fwd = [ln for ln in g1 if not ln.startswith(("deg", "in", "n ", "total"))]
    assert fwd == open(ref).read().splitlines()[1:]
    with open(cache, "r+b") as f:                       # damaged: cut short
        f.truncate(os.path.getsize(cache) - 8)
    h3, g3 = dump("c.txt")
    assert h3 == "hit 0" and g3 == g1
    src.write_text(text + extra + "5 6 1\n")            # a changed input
    h4, g4 = dump("d.txt")
    assert h4 == "hit 0" and g4 != g1
    assert dump("e.txt") == ("hit 1", g4)
