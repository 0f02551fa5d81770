import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "fast-bc_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # The oracle is the checker of every test; build it if the tree has none.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def read_arcs(path):
    import numpy as np
    a = np.loadtxt(path, dtype=np.int64, ndmin=2)
    return a[:, 0], a[:, 1], a[:, 2]


@pytest.fixture(scope="session")
def gpu_ctx():
    import fastbc_amd
    ctx = fastbc_amd.Context(0)
    yield ctx
    ctx.close()
