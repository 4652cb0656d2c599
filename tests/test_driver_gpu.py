"""The C++ host driver example (biocma-mcst_amd/examples/ring_driver.cpp),
which uses only the C ABI, against the oracle on the same ring case."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "..", "biocma-mcst_amd", "examples", "ring_driver")


def _oracle(n_c, n_p, steps, seed):
    from oracle.oracle import OracleRun

    o = OracleRun("monod_mass", n_c, 1, 0.01, seed)
    o.set_params([0.5, 0.01, 0.5, 2.0, 1e-15, 0.1])
    rows = np.array(sum(([i, (i + 1) % n_c] for i in range(n_c)), []), np.uint32)
    cols = np.array(sum(([(i + 1) % n_c, i] for i in range(n_c)), []), np.uint32)
    o.add_flowmap(np.full(n_c, 1e-3), rows, cols, np.full(2 * n_c, 2e-3))
    o.add_feed("constant", 2e-3, 0.0, 0, n_c // 2, [0], [10.0])
    o.set_concentrations(np.full((1, n_c), 5.0))
    o.init_particles(n_p, 10 * n_p, 1.0)
    o.step(steps)
    return o


@pytest.mark.parametrize("args", [(32, 20000, 50, 7), (5, 3000, 33, 11)])
def test_cpp_driver_matches_oracle(args):
    assert os.path.exists(DRIVER), "build the library first (python biocma-mcst_amd/build.py)"
    out = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d = json.loads(out.stdout.strip().splitlines()[-1])
    o = _oracle(*args)
    c = o.counters()
    for k in ("steps", "births", "exits", "moves"):
        assert d[k] == c[k], (k, d[k], c[k])
    p = o.particles()
    n = len(p["pos"])
    assert d["n"] == n and d["idle"] == int((p["status"] == 0).sum())
    w = (np.arange(n) % 7 + 1).astype(np.uint64)
    assert d["pos_checksum"] == int((p["pos"].astype(np.uint64) * w).sum())
    assert math.isclose(d["glucose_mass"], float(o.mass().sum()), rel_tol=1e-12)
