"""A C++ driver loads a case directory with the C++ case loader and runs it
over the C ABI (biocma-mcst_amd/examples/case_driver); the same case read by
the Python reader runs on the oracle: counters and the position checksum
bit-exact, liquid masses within 1e-12.  Cases: the reference's 0-D fixture
and a three-flowmap network cycled every few steps."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "biocma-mcst_amd", "examples", "case_driver")


def _oracle(path, model, dt, t_map, n_p, steps, seed, c0, feed):
    from mcst import case_io
    from oracle.oracle import OracleRun

    cma = case_io.load_case(path)
    d_t = case_io.time_step(cma, dt)
    nv, ns, _ = __import__("oracle.oracle", fromlist=["x"]).model_info(model)
    o = OracleRun(model, cma.n_c, ns, d_t, seed)
    for V, (r, c, q) in zip(cma.volumes, cma.flows):
        o.add_flowmap(V, r, c, q)
    o.set_steps_per_map(case_io.steps_per_flowmap(t_map, cma.n_maps, d_t))
    if feed:
        o.add_feed("constant", feed[2], 0.0, feed[0], None if feed[1] < 0 else feed[1], [0], [feed[3]])
    o.set_concentrations(np.full((ns, cma.n_c), c0))
    o.init_particles(n_p, 4 * n_p + 1024, 1.0)
    o.step(steps, threads=4)
    return o


def _check(path, model, dt, t_map, n_p, steps, seed, c0, feed=None):
    args = [DRIVER, path, model, repr(dt), repr(t_map), str(n_p), str(steps), str(seed), repr(c0)]
    if feed:
        args += [str(x) for x in feed]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    o = _oracle(path, model, dt, t_map, n_p, steps, seed, c0, feed)
    assert got["counters"] == [v for v in o.counters().values()]
    p = o.particles()["pos"]
    assert got["pos_checksum"] == int(np.sum(p.astype(np.uint64) * (np.arange(len(p)) % 7 + 1).astype(np.uint64)))
    np.testing.assert_allclose(np.array(got["mass"]), o.mass().ravel(), rtol=1e-12)
    return got


def test_driver_on_the_reference_0d_fixture():
    got = _check(os.path.join(HERE, "golden", "case_0d"), "monod_mass", 0.5, 0.0, 5000, 40, 3, 5.0)
    assert got["n_c"] == 1 and got["counters"][0] == 40


def test_driver_on_a_multi_flowmap_case(tmp_path):
    from mcst import case_io, cases

    c = cases.small_case("monod_mass", n_c=60, n_maps=3, seed=12)
    case_io.write_case(str(tmp_path), [m[0] for m in c.maps], [m[1:] for m in c.maps])
    got = _check(str(tmp_path), "dual_monod", 0.01, 0.035, 30000, 37, 9, 2.0, feed=(0, 30, 2e-3, 10.0))
    assert got["n_maps"] == 3 and got["steps_per_map"] == 4
    assert got["counters"][4] > 0 and got["counters"][2] > 0  # moves and exits happened
