"""The C++ case loader (biocma-mcst_amd/host/case_loader.cpp), through the
case driver's --dump mode (no GPU): the same volumes, flow entries, time
step and flowmap period as the Python reader (mcst/case_io.py) on the
reference's 0-D fixture (apps/api/tests/data/0d) and on multi-flowmap cases
in both flow encodings; malformed files are refused with the file named."""
import json
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "biocma-mcst_amd", "examples", "case_driver")


def dump(path, dt=0.0, t_map=0.0):
    out = subprocess.run([DRIVER, "--dump", path, repr(dt), repr(t_map)], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(out.stderr.strip())
    return json.loads(out.stdout)


def _compare(path, dt, t_map):
    from mcst import case_io

    got = dump(path, dt, t_map)
    cma = case_io.load_case(path)
    d_t = case_io.time_step(cma, dt)
    assert got["n_c"] == cma.n_c and got["n_maps"] == cma.n_maps
    assert got["dt"] == d_t
    assert got["steps_per_map"] == case_io.steps_per_flowmap(t_map, cma.n_maps, d_t)
    for m, V, (r, c, q) in zip(got["maps"], cma.volumes, cma.flows):
        np.testing.assert_array_equal(np.array(m["V"]), V)
        np.testing.assert_array_equal(np.array(m["rows"], np.uint32), r)
        np.testing.assert_array_equal(np.array(m["cols"], np.uint32), c)
        np.testing.assert_array_equal(np.array(m["flows"]), q)


def test_reference_0d_fixture():
    _compare(os.path.join(HERE, "golden", "case_0d"), 0.5, 0.0)
    got = dump(os.path.join(HERE, "golden", "case_0d"), 0.5)
    assert got["maps"][0]["V"] == [0.02] and got["maps"][0]["flows"] == []


@pytest.mark.parametrize("sparse", [False, True])
def test_multi_flowmap_case(tmp_path, sparse):
    from mcst import case_io, cases

    c = cases.small_case("monod_mass", n_c=40, n_maps=3, seed=8)
    case_io.write_case(str(tmp_path), [m[0] for m in c.maps], [m[1:] for m in c.maps], sparse=sparse)
    _compare(str(tmp_path), 0.0, 0.0)      # dt from the residence times, one step per map
    _compare(str(tmp_path), 0.01, 0.037)   # the user's dt, 4 steps per map


@pytest.mark.parametrize("bad", ["short_vector", "short_matrix", "shape", "entry"])
def test_malformed_case_files_are_refused(tmp_path, bad):
    from mcst import case_io

    d = str(tmp_path)
    case_io.write_raw_vector(os.path.join(d, "vofL.raw"), np.full(3, 1e-3))
    case_io.write_raw_matrix(os.path.join(d, "flowL.raw"), np.full((3, 3), 1e-4))
    if bad == "short_vector":
        with open(os.path.join(d, "vofL.raw"), "r+b") as f:
            f.truncate(12)
    elif bad == "short_matrix":
        with open(os.path.join(d, "flowL.raw"), "r+b") as f:
            f.truncate(20)
    elif bad == "shape":
        case_io.write_raw_matrix(os.path.join(d, "flowL.raw"), np.full((2, 3), 1e-4))
    else:
        os.remove(os.path.join(d, "flowL.raw"))
        case_io.write_coo(os.path.join(d, "flowL.coo"), [0, 5], [1, 0], [1e-4, 1e-4])
    with pytest.raises(RuntimeError, match="vofL|flowL|compartments"):
        dump(d, 0.01)
