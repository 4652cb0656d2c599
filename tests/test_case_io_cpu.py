"""Case reader: the reference's 0-D case files, write/read round trips, and
runs built from a case on disk equal runs built from the same arrays."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_reads_the_reference_0d_case():
    import json

    from mcst import case_io

    d = os.path.join(GOLD, "case_0d")
    assert case_io.read_case_index(os.path.join(d, "cma_case")) == ["./vofL.raw", "./flowL.raw"]
    cma = case_io.load_case(d)
    kat = json.load(open(os.path.join(GOLD, "reference_kats.json")))["case_0d"]
    np.testing.assert_array_equal(cma.volumes[0], kat["volumes"])
    assert cma.n_c == 1 and cma.n_maps == 1 and len(cma.flows[0][2]) == 0
    np.testing.assert_array_equal(case_io.read_raw_vector(os.path.join(d, "turbE.raw.raw")), [0.5])
    # a 0-D case keeps the user's dt and has one step per flowmap
    assert case_io.time_step(cma, 0.5) == 0.5
    assert case_io.steps_per_flowmap(10.0, cma.n_maps, 0.5) == 1


@pytest.mark.parametrize("sparse", [False, True])
def test_round_trip_of_a_cyclic_case(tmp_path, sparse):
    from mcst import case_io, cases

    c = cases.rgg5000(n_particles=0, n_c=300, n_maps=4, steps_per_map=7)
    case_io.write_case(str(tmp_path), [m[0] for m in c.maps], [m[1:] for m in c.maps], sparse=sparse)
    cma = case_io.load_case(str(tmp_path))
    assert cma.n_maps == 4 and cma.n_c == 300
    for (V, r, col, q), V2, (r2, c2, q2) in zip(c.maps, cma.volumes, cma.flows):
        np.testing.assert_array_equal(V, V2)
        a = {(int(i), int(j)): v for i, j, v in zip(r, col, q)}
        b = {(int(i), int(j)): v for i, j, v in zip(r2, c2, q2)}
        assert a == b


def test_time_step_and_flowmap_period():
    from mcst import case_io

    V = np.array([1.0, 2.0, 4.0])
    cma = case_io.CmaCase([V], [(np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32),
                                 np.array([0.5, 0.25, 2.0]))])
    # residence times V/Q_out = 2, 8, 2 -> dt = 2/100
    assert case_io.min_residence_time(cma) == 2.0
    assert case_io.time_step(cma, 0.0) == 0.02
    assert case_io.time_step(cma, 0.1) == 0.1
    assert case_io.steps_per_flowmap(0.0, 5, 0.1) == 1
    assert case_io.steps_per_flowmap(1.0, 1, 0.1) == 1
    assert case_io.steps_per_flowmap(1.05, 5, 0.1) == int(1.05 / 0.1) + 1
    closed = case_io.CmaCase([V, V], [(np.zeros(0, np.uint32),) * 2 + (np.zeros(0),)] * 2)
    with pytest.raises(ValueError):
        case_io.time_step(closed, 0.0)


def test_run_from_disk_equals_run_from_arrays(tmp_path, oracle_lib):
    from mcst import case_io, cases

    c = cases.small_case("monod_mass", n_c=12, n_particles=3000, seed=3, n_maps=2, steps_per_map=3)
    case_io.write_case(str(tmp_path), [m[0] for m in c.maps], [m[1:] for m in c.maps])
    cma = case_io.load_case(str(tmp_path))
    d = case_io.to_case(cma, c.model, c.C0, dt=c.dt, t_per_flowmap=2 * c.dt, feeds=c.feeds, params=c.params,
                        n_particles=c.n_particles, seed=c.seed)
    assert d.steps_per_map == c.steps_per_map
    runs = []
    for case in (c, d):
        r = oracle_lib.OracleRun(case.model, case.n_c, case.n_s, case.dt, case.seed)
        case.apply(r)
        r.step(20)
        runs.append(r)
    a, b = runs
    assert a.counters() == b.counters()
    np.testing.assert_array_equal(a.particles()["pos"], b.particles()["pos"])
    np.testing.assert_array_equal(a.particles()["state"], b.particles()["state"])
    np.testing.assert_array_equal(a.concentrations(), b.concentrations())
