"""A run built from a case directory on disk, on the HIP engine, matches the
oracle run built from the same files."""
import pytest

from parity import assert_parity, make_pair

pytestmark = pytest.mark.gpu


def test_hip_run_from_a_case_directory(tmp_path):
    from mcst import case_io, cases

    c = cases.rgg5000(n_particles=200_000, n_c=400, n_maps=3, steps_per_map=5)
    case_io.write_case(str(tmp_path), [m[0] for m in c.maps], [m[1:] for m in c.maps])
    cma = case_io.load_case(str(tmp_path))
    d = case_io.to_case(cma, c.model, c.C0, dt=c.dt, t_per_flowmap=4 * c.dt, params=c.params,
                        n_particles=c.n_particles, seed=c.seed)
    d.capacity = 2 * c.n_particles
    assert d.steps_per_map == 5
    eng, g, o = make_pair(d)
    eng.step(17)
    o.step(17)
    assert_parity(g, o, probes=False)
    eng.close()
