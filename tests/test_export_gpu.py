"""Device property export (mcst_run_get_properties) against the oracle's,
and the exporter driven by a HIP run."""
import numpy as np
import pytest

from parity import close, make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod", "monod", "fixed_length", "simple_acetate"])
def test_properties_match_oracle(model):
    from mcst import cases

    c = cases.small_case(model, n_c=10, n_particles=5000, seed=9)
    c.capacity = 50000
    eng, g, o = make_pair(c)
    eng.step(30)
    o.step(30)
    vg, sg = g.properties()
    vo, so = o.properties()
    close(vg, vo, 1e-12, "values")           # the state (within the parity bound) and its mass
    np.testing.assert_array_equal(vg == 0.0, vo == 0.0)  # the same slots are Idle
    close(sg, so, 1e-12, "spatial sums")     # FP64 atomics: order differs
    eng.close()


def test_exporter_on_a_hip_run(tmp_path):
    from mcst import cases, export

    c = cases.ring64(n_particles=20000)
    eng, g, o = make_pair(c)
    ex = export.Exporter(str(tmp_path / "run.npz"), c.model)
    for _ in range(2):
        eng.step(50)
        ex.write(g)
    ex.write_probes(g)
    d = export.read(ex.close())
    o.step(100)
    vo, so = o.properties()
    close(d["biological_model/1/m"], vo[0], 1e-12, "m")
    close(d["biological_model/1/spatial/mass"], so[-1], 1e-12, "spatial mass")
    np.testing.assert_array_equal(d["records/number_particle"][1], o.repartition())
    eng.close()
