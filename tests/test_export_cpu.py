"""Property export (get_properties) on the oracle, pinned independently, and
the exporter's dataset layout."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod", "monod", "fixed_length", "simple_acetate",
                                   "quad_monod"])
def test_oracle_properties_match_independent_sums(oracle_lib, model):
    from mcst import cases

    c = cases.small_case(model, n_c=10, n_particles=3000, seed=4)
    c.capacity = 30000
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    r.step(25)
    values, spatial = r.properties()
    p = r.particles()
    idle = p["status"] == 0
    nv = r.n_var
    # per-particle values: the state of Idle slots, zero elsewhere
    np.testing.assert_array_equal(values[:nv], np.where(idle, p["state"], 0.0))
    # mass closed forms (Model::mass): monod_mass/dual_monod carry it in x0;
    # the length models convert length with the linear density of the cell
    if model in ("monod_mass", "dual_monod", "quad_monod"):
        np.testing.assert_array_equal(values[nv], np.where(idle, p["state"][0], 0.0))
    else:
        idx = {"monod": 4, "fixed_length": 3}.get(model)  # d_m in the parameter vector
        d_m = 0.6e-6 if idx is None or c.params is None else c.params[idx]
        lin = 1000.0 * math.pi * d_m * d_m / 4.0
        np.testing.assert_allclose(values[nv], np.where(idle, p["state"][0] * lin, 0.0), rtol=1e-15)
    # spatial sums: an exactly rounded sum per compartment
    for k in range(nv + 1):
        ref = np.array([math.fsum(values[k][(p["pos"] == j) & idle]) for j in range(c.n_c)])
        np.testing.assert_allclose(spatial[k], ref, rtol=1e-13, atol=0)
    # the repartition is the count behind the sums
    assert r.repartition().sum() == idle.sum()


def test_exporter_layout(oracle_lib, tmp_path):
    from mcst import cases, export

    c = cases.small_case("monod", n_c=6, n_particles=1000, seed=8)
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    r.dt = c.dt
    ex = export.Exporter(str(tmp_path / "out.npz"), "monod")
    for _ in range(3):
        r.step(5)
        ex.write(r)
    ex.write_probes(r)
    d = export.read(ex.close())
    for k in range(3):
        for name in ("length", "mu", "mu_eff", "mass", "age", "age_hydro"):
            assert d[f"biological_model/{k}/{name}"].shape == (r.n,) or k < 2
        for name in ("length", "mu", "mu_eff", "mass"):
            assert d[f"biological_model/{k}/spatial/{name}"].shape == (c.n_c,)
    values, spatial = r.properties()
    np.testing.assert_array_equal(d["biological_model/2/mu"], values[2])
    np.testing.assert_array_equal(d["biological_model/2/spatial/mass"], spatial[-1])
    assert d["records/number_particle"].shape == (3, c.n_c)
    np.testing.assert_array_equal(d["records/number_particle"][2], r.repartition())
    t = d["records/tallies"]
    assert t.shape == (3, 10) and list(t[:, 0]) == [5, 10, 15]
    np.testing.assert_allclose(d["records/time"], [0.05, 0.10, 0.15])
    assert d["records/concentration_liquid"].shape == (3, 1, c.n_c)
    assert "probes/exit_ages" in d and "probes/division_ages" in d
