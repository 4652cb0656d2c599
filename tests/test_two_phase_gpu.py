"""Two-phase runs on the HIP engine (a transported gas phase exchanging
oxygen with the liquid: SimulationUnit::ode_step with a gas phase,
apps/libs/simulation/src/simulation.model.cpp:131-150) against the oracle:
particles and events bit-exact, liquid and gas concentrations and masses
within 1e-12; a checkpoint restores the gas phase too."""
import numpy as np
import pytest

from parity import assert_parity, close, make_pair

pytestmark = pytest.mark.gpu


def _gas_parity(g, o):
    assert_parity(g, o)
    close(g.gas_concentrations(), o.gas_concentrations(), 1e-12, "gas concentrations")
    close(g.gas_mass(), o.gas_mass(), 1e-12, "gas mass")


@pytest.mark.parametrize("kla", ["turbulence", "fixed"])
def test_bubble_column_matches_oracle(kla):
    from mcst import cases

    c = cases.bubble_column(n_particles=30000, kla=kla)
    if kla == "fixed":  # a linear gas feed and a linear liquid feed on top
        c.two_phase["gas_feeds"].append(("linear", 1e-5, 1e-5, 4, None, [1], [0.3]))
        c.feeds.append(("linear", 1e-6, 1e-6, 2, None, [0], [5.0]))
    eng, g, o = make_pair(c)
    for _ in range(4):
        eng.step(15)
        o.step(15)
        _gas_parity(g, o)
    # the gas phase moved and exchanged oxygen with the liquid
    cg = g.gas_concentrations()[1]
    assert cg.max() - cg.min() > 1e-3
    eng.close()


def test_two_phase_checkpoint_restores_the_gas_phase():
    import mcst
    from mcst import cases

    c = cases.bubble_column(n_particles=15000)
    eng, g, o = make_pair(c)
    eng.step(23)
    blob = g.checkpoint()
    eng.step(17)
    o.step(40)
    _gas_parity(g, o)
    eng2 = mcst.Engine(c.model)
    h = eng2.add_run(c.n_c, c.dt, c.seed)
    c.apply(h, particles=False)
    h.set_probes(1 << 20)
    h.restore(blob)
    eng2.commit()
    eng2.step(17)
    _gas_parity(h, o)
    eng.close()
    eng2.close()


def test_two_phase_definition_errors():
    import mcst
    from mcst import cases

    c = cases.bubble_column(n_particles=2000)
    c.two_phase["gas_maps"] = c.two_phase["gas_maps"][:1]  # one gas flowmap for two flowmaps
    eng = mcst.Engine(c.model)
    r = eng.add_run(c.n_c, c.dt, c.seed)
    c.apply(r)
    with pytest.raises(mcst.MCSTError, match="one gas flowmap per flowmap"):
        eng.commit()
    eng.close()
    c = cases.bubble_column(n_particles=2000)
    eng = mcst.Engine(c.model)
    r = eng.add_run(c.n_c, c.dt, c.seed)
    c.apply(r)
    r.set_gas_transfer([0.0, 0.1], [0.0, 8e-3])
    with pytest.raises(mcst.MCSTError, match="both a gas phase"):
        eng.commit()
    eng.close()
    eng = mcst.Engine("dual_monod")
    r = eng.add_run(4, 0.01, 1)
    with pytest.raises(mcst.MCSTError):
        r.set_kla("turbulence", species=2)
    with pytest.raises(mcst.MCSTError):
        r.set_gas_phase([0.0, -1.0], np.zeros((2, 4)))
    r.set_kla("turbulence", species=1)
    r.add_gas_flowmap(np.full(4, 1e-4), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    with pytest.raises(mcst.MCSTError):  # no run without a gas phase has gas results
        r.gas_concentrations()
    eng.close()
