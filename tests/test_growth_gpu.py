"""Container growth (mcst_run_set_growth; the reference's merge_buffer /
_resize with its 2.5 allocation factor): runs that outgrow their capacity get
more slots, in place or in a new range of the engine's pool (grown when
full), and match the oracle applying the same rule on the same steps."""
import numpy as np
import pytest

from parity import assert_parity, make_pair

pytestmark = pytest.mark.gpu


def test_tight_run_grows_and_matches_oracle():
    import mcst
    from mcst import cases

    c = cases.small_case("monod_mass", n_particles=3000, seed=8)
    eng, g, o = make_pair(c, capacity=3100)
    for r in (g, o):
        r.set_growth(1, 2.5)
    with pytest.raises(mcst.MCSTError):
        eng.capture(2)
    for _ in range(5):
        eng.step(4)
        o.step(4)
        assert_parity(g, o)
    assert g.n > 3100
    eng.close()


def test_all_dividing_population_doubles_every_step():
    """Every cell divides every step; with a check before every step nothing
    is refused: n = 1000 * 2^k, as in the oracle."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    c = cases.small_case("monod_mass", n_c=1, n_particles=1000, seed=2, feeds=False)
    c.params = np.array([0.5, 0.01, 0.5, 0.0, 1e-15, 0.1])
    eng, g, o = make_pair(c, capacity=1000)
    for r in (g, o):
        r.set_growth(1, 2.5)
    for k in range(1, 8):
        eng.step(1)
        o.step(1)
        assert g.n == 1000 * 2 ** k and g.counters()["overflow"] == 0
    assert_parity(g, o)
    eng.close()


def test_batch_runs_relocate_and_the_pool_grows():
    """Two runs filling the pool: the first cannot extend in place (the second
    follows it), so it moves to a new range after the pool grows; the second
    then extends in place.  Checks on different intervals."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    eng = mcst.Engine("monod_mass")
    pairs = []
    for seed, n, every in ((31, 2000, 2), (32, 2500, 3)):
        c = cases.small_case("monod_mass", n_particles=n, seed=seed)
        c.capacity = n + 100
        g = eng.add_run(c.n_c, c.dt, c.seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        for r in (g, o):
            c.apply(r)
            r.set_growth(every, 2.5)
        g.set_probes(1 << 16)
        pairs.append((g, o))
    eng.commit()
    for _ in range(4):
        eng.step(6)
        for g, o in pairs:
            o.step(6)
            assert_parity(g, o)
    for (g, o), n in zip(pairs, (2000, 2500)):
        assert g.n > n + 100
    eng.close()


def test_growth_arguments_and_profiled_steps():
    """A factor outside (1, 16] is refused; steps taken by the profiler are
    ordinary steps (capacity checks included) and match the oracle."""
    import mcst
    from mcst import cases

    c = cases.small_case("monod_mass", n_particles=3000, seed=9)
    eng, g, o = make_pair(c, capacity=3100)
    for bad in (1.0, 0.5, 17.0):
        with pytest.raises(mcst.MCSTError):
            g.set_growth(1, bad)
    for r in (g, o):
        r.set_growth(2, 2.5)
    eng.profile(10)
    o.step(10)
    assert_parity(g, o)
    assert g.counters()["steps"] == 10 and g.n > 3100
    eng.close()
