"""mcst_run_get_repartition_noalloc, the entry point that allocates no device
memory (its counters live in a library-owned device array per GPU), gives
what mcst_run_get_repartition gives: two calls in a row, the second on a
smaller run (the array keeps the first call's counts and must be cleared),
then again after more steps; a run beyond the array's size is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, steps):
    import mcst

    eng = mcst.Engine(case.model)
    r = eng.add_run(case.n_c, case.dt, case.seed)
    case.apply(r)
    eng.commit()
    eng.step(steps)
    return eng, r


def test_noalloc_repartition_matches_the_allocating_one():
    from mcst import cases

    big = cases.small_case("monod_mass", n_c=3000, n_particles=60000, seed=4)
    small = cases.small_case("dual_monod", n_c=40, n_particles=9000, seed=5)
    e1, r1 = _run(big, 6)
    e2, r2 = _run(small, 9)
    a = r1.repartition_noalloc()
    np.testing.assert_array_equal(a, r1.repartition())
    assert a.sum() == r1.counters()["n"] - r1.counters()["inactive"]
    b = r2.repartition_noalloc()  # second call, smaller input
    np.testing.assert_array_equal(b, r2.repartition())
    assert b.sum() > 0
    e2.step(4)
    e1.step(2)
    np.testing.assert_array_equal(r2.repartition_noalloc(), r2.repartition())
    np.testing.assert_array_equal(r1.repartition_noalloc(), r1.repartition())
    e1.close()
    e2.close()


def test_noalloc_repartition_refuses_runs_beyond_its_array():
    import mcst

    n_c = (1 << 20) + 1
    eng = mcst.Engine("monod_mass")
    r = eng.add_run(n_c, 0.01, 3)
    r.add_flowmap(np.full(n_c, 1e-3), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r.set_concentrations(np.zeros((1, n_c)))
    r.init_particles(2000, 4000, 1.0)
    eng.commit()
    with pytest.raises(mcst.MCSTError, match="allocation-free repartition holds"):
        r.repartition_noalloc()
    assert r.repartition().sum() == 2000  # the allocating entry point still serves it
    eng.close()
