"""mcst_run_add_flowmap_tables: the HIP step on flowmaps given as the
reference's tables (neighbour rows in rcmtool's order and padding, their
cumulative probabilities, out-flows, the transition matrix as COO) matches
the oracle given the same tables: events bit-exact, FP64 within 1e-12."""
import numpy as np
import pytest

from flow_tables import apply_tables, tables
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _pair_tables(case, **kw):
    import mcst
    from oracle.oracle import OracleRun

    eng = mcst.Engine(case.model)
    g = eng.add_run(case.n_c, case.dt, case.seed)
    o = OracleRun(case.model, case.n_c, case.n_s, case.dt, case.seed)
    for r in (g, o):
        apply_tables(case, r, **kw)
    g.set_probes(1 << 20)
    eng.commit()
    return eng, g, o


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod", "quad_monod"])
def test_shuffled_self_padded_tables_match_oracle(model):
    from mcst import cases

    c = cases.small_case(model, n_c=24, n_particles=12000, seed=17)
    eng, g, o = _pair_tables(c, order="shuffled", pad="self", extra=3, split=True)
    for _ in range(3):
        eng.step(8)
        o.step(8)
        assert_parity(g, o)
    assert g.counters()["moves"] > 0 and g.counters()["exits"] > 0
    eng.close()


def test_tables_of_the_flows_builder_give_the_same_hip_run():
    import mcst
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=30, n_particles=20000, seed=2)
    runs = []
    for use_tables in (False, True):
        eng = mcst.Engine(c.model)
        r = eng.add_run(c.n_c, c.dt, c.seed)
        if use_tables:
            apply_tables(c, r, order="ascending", pad="last")
        else:
            c.apply(r)
        eng.commit()
        eng.step(25)
        runs.append((eng, r))
    a, b = runs[0][1], runs[1][1]
    assert a.counters() == b.counters()
    pa, pb = a.particles(), b.particles()
    for k in ("pos", "status"):
        np.testing.assert_array_equal(pa[k], pb[k])
    np.testing.assert_allclose(pa["state"], pb["state"], rtol=1e-12)
    for eng, _ in runs:
        eng.close()


@pytest.mark.parametrize("bad", ["decreasing", "neighbour", "volume", "transition", "nan", "no_columns"])
def test_malformed_tables_are_refused(bad):
    import mcst

    n = 4
    rows = np.array([0, 1, 1, 2, 2, 3], np.uint32)
    cols = np.array([1, 0, 2, 1, 3, 2], np.uint32)
    t = list(tables(n, rows, cols, np.full(6, 1e-4)))
    V = np.full(n, 1e-3)
    if bad == "decreasing":
        t[1][1, 0], t[1][1, 1] = 0.9, 0.1
    elif bad == "neighbour":
        t[0][2, 0] = n
    elif bad == "volume":
        V[3] = -1.0
    elif bad == "transition":
        t[4][0] = n
    elif bad == "nan":
        t[5][0] = np.nan
    else:
        t[0], t[1] = np.zeros((n, 0), np.uint64), np.zeros((n, 0))
    eng = mcst.Engine("monod_mass")
    r = eng.add_run(n, 0.01, 1)
    with pytest.raises(mcst.MCSTError):
        r.add_flowmap_tables(V, *t)
    eng.close()
