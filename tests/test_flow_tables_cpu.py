"""The oracle's flowmap from ready-made tables (the reference's
ReactorDomain::update + set_transition inputs): tables equal to what the
flows-based builder derives give the same run bit for bit; tables in another
neighbour order and padding give the reference's move law
P(j | leave) = F(i,j) / sum_k F(i,k); malformed tables are refused."""
import json
import os

import numpy as np
import pytest
from scipy import stats

from flow_tables import apply_tables, tables

HERE = os.path.dirname(os.path.abspath(__file__))


def _same(a, b):
    assert a.counters() == b.counters()
    pa, pb = a.particles(), b.particles()
    for k in ("pos", "status", "state", "age_hyd", "age_div"):
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    np.testing.assert_array_equal(a.concentrations(), b.concentrations())
    np.testing.assert_array_equal(a.mass(), b.mass())


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod"])
def test_tables_of_the_flows_builder_give_the_same_run(oracle_lib, model):
    from mcst import cases

    c = cases.small_case(model, n_c=14, n_particles=6000, seed=9)
    a = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    b = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(a)
    apply_tables(c, b, order="ascending", pad="last")
    for _ in range(3):
        a.step(7)
        b.step(7)
        _same(a, b)


def test_neighbour_order_and_padding_keep_the_move_law(oracle_lib):
    """The reference's 16-compartment flow fixture
    (apps/libs/cma_utils/tests/test_transport.cpp:96-373) given with rows in
    a shuffled order, padded with the compartment's own id: the destinations
    of the particles that leave compartment i follow F(i,j)/sum_k F(i,k)."""
    kats = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))
    F = np.array(kats["transport16_flows"]["flows"])
    n = 16
    ri, ci = np.nonzero(F)
    keep = ri != ci
    rows, cols, flows = ri[keep].astype(np.uint32), ci[keep].astype(np.uint32), F[ri, ci][keep]
    V = np.full(n, 1e-3)
    t = tables(n, rows, cols, flows, order="shuffled", pad="self", extra=2, split=True, seed=4)
    out = t[2]
    dt = 0.2 / (out.max() / V[0])  # p_move up to 0.2 per step
    r = oracle_lib.OracleRun("monod_mass", n, 1, dt, 21)
    r.add_flowmap_tables(V, *t)
    r.set_concentrations(np.zeros((1, n)))
    n_p = 32000
    pos0 = (np.arange(n_p) % n).astype(np.uint32)
    x = np.stack([np.full(n_p, 1e-15), np.full(n_p, 1e-15)])
    r.set_particles(n_p, pos0, np.zeros(n_p, np.uint8), x, np.zeros(n_p), np.zeros(n_p), 1.0)
    r.step_stages(8)  # move only
    pos1 = r.particles()["pos"]
    moved = pos1 != pos0
    for i in range(n):
        sel = (pos0 == i) & moved
        if out[i] == 0.0:
            assert not sel.any()
            continue
        p_exp = F[i] / out[i]
        p_exp[i] = 0.0
        obs = np.bincount(pos1[sel], minlength=n)
        assert np.all(obs[p_exp == 0] == 0), i  # never to a compartment without flow
        nz = p_exp > 0
        if obs.sum() > 50 and nz.sum() > 1:
            assert stats.chisquare(obs[nz], p_exp[nz] * obs.sum()).pvalue > 1e-4, i
        # the leaving fraction: dt*Q/V
        frac = sel.sum() / (pos0 == i).sum()
        pm = dt * out[i] / V[i]
        assert abs(frac - pm) < 5 * np.sqrt(pm * (1 - pm) / (pos0 == i).sum()) + 1e-12, i


def test_reference_neighbour_fixture_table_is_accepted(oracle_lib):
    """The neighbour table of the reference's test, as written (three columns,
    rcmtool's order, padded with the compartment's own id), with flows on its
    listed neighbours: moves only go to listed neighbours."""
    kats = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))
    nbr = np.array(kats["transport16_neighbors"]["neighbors"], np.uint64)
    n, m = nbr.shape
    rng = np.random.default_rng(3)
    cdf = np.zeros((n, m))
    out = np.zeros(n)
    trip = []
    for i in range(n):
        q = [rng.uniform(1e-5, 1e-4) if (nbr[i, k] != i and nbr[i, k] not in nbr[i, :k]) else 0.0 for k in range(m)]
        out[i] = sum(q)
        run = 0.0
        for k in range(m):
            run += q[k]
            cdf[i, k] = run / out[i] if out[i] > 0 else 0.0
            if q[k] > 0:
                trip.append((i, int(nbr[i, k]), q[k]))
        trip.append((i, i, -out[i]))
    tr, tc, tv = (np.array([x[0] for x in trip], np.uint64), np.array([x[1] for x in trip], np.uint64),
                  np.array([x[2] for x in trip]))
    r = oracle_lib.OracleRun("monod_mass", n, 1, 2.0, 5)
    r.add_flowmap_tables(np.full(n, 1e-3), nbr, cdf, out, tr, tc, tv)
    r.set_concentrations(np.zeros((1, n)))
    n_p = 16000
    pos0 = (np.arange(n_p) % n).astype(np.uint32)
    x = np.stack([np.full(n_p, 1e-15), np.full(n_p, 1e-15)])
    r.set_particles(n_p, pos0, np.zeros(n_p, np.uint8), x, np.zeros(n_p), np.zeros(n_p), 1.0)
    r.step_stages(8)
    pos1 = r.particles()["pos"]
    for i in range(n):
        dest = set(pos1[(pos0 == i) & (pos1 != pos0)].tolist())
        assert dest <= set(int(v) for v in nbr[i] if v != i), (i, dest)
    assert (pos1 != pos0).sum() > 0


@pytest.mark.parametrize("bad", ["decreasing", "neighbour", "volume", "transition", "nan"])
def test_malformed_tables_are_refused(oracle_lib, bad):
    n = 4
    rows = np.array([0, 1, 1, 2, 2, 3], np.uint32)
    cols = np.array([1, 0, 2, 1, 3, 2], np.uint32)
    t = list(tables(n, rows, cols, np.full(6, 1e-4)))
    V = np.full(n, 1e-3)
    if bad == "decreasing":
        t[1] = t[1].copy()
        t[1][1, 0], t[1][1, 1] = 0.9, 0.1
    elif bad == "neighbour":
        t[0] = t[0].copy()
        t[0][2, 0] = n
    elif bad == "volume":
        V[3] = 0.0
    elif bad == "transition":
        t[4] = t[4].copy()
        t[4][0] = n
    else:
        t[5] = t[5].copy()
        t[5][0] = np.nan
    r = oracle_lib.OracleRun("monod_mass", n, 1, 0.01, 1)
    with pytest.raises(RuntimeError):
        r.add_flowmap_tables(V, *t)
