"""Known-answer and library cross-checks of the oracle's building blocks."""
import json
import os

import numpy as np
import pytest
from scipy import sparse, stats

from philox_ref import philox4

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# Random123 kat_vectors, philox4x32 10 rounds (ctr, key, expected)
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(oracle_lib, ctr, key, want):
    assert tuple(int(x) for x in oracle_lib.philox(ctr, key)) == want
    assert tuple(philox4(ctr, key, 32)) == want


def test_python_philox_matches_numpy_4x64():
    # the same round code, run with 64-bit words, reproduces numpy's Philox
    key = (0x0123456789ABCDEF, 0xFEDCBA9876543210)
    for start in (1, 7, 12345):
        bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.array([start - 1, 0, 0, 0], np.uint64))
        got = [int(x) for x in bg.random_raw(4)]
        assert got == philox4((start, 0, 0, 0), key, 64)


def test_oracle_philox_random_counters(oracle_lib):
    rng = np.random.default_rng(5)
    for _ in range(200):
        ctr = rng.integers(0, 2**32, 4, dtype=np.uint64)
        key = rng.integers(0, 2**32, 2, dtype=np.uint64)
        assert [int(x) for x in oracle_lib.philox(ctr, key)] == philox4(ctr, key, 32)


def test_uniform_maps(oracle_lib):
    L = oracle_lib.lib()
    assert L.orc_u01(0) == 2.0 ** -33
    assert L.orc_u01(0xFFFFFFFF) == 1.0 - 2.0 ** -33
    assert 0.0 < L.orc_u01_52(0, 0) < 1e-15
    assert L.orc_u01_52(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    xs = np.array([L.orc_u01(int(x)) for x in np.random.default_rng(1).integers(0, 2**32, 20000)])
    assert stats.kstest(xs, "uniform").pvalue > 1e-3


def test_norm_inv_against_scipy(oracle_lib):
    L = oracle_lib.lib()
    p = np.concatenate([np.geomspace(1e-300, 1e-3, 200), np.linspace(1e-3, 1 - 1e-3, 4001),
                        1 - np.geomspace(1e-3, 1e-15, 100)])
    mine = np.array([L.orc_norm_inv(x) for x in p])
    ref = stats.norm.ppf(p)
    np.testing.assert_allclose(mine, ref, rtol=5e-15, atol=1e-300)
    x = np.linspace(-30, 8, 2001)
    # relative error of erfc(-x/sqrt2) grows like x^2 * eps in the far tail
    np.testing.assert_allclose([L.orc_norm_cdf(v) for v in x], stats.norm.cdf(x), rtol=1e-12, atol=0)


def test_transition_matrix_kats(oracle_lib):
    kat = json.load(open(os.path.join(GOLD, "reference_kats.json")))
    for name in ("transition_2x2", "transition_3x3"):
        F = np.array(kat[name]["flows"], float)
        n = F.shape[0]
        r, c = np.nonzero(np.ones_like(F))
        got = oracle_lib.transition_dense(n, r, c, F[r, c])
        np.testing.assert_array_equal(got, np.array(kat[name]["expected"], float))


def test_transport16_move_tables(oracle_lib):
    """CDF rows of the reference's 16-compartment fixture: monotone, in [0,1],
    last column 1 (or 0 without flow), and P(j|leave) = F(i,j)/sum_k F(i,k)."""
    kat = json.load(open(os.path.join(GOLD, "reference_kats.json")))
    F = np.array(kat["transport16_flows"]["flows"])
    np.fill_diagonal(F, 0.0)
    r, c = np.nonzero(F)
    V = np.linspace(1.0, 2.5, 16)
    dt = 1e-3
    out, pm, cdf, nbr = oracle_lib.move_tables(V, r, c, F[r, c], dt)
    np.testing.assert_allclose(out, F.sum(axis=1), rtol=1e-15)
    np.testing.assert_array_equal(pm, (dt * out) / V)
    assert np.all(np.diff(cdf, axis=1) >= 0) and np.all((cdf >= 0) & (cdf <= 1))
    for i in range(16):
        assert cdf[i, -1] == (1.0 if out[i] > 0 else 0.0)
        p = np.diff(np.concatenate([[0.0], cdf[i]]))
        probs = np.zeros(16)
        np.add.at(probs, nbr[i], p)
        if out[i] > 0:
            np.testing.assert_allclose(probs, F[i] / F[i].sum(), rtol=1e-13, atol=1e-15)


def test_transport_against_scipy(oracle_lib):
    g = np.load(os.path.join(GOLD, "transport_scipy.npz"))
    for k in range(int(g["n_cases"])):
        a = {n: g[f"case{k}_{n}"] for n in ("rows", "cols", "flows", "V", "C", "m", "S", "sink", "m2", "C2", "dt")}
        C2, m2 = oracle_lib.transport_once(a["V"], a["rows"], a["cols"], a["flows"], a["sink"], a["S"],
                                           float(a["dt"]), a["C"], a["m"])
        np.testing.assert_allclose(m2, a["m2"], rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(C2, a["C2"], rtol=1e-12, atol=1e-15)


def test_transport_against_live_scipy(oracle_lib):
    rng = np.random.default_rng(99)
    n, ns = 50, 2
    nnz = 300
    rows = rng.integers(0, n, nnz).astype(np.uint32)
    cols = rng.integers(0, n, nnz).astype(np.uint32)
    flows = rng.uniform(0, 1e-3, nnz)
    V, C = rng.uniform(1e-3, 2e-3, n), rng.uniform(0, 5, (ns, n))
    S, sink = rng.normal(0, 1e-4, (ns, n)), rng.uniform(0, 1e-4, n)
    keep = rows != cols
    Fm = sparse.coo_matrix((flows[keep], (rows[keep], cols[keep])), shape=(n, n)).tocsr()
    M = Fm - sparse.diags(np.asarray(Fm.sum(axis=1)).ravel())
    m2 = C * V + 0.01 * ((M.T @ C.T).T - C * sink + S)
    C2o, m2o = oracle_lib.transport_once(V, rows, cols, flows, sink, S, 0.01, C, C * V)
    np.testing.assert_allclose(m2o, m2, rtol=1e-12)
    np.testing.assert_allclose(C2o, m2 / V, rtol=1e-12)


def test_feed_closed_forms(oracle_lib):
    kat = json.load(open(os.path.join(GOLD, "reference_kats.json")))["feed"]
    L = oracle_lib.lib()
    for t, lin, ex in zip(kat["t"], kat["linear"], kat["exponential"]):
        assert abs(L.orc_feed_flow(0, kat["f0"], 0.0, t) - kat["f0"]) < kat["tolerance"]
        assert abs(L.orc_feed_flow(1, kat["f0"], kat["df"], t) - lin) < kat["tolerance"]
        assert abs(L.orc_feed_flow(2, kat["f0"], kat["alpha"], t) - ex) < kat["tolerance"]


def test_feed_through_a_run(oracle_lib):
    """One compartment without outlet: m_{k+1} - m_k = dt * f(t_k) * c_feed."""
    for kind, param in (("linear", 0.1), ("exponential", 0.1), ("constant", 0.0)):
        r = oracle_lib.OracleRun("monod_mass", 1, 1, 0.1, 1)
        r.add_flowmap(np.array([1.0]), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
        r.add_feed(kind, 1.0, param, 0, None, [0], [2.0])
        r.set_concentrations(np.zeros((1, 1)))
        r.init_particles(0, 0, 1.0)
        prev = r.mass()[0, 0]
        for k in range(12):
            r.step(1)
            m = r.mass()[0, 0]
            f = oracle_lib.lib().orc_feed_flow(oracle_lib.FEED_TYPES[kind], 1.0, param, k * 0.1)
            assert abs((m - prev) / 0.1 - 2.0 * f) < 1e-11 * max(1.0, abs(2.0 * f))
            prev = m


def test_case_0d_fixture(oracle_lib):
    """The reference's 0-D case (one 0.02 m3 compartment, no flow): nothing
    moves or leaves and the liquid stays put without particles' uptake."""
    kat = json.load(open(os.path.join(GOLD, "reference_kats.json")))["case_0d"]
    V = np.array(kat["volumes"])
    F = np.array(kat["flow_matrix"])
    r, c = np.nonzero(F)
    run = oracle_lib.OracleRun("monod_mass", 1, 1, 0.01, 4)
    run.add_flowmap(V, r, c, F[r, c])
    run.set_concentrations(np.full((1, 1), 5.0))
    run.init_particles(2000, 4000, 1.0)
    run.step_stages(1 | 8 | 16 | 32)  # transport, move, leave, remove (no model)
    cnt = run.counters()
    assert cnt["moves"] == 0 and cnt["exits"] == 0 and run.n == 2000
    assert run.mass()[0, 0] == 5.0 * 0.02
    assert np.all(run.particles()["pos"] == 0)
