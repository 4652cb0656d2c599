"""The oracle's step against closed forms, conservation laws and exact
distributions (checks that a dropped term, a wrong index or a swapped
operand in the restatement would fail)."""
import math

import numpy as np
import pytest
from scipy import stats

TRANSPORT, MODEL, CONTRIB, MOVE, LEAVE, REMOVE = 1, 2, 4, 8, 16, 32


def one_box(oracle_lib, model="monod_mass", V=1e-3, seed=1, dt=0.01):
    r = oracle_lib.OracleRun(model, 1, oracle_lib.model_info(model)[1], dt, seed)
    r.add_flowmap(np.array([V]), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    return r


def test_monod_euler_closed_form(oracle_lib):
    r = one_box(oracle_lib)
    mu_max, ks, Y = 0.5, 0.01, 0.5
    r.set_params([mu_max, ks, Y, 2.0, 1e-15, 0.1])
    s = 3.0
    r.set_concentrations(np.full((1, 1), s))
    n = 50
    m0 = np.linspace(1e-15, 1.2e-15, n)
    w = 7.0
    r.set_particles(2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), np.stack([m0, m0 / 1.0]),
                    np.zeros(n), np.zeros(n), w)
    r.set_params([mu_max, ks, Y, 1e9, 1e-15, 0.1])  # no division
    mu = mu_max * s / (ks + s)
    for k in range(1, 41):
        r.step_stages(MODEL | CONTRIB)
        st = r.particles()["state"]
        np.testing.assert_allclose(st[0], m0 * (1 + 0.01 * mu) ** k, rtol=1e-12)
        np.testing.assert_allclose(r.particles()["age_div"], np.full(n, 0.01 * k), rtol=1e-12)
    # the last step's sources: S = w * sum(-mu*m/Y) at the pre-growth mass
    m_prev = m0 * (1 + 0.01 * mu) ** 39
    np.testing.assert_allclose(r.contributions()[0, 0], -w * np.sum(mu * m_prev / Y), rtol=1e-12)


QUAD = [2.0, 0.01, 1e-4, 0.01, 0.5, 0.5, 0.3, 0.3, 0.4, 0.3, 0.2, 0.4, 1e9, 1.4e-15, 0.2e-15]  # no division


def _quad_box(oracle_lib, C, n=40, w=3.0):
    r = one_box(oracle_lib, "quad_monod")
    r.set_params(QUAD)
    r.set_concentrations(np.array(C, float).reshape(4, 1))
    m0 = np.linspace(1.0e-15, 1.8e-15, n)
    x = np.zeros((8, n))
    x[0], x[1], x[6] = m0, 0.5 * QUAD[0], m0
    r.set_particles(2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), x, np.zeros(n), np.zeros(n), w)
    return r, m0


def test_quad_monod_without_oxygen_overflows_everything(oracle_lib):
    """O = 0: nothing is respired, so no growth, no O2 or CO2 term, and the
    acetate excreted is Yas times the glucose taken up; the rest is stored."""
    r, m0 = _quad_box(oracle_lib, [5.0, 0.0, 0.0, 0.0])
    r.step_stages(MODEL | CONTRIB)
    S = r.contributions()[:, 0]
    x = r.particles()["state"]
    assert S[1] == 0.0 and S[3] == 0.0
    np.testing.assert_allclose(S[2], -QUAD[9] * S[0], rtol=1e-14)
    np.testing.assert_allclose(S[0], -3.0 * np.sum(0.5 * QUAD[0] * m0), rtol=1e-13)
    np.testing.assert_array_equal(x[0], m0)
    np.testing.assert_allclose(x[3], 0.01 * 0.5 * QUAD[0] * m0 * (1 - QUAD[9]), rtol=1e-13)
    np.testing.assert_allclose(x[5], 0.5 * QUAD[0] * m0, rtol=1e-14)


def test_quad_monod_uptake_lag_and_aerobic_growth(oracle_lib):
    """No glucose: the lagged uptake rate decays geometrically by (1 - dt/tau)
    per step; with oxygen and no acetate every cell grows by the same factor
    1 + dt*Yxs*qs*fo, and O2/CO2 terms stand in the ratio Yos/Yco2."""
    O = 2e-4
    r, m0 = _quad_box(oracle_lib, [0.0, O, 0.0, 0.0])
    fo = O / (QUAD[2] + O)
    q = 0.5 * QUAD[0]
    m = m0.copy()
    for k in range(1, 6):
        r.step_stages(MODEL | CONTRIB)
        x = r.particles()["state"]
        m = m * (1 + 0.01 * QUAD[5] * q * fo)
        q = q * (1 - 0.01 / QUAD[10])
        np.testing.assert_allclose(x[1], q, rtol=1e-13)
        np.testing.assert_allclose(x[0], m, rtol=1e-12)
    S = r.contributions()[:, 0]
    np.testing.assert_allclose(S[1] / S[3], -QUAD[7] / QUAD[11], rtol=1e-13)


@pytest.mark.parametrize("free,cap", [(50, None), (5000, 70), (5000, None)])
def test_division_conserves_mass_and_respects_budget(oracle_lib, free, cap):
    """Every cell of three tiles asks to divide: the first min(free slots,
    per-step cap) requests of the run in slot order are accepted (the
    reference refuses only when its division buffer is full,
    particles_container.hpp:559-573), the k-th daughter takes slot n + k, and
    mass is conserved."""
    r = one_box(oracle_lib)
    r.set_params([0.5, 0.01, 0.5, 2.0, 1e-15, 0.1])
    r.set_concentrations(np.full((1, 1), 1.0))
    n = 3000  # three tiles
    m = np.full(n, 2.5e-15)
    r.set_particles(n + free, np.zeros(n, np.uint32), np.zeros(n, np.uint8), np.stack([m, np.full(n, 1e-15)]),
                    np.zeros(n), np.full(n, 3.0), 1.0)
    if cap:
        r.set_division_cap(cap)
    r.step_stages(MODEL)
    c = r.counters()
    born = min(n, free, cap or n)
    assert c["births"] == born and c["overflow"] == n - born and r.n == n + born
    p = r.particles()
    grown = 2.5e-15 + 0.01 * (0.5 * 1.0 / 1.01) * 2.5e-15
    np.testing.assert_allclose(p["state"][0].sum(), n * grown, rtol=1e-13)
    div_slots = np.nonzero(p["state"][0][:n] < 2e-15)[0]
    assert list(div_slots) == list(range(born))  # the first requests in slot order
    assert np.all(p["age_div"][div_slots] == 0.0) and np.all(p["age_div"][n:] == 0.0)
    np.testing.assert_array_equal(p["state"][0][n:], p["state"][0][div_slots])
    assert np.all(p["age_div"][born:n] == 3.01)  # refused: the division age runs on


def test_closed_system_mass_balance(oracle_lib):
    """No feed, no outlet: liquid substrate + pending sources + biomass/Y is
    invariant, to 1e-10 relative."""
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=10, n_particles=4000, feeds=False)
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    Y = c.params[2]

    def total():
        return r.mass().sum() + c.dt * r.contributions().sum() + r.weight * r.particles()["state"][0].sum() / Y

    t0 = total()
    for _ in range(25):
        r.step(1)
        assert abs(total() - t0) <= 1e-10 * abs(t0)
    assert r.counters()["births"] > 100  # divisions happened


def test_cstr_exit_probability_and_residence_time(oracle_lib):
    """0-D reactor with an outlet: exit probability per step is
    1 - exp(-dt Q/V); residence times (in steps) are geometric."""
    V, Q, dt = 1e-3, 2e-3, 0.01
    r = one_box(oracle_lib, V=V, dt=dt)
    r.add_feed("constant", Q, 0.0, 0, 0, [0], [1.0])
    r.set_concentrations(np.full((1, 1), 1.0))
    r.init_particles(200000, 200000, 1.0)
    r.set_compaction(0.0, 0)
    p = 1 - math.exp(-dt * Q / V)
    alive = 200000
    exits = []
    for _ in range(60):
        r.step_stages(TRANSPORT | LEAVE | REMOVE)
        n_now = r.n
        exits.append(alive - n_now)
        alive = n_now
    total_exit = 200000 - alive
    expect = 200000 * (1 - (1 - p) ** 60)
    sd = math.sqrt(200000 * (1 - (1 - p) ** 60) * (1 - p) ** 60)
    assert abs(total_exit - expect) < 5 * sd
    ages = np.rint(r.exit_ages() / dt).astype(int)
    assert len(ages) == total_exit and ages.min() >= 1 and ages.max() <= 60
    # geometric(p) truncated at 60 steps: chi-square on the first 40 steps
    k = np.arange(1, 41)
    pk = (1 - p) ** (k - 1) * p
    obs = np.bincount(ages, minlength=61)[1:41]
    expct = 200000 * pk
    chi2 = np.sum((obs - expct) ** 2 / expct)
    assert stats.chi2.sf(chi2, len(k) - 1) > 1e-3


def two_box(oracle_lib, F=(3e-3, 1e-3), V=(1e-3, 2e-3), n=100000, dt=0.01, seed=2):
    r = oracle_lib.OracleRun("monod_mass", 2, 1, dt, seed)
    r.add_flowmap(np.array(V), np.array([0, 1], np.uint32), np.array([1, 0], np.uint32), np.array(F))
    r.set_concentrations(np.ones((1, 2)))
    r.init_particles(n, n, 1.0)
    return r


def test_move_probability_is_linear(oracle_lib):
    """Move test dt*Q/V > u (probability_leaving.hpp:33-46)."""
    r = two_box(oracle_lib)
    before = r.particles()["pos"]
    r.step_stages(MOVE)
    after = r.particles()["pos"]
    for i, (q, v) in enumerate(((3e-3, 1e-3), (1e-3, 2e-3))):
        n_i = int(np.sum(before == i))
        moved = int(np.sum((before == i) & (after != i)))
        p = 0.01 * q / v
        assert abs(moved - n_i * p) < 5 * math.sqrt(n_i * p * (1 - p))


def test_inactive_slots_keep_moving(oracle_lib):
    """move_kernel.hpp:209-273,393-434: the move stage runs over every slot
    below n whatever its status (Exit and Dead slots waiting for removal keep
    moving) and counts a Move event for each.  Half the slots Exit, removal
    never due: both halves move with probability dt*Q/V, and the counter
    holds every move.  A restatement that moved only Idle particles would
    leave the Exit half in place and count half the moves."""
    n = 200000
    r = oracle_lib.OracleRun("monod_mass", 2, 1, 0.01, 5)
    r.add_flowmap(np.array([1e-3, 2e-3]), np.array([0, 1], np.uint32), np.array([1, 0], np.uint32),
                  np.array([3e-3, 1e-3]))
    r.set_concentrations(np.ones((1, 2)))
    st = np.zeros(n, np.uint8)
    st[1::2] = 2  # Exit
    m = np.full(n, 1e-15)
    r.set_particles(n, np.zeros(n, np.uint32), st, np.stack([m, m]), np.zeros(n), np.zeros(n), 1.0)
    r.set_compaction(1e9, 10**9)
    r.step_stages(MOVE)
    p = r.particles()
    moved = p["pos"] != 0
    pm = 0.01 * 3e-3 / 1e-3
    for sel in (st == 0, st == 2):
        k, x = int(sel.sum()), int((moved & sel).sum())
        assert abs(x - k * pm) < 5 * math.sqrt(k * pm * (1 - pm)), (k, x)
    assert r.counters()["moves"] == int(moved.sum())
    np.testing.assert_array_equal(p["status"], st)


def test_neighbour_choice_follows_flows(oracle_lib):
    n = 4
    rows = np.array([0, 0, 0, 1, 2, 3], np.uint32)
    cols = np.array([1, 2, 3, 0, 0, 0], np.uint32)
    F = np.array([1e-2, 3e-2, 6e-2, 1e-2, 3e-2, 6e-2])
    r = oracle_lib.OracleRun("monod_mass", n, 1, 0.1, 9)
    r.add_flowmap(np.full(n, 1.0), rows, cols, F)
    r.set_concentrations(np.ones((1, n)))
    N = 120000
    r.set_particles(N, np.zeros(N, np.uint32), np.zeros(N, np.uint8), np.ones((2, N)) * 1e-15,
                    np.zeros(N), np.zeros(N), 1.0)
    r.step_stages(MOVE)
    pos = r.particles()["pos"]
    p_move = 0.1 * F[:3].sum() / 1.0
    counts = np.bincount(pos, minlength=n)
    expct = N * np.array([1 - p_move, p_move * 0.1, p_move * 0.3, p_move * 0.6])
    chi2 = np.sum((counts - expct) ** 2 / expct)
    assert stats.chi2.sf(chi2, 3) > 1e-3


def test_stationary_distribution_follows_volumes(oracle_lib):
    """A conservative closed network has the volume-proportional stationary
    distribution: each step's inflow dt*sum_i F(i,j) equals outflow dt*Q_j."""
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=12, n_particles=60000, feeds=False, n_maps=1)
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    for _ in range(300):  # mixing: the slowest compartment moves with p ~ 0.045 per step
        r.step_stages(MOVE)
    h = r.repartition().astype(float)
    V = c.maps[0][0]
    expct = h.sum() * V / V.sum()
    chi2 = np.sum((h - expct) ** 2 / expct)
    assert stats.chi2.sf(chi2, c.n_c - 1) > 1e-3


def test_removal_bookkeeping(oracle_lib):
    """apps/libs/mc/tests/test_container.cpp:108-148: removing k inactive
    particles leaves n-k live ones, all below n, with the same states."""
    for size, dead in ((1000, 10), (100, 99), (5000, 2500)):
        r = one_box(oracle_lib)
        rng = np.random.default_rng(size)
        st = np.zeros(size, np.uint8)
        idx = rng.choice(size, dead, replace=False)
        st[idx] = 3
        m = rng.uniform(1e-15, 2e-15, size)
        r.set_particles(size, np.zeros(size, np.uint32), st, np.stack([m, m]), np.arange(size) * 1.0,
                        np.zeros(size), 1.0)
        r.set_compaction(0.0, 0)
        r.step_stages(REMOVE)
        p = r.particles()
        assert r.n == size - dead and np.all(p["status"] == 0)
        np.testing.assert_array_equal(np.sort(p["state"][0]), np.sort(m[st == 0]))
        # gaps are filled from the tail: live particles below n - dead stay put
        keep = np.nonzero(st[: size - dead] == 0)[0]
        np.testing.assert_array_equal(p["age_hyd"][keep], keep * 1.0)


def test_removal_threshold(oracle_lib):
    """Removal only when inactive > max(min_removal, ratio * n)."""
    r = one_box(oracle_lib)
    n = 1000
    st = np.zeros(n, np.uint8)
    st[:10] = 2
    m = np.full(n, 1e-15)
    r.set_particles(n, np.zeros(n, np.uint32), st, np.stack([m, m]), np.zeros(n), np.zeros(n), 1.0)
    r.set_compaction(0.01, 0)  # threshold 10: 10 inactive is not > 10
    r.step_stages(REMOVE)
    assert r.n == 1000 and r.counters()["inactive"] == 10
    r.set_compaction(0.009, 0)  # threshold 9
    r.step_stages(REMOVE)
    assert r.n == 990 and r.counters()["inactive"] == 0 and r.counters()["compactions"] == 1


def test_init_distribution_and_weight(oracle_lib):
    r = one_box(oracle_lib, V=2e-3)
    n = 50000
    r.init_particles(n, n, 3.0)
    m = r.particles()["state"][0]
    a = (0 - 1e-15) / 1e-16
    assert stats.kstest(m, stats.truncnorm(a, np.inf, loc=1e-15, scale=1e-16).cdf).pvalue > 1e-3
    assert r.weight == (3.0 * 2e-3) / math.fsum(m) or abs(r.weight - 3.0 * 2e-3 / math.fsum(m)) < 1e-12 * r.weight
    mb = r.particles()["state"][1]
    ratio = m / mb - 1.0
    assert stats.kstest(ratio, "uniform").pvalue > 1e-3


def test_initial_positions_uniform(oracle_lib):
    r = oracle_lib.OracleRun("monod_mass", 37, 1, 0.01, 5)
    r.add_flowmap(np.full(37, 1e-3), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r.set_concentrations(np.ones((1, 37)))
    r.init_particles(74000, 74000, 1.0)
    h = np.bincount(r.particles()["pos"], minlength=37)
    assert stats.chisquare(h).pvalue > 1e-3


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod", "monod", "fixed_length", "simple_acetate",
                                   "quad_monod"])
def test_threads_do_not_change_events(oracle_lib, model):
    from mcst import cases

    c = cases.small_case(model, n_particles=6000)
    runs = []
    for th in (1, 4):
        r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        c.apply(r)
        r.step(12, threads=th)
        runs.append(r)
    a, b = (x.particles() for x in runs)
    assert runs[0].counters() == runs[1].counters()
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["status"], b["status"])
    np.testing.assert_allclose(a["state"], b["state"], rtol=1e-12)
    np.testing.assert_allclose(runs[0].concentrations(), runs[1].concentrations(), rtol=1e-12)
    # the probes are recorded in tile order whatever the thread count
    np.testing.assert_array_equal(runs[0].exit_ages(), runs[1].exit_ages())
    np.testing.assert_array_equal(runs[0].division_ages(), runs[1].division_ages())
    assert len(runs[0].division_ages()) > 0


def test_gas_liquid_transfer_closed_form(oracle_lib):
    """One well-mixed box, no particles, no flow: C_k = C* (1 - (1 - dt kla)^k)
    for the transferred species; a species with kla = 0 stays put; mass
    changes by exactly dt*kla*(C* - C)*V per step."""
    r = one_box(oracle_lib, "dual_monod", V=2e-3)
    r.set_concentrations(np.array([[1.5], [0.0]]))
    kla, cs = 0.3, 8e-3
    r.set_gas_transfer([0.0, kla], [0.0, cs])
    r.init_particles(0, 0, 1.0)
    for k in range(1, 31):
        m_prev = r.mass()[1, 0]
        c_prev = r.concentrations()[1, 0]
        r.step_stages(TRANSPORT)
        C = r.concentrations()
        np.testing.assert_allclose(C[1, 0], cs * (1 - (1 - 0.01 * kla) ** k), rtol=1e-12)
        np.testing.assert_allclose(r.mass()[1, 0] - m_prev, 0.01 * kla * (cs - c_prev) * 2e-3, rtol=1e-9)
        assert C[0, 0] == 1.5


def test_udf_minimal_closed_form(oracle_lib):
    """The example plugin's model (minimal.cpp:65-99) at constant substrate:
    the length grows by the same dt*ldot/(1 + dt*ldot) every step, so it is
    linear in the step count; each particle takes phi_s_max*g; at l_max the
    cell divides into two halves that both keep l_max."""
    p = [0.5, 2e-7, 2e-6, 1e-3, 0.6e-6]
    r = one_box(oracle_lib, "udf_minimal")
    r.set_params(p)
    s = 0.7
    r.set_concentrations(np.full((1, 1), s))
    n, w = 30, 2.0
    l0 = np.linspace(1.0e-6, 1.2e-6, n)
    r.set_particles(2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), np.stack([l0, np.full(n, 1.0)]),
                    np.zeros(n), np.zeros(n), w)  # l_max = 1 m: no division
    g = s / (p[3] + s)
    ldot = p[1] * g
    dl = 0.01 * ldot / (1 + 0.01 * ldot)
    for k in range(1, 21):
        r.step_stages(MODEL | CONTRIB)
        np.testing.assert_allclose(r.particles()["state"][0], l0 + k * dl, rtol=1e-12)
    lin = 1000 * math.pi * p[4] ** 2 / 4
    np.testing.assert_allclose(r.contributions()[0, 0], -w * n * (p[1] * lin / p[0]) * g, rtol=1e-12)
    # division: lengths just below l_max cross it in one step and halve
    r2 = one_box(oracle_lib, "udf_minimal")
    r2.set_params(p)
    r2.set_concentrations(np.full((1, 1), s))
    x0 = np.full(n, p[2] - 0.5 * dl)
    r2.set_particles(2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), np.stack([x0, np.full(n, p[2])]),
                     np.zeros(n), np.zeros(n), w)
    r2.step_stages(MODEL)
    st = r2.particles()["state"]
    assert r2.n == 2 * n
    np.testing.assert_allclose(st[0], np.full(2 * n, (x0[0] + dl) / 2), rtol=1e-14)
    np.testing.assert_array_equal(st[1], np.full(2 * n, p[2]))


def test_udf_starvation_death_times_and_contributions(oracle_lib):
    """The dying example plugin's model at constant substrate: viability falls
    by delta = dt*kd*(1 - g) per step, so a cell starting at (j + 1/2)*delta
    dies at step j + 1; dead cells stop contributing, and the living ones
    contribute w*mu*m/Y with m growing by (1 + dt*mu) per step."""
    p = [0.5, 0.01, 0.5, 2.0, 3.0, 1e-15, 0.1]
    s, dt, w = 1e-3, 0.01, 4.0
    g = s / (p[1] + s)
    mu, delta = p[0] * g, dt * (p[4] * (1 - g))
    n = 40
    j = np.arange(n) % 8
    v0 = (j + 0.5) * delta
    m0 = np.linspace(1e-15, 1.5e-15, n)
    r = one_box(oracle_lib, "udf_starvation", dt=dt)
    r.set_params(p)
    r.set_concentrations(np.full((1, 1), s))
    r.set_particles(2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), np.stack([m0, np.ones(n), v0]),
                    np.zeros(n), np.zeros(n), w)
    r.set_compaction(1.0, 0)  # keep dead slots in place
    for k in range(1, 10):
        r.step_stages(MODEL | CONTRIB)
        alive = j >= k
        assert r.counters()["deaths"] == int(np.sum(~alive))
        st = r.particles()["status"]
        np.testing.assert_array_equal(st == 3, ~alive)
        m_prev = m0 * (1 + dt * mu) ** (k - 1)
        want = -w * np.sum(mu * m_prev[alive] / p[2]) if alive.any() else 0.0
        np.testing.assert_allclose(r.contributions()[0, 0], want, rtol=1e-12)


def test_container_growth_doubles_an_all_dividing_population(oracle_lib):
    """Every cell divides every step (division ratio 0): with a capacity check before every step (grow to 2.5 n when
    fewer slots than cells are free) nothing is refused and n doubles each
    step; with the capacity fixed at the initial count nothing is born."""
    for every in (1, 0):
        r = one_box(oracle_lib)
        r.set_params([0.5, 0.01, 0.5, 0.0, 1e-15, 0.1])
        r.set_concentrations(np.full((1, 1), 1.0))
        r.init_particles(1000, 1000, 1.0)
        if every:
            r.set_growth(every, 2.5)
        for k in range(1, 7):
            r.step_stages(MODEL)
            c = r.counters()
            if every:
                assert r.n == 1000 * 2 ** k and c["overflow"] == 0
            else:
                assert r.n == 1000 and c["overflow"] == 1000 * k
    p = r.particles()
    assert p["state"].shape[1] == 1000


def test_time_dependent_feed_closed_form(oracle_lib):
    """One box, no particles, no outlet: a feed of flow q(t) at concentration
    c adds dt*q(k*dt)*c of mass at step k (t is the step's start), so
    after n steps the mass grew by c*dt*sum_k q(k*dt): for a linear feed
    c*p*dt^2*n(n-1)/2, for an exponential one c*dt*(n*f0 + (e^{a n dt}-1)/(e^{a dt}-1))."""
    dt, V, c_in = 0.01, 2.0, 4.0
    for kind, f0, a in (("linear", 0.0, 0.3), ("exponential", 0.5, -2.0)):
        r = one_box(oracle_lib, "monod_mass", V=V, dt=dt)
        r.set_concentrations(np.array([[0.25]]))
        r.add_feed(kind, f0, a, 0, None, [0], [c_in])
        r.init_particles(0, 0, 1.0)
        m0 = r.mass()[0, 0]
        r.step_stages(TRANSPORT)
        if kind == "linear":
            assert r.mass()[0, 0] == m0  # q(0) = 0: the first step adds nothing
        for n in range(2, 41):
            r.step_stages(TRANSPORT)
            if kind == "linear":
                want = c_in * a * dt * dt * n * (n - 1) / 2
            else:
                want = c_in * dt * (n * f0 + math.expm1(a * n * dt) / math.expm1(a * dt))
            np.testing.assert_allclose(r.mass()[0, 0] - m0, want, rtol=1e-11)
            np.testing.assert_allclose(r.concentrations()[0, 0], r.mass()[0, 0] / V, rtol=1e-14)


def test_reorder_is_a_stable_sort_by_compartment(oracle_lib):
    """sort_particles permutes whole particles into compartment order, slot
    order among equal positions; an interval k sorts after every k-th step."""
    from mcst import cases

    c = cases.small_case("dual_monod", n_c=9, n_particles=4000, seed=12)
    c.capacity = 40000
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    r.step(15)
    before = r.particles()
    n = r.n
    r.sort_particles()
    after = r.particles()
    order = np.argsort(before["pos"], kind="stable")
    np.testing.assert_array_equal(after["pos"], before["pos"][order])
    np.testing.assert_array_equal(after["status"], before["status"][order])
    np.testing.assert_array_equal(after["state"], before["state"][:, order])
    np.testing.assert_array_equal(after["age_hyd"], before["age_hyd"][order])
    np.testing.assert_array_equal(after["age_div"], before["age_div"][order])
    assert np.all(np.diff(after["pos"].astype(np.int64)) >= 0) and r.n == n
    # an interval equals explicit sorts at the same steps
    a = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    b = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    for x in (a, b):
        c.apply(x)
    a.set_sort_interval(4)
    a.step(12)
    for _ in range(3):
        b.step(4)
        b.sort_particles()
    pa, pb = a.particles(), b.particles()
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k])
    assert a.counters() == b.counters()


def test_ages_count_the_steps_that_age_them(oracle_lib):
    """The reference adds dt to an Idle particle's division age in every step
    that runs the model (model_kernel.hpp:193) and to its hydraulic age in
    every step that runs the leave stage of a run with a leaving flow
    (move_kernel.hpp:347-359), and resets them at division and exit
    (particles_container.hpp:559-573, move_kernel.hpp:633-636).  Without
    divisions and exits, every Idle particle's ages are then exactly k
    steps' worth; a step without the model (leave) stage leaves the division
    (hydraulic) age where it was; a run without an outlet keeps hydraulic
    age 0; and with divisions and exits every age and every recorded age is
    a whole number of steps."""
    from mcst import STAGE_ALL, STAGES
    from mcst import cases

    def ages_in_steps(p, dt):
        k = np.asarray(p) / dt
        assert np.all(np.abs(k - np.rint(k)) < 1e-9 * np.maximum(1.0, k))
        return np.rint(k)

    # slow growth, an outlet: no division within 40 steps, some exits
    c = cases.small_case("monod_mass", n_c=12, n_particles=4000, seed=7, fast=False)
    c.capacity = 20000
    r = oracle_lib.OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    r.set_compaction(1.0, 1 << 40)  # no removal: exited slots stay
    r.step(30)
    for _ in range(4):  # no model stage: the division age stands still
        r.step_stages(STAGE_ALL & ~STAGES["model"])
    for _ in range(6):  # no leave stage: the hydraulic age stands still
        r.step_stages(STAGE_ALL & ~STAGES["leave"])
    cnt = r.counters()
    assert cnt["births"] == 0 and cnt["exits"] > 0
    p = r.particles()
    idle = p["status"] == 0
    assert np.all(ages_in_steps(p["age_div"][idle], c.dt) == 36)
    assert np.all(ages_in_steps(p["age_hyd"][idle], c.dt) == 34)
    np.testing.assert_allclose(p["age_div"][idle], 36 * c.dt, rtol=1e-13)
    gone = p["status"] == 2  # exited: hydraulic age reset, division age frozen
    assert gone.any() and np.all(p["age_hyd"][gone] == 0.0)
    assert np.all(ages_in_steps(p["age_div"][gone], c.dt) <= 36)

    # no outlet: the hydraulic age stays 0
    c2 = cases.small_case("monod_mass", n_c=12, n_particles=2000, seed=8, fast=False, feeds=False)
    r2 = oracle_lib.OracleRun(c2.model, c2.n_c, c2.n_s, c2.dt, c2.seed)
    c2.apply(r2)
    r2.step(12)
    p2 = r2.particles()
    assert np.all(p2["age_hyd"] == 0.0)
    assert np.all(ages_in_steps(p2["age_div"], c2.dt) == 12)

    # divisions and exits: whole numbers of steps everywhere
    c3 = cases.small_case("monod_mass", n_c=12, n_particles=4000, seed=9)
    c3.capacity = 60000
    r3 = oracle_lib.OracleRun(c3.model, c3.n_c, c3.n_s, c3.dt, c3.seed)
    c3.apply(r3)
    r3.set_probes(1 << 20)
    r3.step(25)
    cnt3 = r3.counters()
    assert cnt3["births"] > 0 and cnt3["exits"] > 0
    p3 = r3.particles()
    for f in ("age_div", "age_hyd"):
        k = ages_in_steps(p3[f], c3.dt)
        assert np.all((k >= 0) & (k <= 25))
    for rec in (r3.exit_ages(), r3.division_ages()):
        k = ages_in_steps(rec, c3.dt)
        assert len(k) and np.all((k >= 1) & (k <= 25))

