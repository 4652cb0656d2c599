"""The oracle's restatements of the particle models, pinned by what the
reference's equations imply rather than by the equations themselves:
closed forms of the explicit Euler recursions at constant concentration
(geometric lags, linear growth), half-saturation points where the Monod
factor is exactly 1/2, balances between outputs (uptake against growth and
storage), division invariants, and the distributions of the random draws
(KS against scipy).  A dropped term, a swapped operand or a wrong constant
in model_update / model_division / model_init fails one of these.

Reference models: apps/libs/models/public/models/monod.hpp:80-193,
fixed_length.hpp:108-161 with models/src/config_loader.cpp:13-60,
simple_acetate.hpp:132-248; dual_monod is the 1e7-particle workload's own
model (4 FP64 states: mass, uptake rate, storage pool, division age)."""
import math

import numpy as np
import pytest
from scipy import stats

MODEL, CONTRIB = 2, 4
LN2 = math.log(2.0)


def lin_density(d):
    return 1000.0 * math.pi * d * d / 4.0


def box(oracle_lib, model, C, state, w=1.0, dt=0.01, params=None, cap=None):
    """One compartment, concentrations held fixed (no transport stage)."""
    nv, ns, _ = oracle_lib.model_info(model)
    r = oracle_lib.OracleRun(model, 1, ns, dt, 5)
    r.add_flowmap(np.array([1e-3]), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    if params is not None:
        r.set_params(params)
    r.set_concentrations(np.asarray(C, float).reshape(ns, 1))
    state = np.asarray(state, float)
    n = state.shape[1]
    r.set_particles(cap or 2 * n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), state, np.zeros(n),
                    np.zeros(n), w)
    return r


def step(r):
    r.step_stages(MODEL | CONTRIB)
    return r.particles()["state"], r.contributions()[:, 0]


# ------------------------------------------------------------------ monod
MONOD = [0.77 / 3600.0, 1e-3, 2.0, 2e-6, 0.6e-6, 1.0]  # mu_max, k_s, y_s_x, l_max, d_m, tau_meta (short lag)


def monod_state(n, l0, mu_p0, l_max=2e-6):
    x = np.zeros((6, n))
    x[0] = l0
    x[1] = l_max
    x[2] = mu_p0
    x[4] = (MONOD[3] / 2.0) / LN2  # dl / ln2 (monod.hpp:103-110)
    return x


def test_monod_lagging_growth_rate_closed_form(oracle_lib):
    """s = k_s, so mu = mu_max/2; mu_p starts at 0 < mu, so mu_eff = mu_p:
    mu_p_k = mu (1 - a^k) with a = 1 - dt/tau, and the length follows the
    geometric sum l_k = l_0 + dt*(dl/ln2)*mu*(k - (1 - a^k)/(1 - a)).
    The uptake is -mu_eff*y*mass at the new length (monod.hpp:122-160)."""
    n, dt, w = 16, 0.01, 3.0
    l0 = np.linspace(1.0e-6, 1.4e-6, n)
    r = box(oracle_lib, "monod", [MONOD[1]], monod_state(n, l0, 0.0), w=w, dt=dt, params=MONOD)
    mu = MONOD[0] / 2.0
    a = 1.0 - dt / MONOD[5]
    dl = (MONOD[3] / 2.0) / LN2
    ld = lin_density(MONOD[4])
    for k in range(1, 41):
        x, S = step(r)
        mu_eff = mu * (1.0 - a ** (k - 1))  # the rate of step k: mu_p before its update
        np.testing.assert_allclose(x[2], mu * (1.0 - a ** k), rtol=1e-12)
        np.testing.assert_allclose(x[0], l0 + dt * dl * mu * (k - (1.0 - a ** k) / (1.0 - a)), rtol=1e-12)
        np.testing.assert_allclose(x[3], mu_eff, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(x[5], -mu_eff * MONOD[2] * x[0] * ld, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(S[0], w * x[5].sum(), rtol=1e-12, atol=1e-300)


def test_monod_rate_capped_by_substrate(oracle_lib):
    """mu_p above the substrate's mu: mu_eff = mu, so the length grows
    linearly while mu_p relaxes to mu geometrically."""
    n, dt = 8, 0.01
    mu = MONOD[0] * 4.0 / (MONOD[1] + 4.0)
    mu_p0 = 2.0 * MONOD[0]
    l0 = np.full(n, 1.2e-6)
    r = box(oracle_lib, "monod", [4.0], monod_state(n, l0, mu_p0), dt=dt, params=MONOD)
    dl = (MONOD[3] / 2.0) / LN2
    a = 1.0 - dt / MONOD[5]
    for k in range(1, 31):
        x, _ = step(r)
        np.testing.assert_allclose(x[0], l0 + k * dt * dl * mu, rtol=1e-12)
        np.testing.assert_allclose(x[2], mu + (mu_p0 - mu) * a ** k, rtol=1e-12)
        np.testing.assert_allclose(x[3], mu, rtol=1e-14)


def test_monod_clamps_negative_concentration(oracle_lib):
    """A heavy population drives the box's concentration below zero in one
    transport step; the model then reads it as 0 (monod.hpp:130-131): no
    growth, no uptake, and mu_p decays as mu_p * a per step."""
    n, dt = 4, 0.01
    r = box(oracle_lib, "monod", [1e-9], monod_state(n, np.full(n, 1.5e-6), MONOD[0]), w=1e15, dt=dt,
            params=MONOD)
    r.step_stages(1 | MODEL | CONTRIB)  # transport with no sources yet, then uptake
    r.step_stages(1 | MODEL | CONTRIB)  # transport takes the uptake: the concentration goes negative
    assert r.concentrations()[0, 0] < 0.0
    x0 = r.particles()["state"]
    assert r.contributions()[0, 0] == 0.0 and np.all(x0[3] == 0.0)
    a = 1.0 - dt / MONOD[5]
    for k in range(1, 4):
        x, S = step(r)
        assert np.all(x[0] == x0[0]) and S[0] == 0.0
        np.testing.assert_allclose(x[2], x0[2] * a ** k, rtol=1e-13)


def test_monod_division(oracle_lib):
    """A cell past its own l_max halves; the daughter takes the same length,
    the model's l_max, the mother's mu_p and dl/ln2, zero mu_eff and uptake;
    the mother keeps her l_max (monod.hpp:162-193)."""
    n, dt = 10, 0.01
    x0 = monod_state(n, np.full(n, 1.6e-6), MONOD[0], l_max=1.5e-6)
    r = box(oracle_lib, "monod", [1.0], x0, dt=dt, params=MONOD)
    x, _ = step(r)
    assert r.n == 2 * n and r.counters()["births"] == n
    mu = MONOD[0] * 1.0 / (MONOD[1] + 1.0)
    l_new = 1.6e-6 + dt * x0[4, 0] * min(MONOD[0], mu)
    mother, daughter = x[:, :n], x[:, n:]
    np.testing.assert_allclose(mother[0], l_new / 2.0, rtol=1e-15)
    np.testing.assert_array_equal(daughter[0], mother[0])
    assert np.all(mother[1] == 1.5e-6) and np.all(daughter[1] == MONOD[3])
    np.testing.assert_array_equal(daughter[2], mother[2])
    np.testing.assert_array_equal(daughter[4], mother[4])
    assert np.all(daughter[3] == 0.0) and np.all(daughter[5] == 0.0)
    assert np.all(r.particles()["age_div"] == 0.0)


# ------------------------------------------------------------------ fixed_length
FIXED = [2e-6 / 3600.0, 2e-6, 1e-3, 0.6e-6]  # l_dot_max, l_max, k, d_m


@pytest.mark.parametrize("s,g", [(1e-3, 0.5), (0.0, 0.0), (3e-3, 0.75)])
def test_fixed_length_linear_growth_and_yield(oracle_lib, s, g):
    """g = s/(k + s) (1/2 at s = k): the length grows by dt*l_dot_max*g per
    step, and the uptake is the mass growth over the yield 0.5:
    phi = -(dl/dt)*lin_density/0.5 (fixed_length.hpp:123-140, utils.hpp:46-52)."""
    n, dt, w = 12, 0.01, 2.0
    l0 = np.linspace(1.0e-6, 1.5e-6, n)
    x0 = np.stack([l0, np.full(n, 2e-6)])
    r = box(oracle_lib, "fixed_length", [s], x0, w=w, dt=dt, params=FIXED)
    ld = lin_density(FIXED[3])
    for k in range(1, 21):
        x, S = step(r)
        np.testing.assert_allclose(x[0], l0 + k * dt * FIXED[0] * g, rtol=1e-12)
    np.testing.assert_allclose(S[0], -w * n * (FIXED[0] * g) * ld / 0.5, rtol=1e-12, atol=1e-300)


def test_fixed_length_division(oracle_lib):
    n = 6
    x0 = np.stack([np.full(n, 2.1e-6), np.full(n, 2e-6)])
    r = box(oracle_lib, "fixed_length", [1e-3], x0, params=FIXED)
    x, _ = step(r)
    assert r.n == 2 * n
    l_new = 2.1e-6 + 0.01 * FIXED[0] * 0.5
    np.testing.assert_allclose(x[0], l_new / 2.0, rtol=1e-15)
    assert np.all(x[1] == FIXED[1])


@pytest.mark.parametrize("lam", [LN2 / 1e-6, 3.0 / 1e-6, 0.0])
def test_fixed_length_initial_length_distribution(oracle_lib, lam):
    """Initial lengths follow the Metropolis target of config_loader.cpp:13-36,
    lambda*exp(-lambda*l) on [l_max/2, l_max] (KS against scipy's truncated
    exponential; lambda = 0 is the uniform limit)."""
    r = oracle_lib.OracleRun("fixed_length", 4, 1, 0.01, 17)
    r.add_flowmap(np.full(4, 1e-3), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r.set_params(FIXED + [lam])
    r.init_particles(20000, 20000, 1.0)
    x = r.particles()["state"]
    a, b = 1e-6, 2e-6
    assert np.all((x[0] >= a) & (x[0] <= b)) and np.all(x[1] == b)
    if lam:
        ref = stats.truncexpon(b=lam * (b - a), loc=a, scale=1.0 / lam)
    else:
        ref = stats.uniform(loc=a, scale=b - a)
    assert stats.kstest(x[0], ref.cdf).pvalue > 1e-3
    # the default: half the cells start below a + ln(4/3)/lambda (the median)
    if lam == LN2 / 1e-6:
        np.testing.assert_allclose(np.median(x[0]), a + math.log(4.0 / 3.0) / lam, rtol=5e-3)


# ------------------------------------------------------------------ simple_acetate
A_MAX, L_MAX, K_S, K_A, Y_S, Y_A = 2e-6 / 3600.0, 2e-6, 1e-3, 1e-4, 2.0, 3.0
LD_SA = lin_density(0.6e-6)


def sa_state(n, a_p, length=1.2e-6, l_max=1.9e-6):
    x = np.zeros((9, n))
    x[0] = length
    x[1] = l_max
    x[2] = a_p
    x[3] = A_MAX
    return x


# (c0, c1, a_p) -> (glucose uptake U0, acetate uptake U1, glucose uptake capacity D0, overflow pa if >= 0)
SA_CASES = {
    # glucose at half saturation (D0 = a_max/2) above the cell's a_p: it takes a_p of
    # glucose and excretes the rest as acetate; acetate in the liquid is ignored
    "overflow": ((K_S, 7.0, A_MAX / 4), (A_MAX / 4, 0.0, A_MAX / 2, A_MAX / 4)),
    # glucose below a_p: acetate at half saturation (D1 = a_max/6) is taken up too
    "co_uptake": ((K_S, K_A, A_MAX), (A_MAX / 2, A_MAX / 6, A_MAX / 2, None)),
    # plenty of acetate: its uptake fills exactly the gap a_p - D0
    "gap_filled": ((K_S, 1e3, 0.6 * A_MAX), (A_MAX / 2, 0.1 * A_MAX, A_MAX / 2, None)),
    # nothing in the liquid: no uptake at all
    "starved": ((0.0, 0.0, A_MAX / 2), (0.0, 0.0, 0.0, None)),
}


@pytest.mark.parametrize("case", sorted(SA_CASES))
def test_simple_acetate_uptake_branches(oracle_lib, case):
    """simple_acetate.hpp:154-206 at fixed concentrations: the length grows
    by dt*(U0 + U1) per step; phi_s = -D0*ld*Y_s; phi_a = -U1*ld*Y_a when
    glucose is short, +(D0 - a_p)*ld*Y_s/Y_a (excretion) when it overflows."""
    (c0, c1, a_p), (u0, u1, d0, pa) = SA_CASES[case]
    n, dt, w = 9, 0.01, 4.0
    x0 = sa_state(n, a_p)
    r = box(oracle_lib, "simple_acetate", [c0, c1], x0, w=w, dt=dt)
    for k in range(1, 11):
        x, S = step(r)
        np.testing.assert_allclose(x[0], x0[0] + k * dt * (u0 + u1), rtol=1e-12)
    np.testing.assert_allclose(x[4], u0 + u1, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(x[5], u0, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(x[6], u1, rtol=1e-12, atol=1e-300)
    phi_s = -d0 * LD_SA * Y_S
    phi_a = pa * LD_SA * Y_S / Y_A if pa is not None else -u1 * LD_SA * Y_A
    np.testing.assert_allclose(x[7], phi_s, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(x[8], phi_a, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(S, [w * n * phi_s, w * n * phi_a], rtol=1e-12, atol=1e-300)


def test_simple_acetate_division_draws(oracle_lib):
    """simple_acetate.hpp:208-248: the length is halved and copied to the
    daughter with a_max; each cell redraws a_p ~ LogNormal(ln(a_e) - s^2/2, s)
    with s = 0.2 (mean a_e) and l_max ~ TruncatedNormal(l_max, l_max/10,
    0.7 l_max, 1.3 l_max).  KS tests against scipy."""
    n, dt = 6000, 0.01
    a_p = A_MAX / 4
    x0 = sa_state(n, a_p, length=1.95e-6, l_max=1.9e-6)
    r = box(oracle_lib, "simple_acetate", [K_S, 0.0], x0, dt=dt)
    x, _ = step(r)
    assert r.n == 2 * n and r.counters()["overflow"] == 0
    mother, daughter = x[:, :n], x[:, n:]
    l_new = 1.95e-6 + dt * a_p
    np.testing.assert_allclose(mother[0], l_new / 2.0, rtol=1e-15)
    np.testing.assert_array_equal(daughter[0], mother[0])
    np.testing.assert_array_equal(daughter[3], mother[3])
    assert np.all(daughter[4:] == 0.0)
    a_e = a_p  # the overflow branch: a_e = a_p
    ap = np.concatenate([mother[2], daughter[2]])
    sig = 0.2
    assert stats.kstest(np.log(ap), stats.norm(math.log(a_e) - sig * sig / 2, sig).cdf).pvalue > 1e-3
    assert abs(ap.mean() / a_e - 1.0) < 4 * sig / math.sqrt(ap.size)
    lm = np.concatenate([mother[1], daughter[1]])
    ref = stats.truncnorm(-3.0, 3.0, loc=L_MAX, scale=L_MAX / 10)
    assert stats.kstest(lm, ref.cdf).pvalue > 1e-3
    assert lm.min() >= 0.7 * L_MAX and lm.max() <= 1.3 * L_MAX
    # mother and daughter draw independently
    assert abs(np.corrcoef(mother[2], daughter[2])[0, 1]) < 0.06


# ------------------------------------------------------------------ dual_monod
DUAL = [2.8e-4, 0.01, 1e-4, 0.5, 1.0, 100.0, 1e-4, 1e9, 1.4e-15, 0.2e-15]  # no division


def dual_state(n):
    x = np.zeros((3, n))
    x[0] = np.linspace(1.2e-15, 1.8e-15, n)
    x[1] = 0.0
    x[2] = np.linspace(0.0, 1e-16, n)
    return x


def test_dual_monod_uptake_lag(oracle_lib):
    """Glucose at half saturation: the target rate is qs_max/2 and the lagged
    uptake rate follows q_k = q_t (1 - a^k), a = 1 - dt/tau."""
    n, dt = 8, 0.01
    r = box(oracle_lib, "dual_monod", [DUAL[1], 1.0], dual_state(n), dt=dt, params=DUAL)
    qt = DUAL[0] / 2.0
    a = 1.0 - dt / DUAL[5]
    for k in range(1, 31):
        x, _ = step(r)
        np.testing.assert_allclose(x[1], qt * (1.0 - a ** k), rtol=1e-12)


@pytest.mark.parametrize("O,fo", [(0.0, 0.0), (1e-4, 0.5), (3e-4, 0.75)])
def test_dual_monod_oxygen_split_and_balances(oracle_lib, O, fo):
    """The oxygen saturation fo = O/(Ko + O) (1/2 at O = Ko) splits the
    glucose taken up, U = q*m, between respiration (growth Yxs*U*fo) and the
    storage pool (U*(1 - fo)), which is respired at k_st*X*fo.  Balances that
    hold for any such split: uptake = (growth/Yxs + storage change)/dt per
    cell, and the O2 source is -Yos times the biomass growth rate."""
    n, dt, w = 10, 0.01, 5.0
    x0 = dual_state(n)
    x0[1] = DUAL[0] / 3.0  # a running uptake rate
    r = box(oracle_lib, "dual_monod", [5.0, O], x0, w=w, dt=dt, params=DUAL)
    prev = x0.copy()
    for _ in range(5):
        x, S = step(r)
        U = prev[1] * prev[0]
        dm, dX = x[0] - prev[0], x[2] - prev[2]
        np.testing.assert_allclose((dm / DUAL[3] + dX) / dt, U, rtol=1e-9)
        np.testing.assert_allclose(S[0], -w * U.sum(), rtol=1e-12)
        np.testing.assert_allclose(S[1], -DUAL[4] * w * dm.sum() / dt, rtol=1e-9, atol=1e-300)
        # storage: filled by the unrespired uptake, emptied by respiration
        np.testing.assert_allclose(dX / dt, U * (1 - fo) - DUAL[6] * prev[2] * fo, rtol=1e-9, atol=1e-40)
        prev = x.copy()
    if O == 0.0:  # no oxygen: no growth, no O2 source
        assert np.all(x[0] == x0[0]) and S[1] == 0.0


def test_dual_monod_division_halves_mass_and_storage(oracle_lib):
    p = list(DUAL)
    p[7] = 1.5e-15
    n = 8
    x0 = dual_state(n)
    x0[0] = 1.6e-15
    x0[1] = 1e-4
    r = box(oracle_lib, "dual_monod", [5.0, 1.0], x0, params=p)
    before = x0.copy()
    x, _ = step(r)
    assert r.n == 2 * n
    mother, daughter = x[:, :n], x[:, n:]
    np.testing.assert_array_equal(daughter, mother)
    fo = 1.0 / (DUAL[2] + 1.0)
    U = before[1] * before[0]
    m_new = before[0] + 0.01 * (DUAL[3] * U * fo + DUAL[3] * DUAL[6] * before[2] * fo)
    np.testing.assert_allclose(mother[0] + daughter[0], m_new, rtol=1e-12)
