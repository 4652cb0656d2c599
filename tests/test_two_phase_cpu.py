"""The oracle's two-phase step (SimulationUnit::ode_step with a gas phase,
apps/libs/simulation/src/simulation.model.cpp:131-150) against what the
mathematics fixes: the explicit scheme of the transfer between a liquid and
a gas box in closed form (the Henry-weighted gap decays geometrically, the
total mass is constant), the turbulence kLa correlation's scaling laws and
value (impl_mtr.cpp:21-50,90-134), gas transport and gas feeds conserving
the gas mass up to what the feeds bring and the gas outlet takes, and
clearNegs clipping only small negative liquid concentrations
(implScalar.cpp:266-297)."""
import math

import numpy as np
import pytest

T, M, C_ = 1, 2, 4
H_O2 = 3.181e-2  # the reference's Henry constant of oxygen (mass_transfer.cpp:125)


def gl_box(oracle_lib, Vl=1e-3, Vg=2e-4, Cl=(5.0, 0.0), Cg=(0.0, 0.3), dt=0.01, model="dual_monod"):
    r = oracle_lib.OracleRun(model, 1, 2, dt, 1)
    none = (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r.add_flowmap(np.array([Vl]), *none)
    r.add_gas_flowmap(np.array([Vg]), *none)
    r.set_concentrations(np.array(Cl, float).reshape(2, 1))
    r.set_gas_phase([0.0, H_O2], np.array(Cg, float).reshape(2, 1))
    r.set_particles(0, np.zeros(0, np.uint32), np.zeros(0, np.uint8), np.zeros((3, 0)), np.zeros(0), np.zeros(0), 1.0)
    return r


def test_transfer_closed_form_and_mass_conservation(oracle_lib):
    """d = H*Cg - Cl obeys d_{n+1} = d_n (1 - dt kla (1 + H Vl/Vg)); the
    total Cl*Vl + Cg*Vg is conserved; a species with zero kLa is untouched."""
    Vl, Vg, dt, kla = 1e-3, 2e-4, 0.01, 20.0
    r = gl_box(oracle_lib, Vl, Vg, dt=dt)
    r.set_kla("fixed", [0.0, kla])
    cl0, cg0 = 0.0, 0.3
    d0 = H_O2 * cg0 - cl0
    tot = cl0 * Vl + cg0 * Vg
    q = 1.0 - dt * kla * (1.0 + H_O2 * Vl / Vg)
    for n in range(1, 60):
        r.step_stages(T)
        cl, cg = r.concentrations()[:, 0], r.gas_concentrations()[:, 0]
        assert cl[0] == 5.0 and cg[0] == 0.0
        d = d0 * q ** n
        cl_exp = (tot - d * Vg / H_O2) / (Vl + Vg / H_O2)
        np.testing.assert_allclose(cl[1], cl_exp, rtol=1e-11)
        np.testing.assert_allclose(H_O2 * cg[1] - cl[1], d, rtol=1e-10, atol=1e-15)
        ml, mg = r.mass()[1, 0], r.gas_mass()[1, 0]
        np.testing.assert_allclose(ml + mg, tot, rtol=1e-14)
    assert abs(H_O2 * cg[1] - cl[1]) < 1e-3 * d0  # near Henry equilibrium


def test_turbulence_kla_correlation(oracle_lib):
    """kLa of the oxygen row from eps: kL = 0.3 (eps nu)^(1/4) Sc^(-1/2)
    (nu = 1.0023e-6 m2/s, the reference correlation's water at 20 C; D =
    1e-9 m2/s), a = 6 alpha/(d_b (1 - alpha)), alpha = Vg/(Vl + Vg); read
    back from one step's transfer dm_l = dt kla (H Cg - Cl) Vl."""
    def kla_of(eps, Vl, Vg, db):
        r2 = oracle_lib.OracleRun("dual_monod", 1, 2, 1e-6, 1)
        none = (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
        r2.add_flowmap(np.array([Vl]), *none)
        r2.add_gas_flowmap(np.array([Vg]), *none, eps=np.array([eps]))
        r2.set_concentrations(np.array([[5.0], [0.0]]))
        r2.set_gas_phase([0.0, H_O2], np.array([[0.0], [0.3]]))
        r2.set_particles(0, np.zeros(0, np.uint32), np.zeros(0, np.uint8), np.zeros((3, 0)), np.zeros(0),
                         np.zeros(0), 1.0)
        r2.set_kla("turbulence", species=1, bubble_diameter=db)
        m0 = r2.mass()[1, 0]
        r2.step_stages(T)
        assert r2.mass()[0, 0] == 5.0 * Vl  # no transfer of the other species
        return (r2.mass()[1, 0] - m0) / (1e-6 * (H_O2 * 0.3) * Vl)

    nu, sc = 1.0023e-6, 1.0023e-6 / 1e-9
    k = kla_of(0.5, 1e-3, 1e-3, 5e-3)
    np.testing.assert_allclose(k, 0.3 * (0.5 * nu) ** 0.25 / math.sqrt(sc) * 6.0 / 5e-3, rtol=1e-9)
    np.testing.assert_allclose(kla_of(8.0, 1e-3, 1e-3, 5e-3) / k, 2.0, rtol=1e-9)  # eps x16: kLa x2
    np.testing.assert_allclose(kla_of(0.5, 3e-3, 1e-3, 5e-3) / k, 1.0 / 3.0, rtol=1e-9)  # alpha 1/4: a = 2/d_b
    np.testing.assert_allclose(kla_of(0.5, 1e-3, 1e-3, 2.5e-3) / k, 2.0, rtol=1e-9)  # a ~ 1/d_b


def test_gas_transport_and_feeds_balance(oracle_lib):
    """Two gas compartments exchanging flow, a gas feed into the first and a
    gas outlet at the second, no transfer: the gas mass changes each step by
    dt*(q*c_feed - q*Cg_out), the liquid stays as it is."""
    r = oracle_lib.OracleRun("dual_monod", 2, 2, 0.01, 1)
    none = (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r.add_flowmap(np.array([1e-3, 1e-3]), *none)
    rows, cols = np.array([0, 1], np.uint32), np.array([1, 0], np.uint32)
    r.add_gas_flowmap(np.array([2e-4, 3e-4]), rows, cols, np.array([4e-5, 1e-5]))
    r.set_concentrations(np.array([[5.0, 5.0], [1e-3, 2e-3]]))
    r.set_gas_phase([0.0, H_O2], np.array([[0.0, 0.0], [0.1, 0.2]]))
    q, cf = 2e-5, 0.28
    r.add_gas_feed("constant", q, 0.0, 0, 1, [1], [cf])
    r.set_kla("fixed", [0.0, 0.0])
    r.set_particles(0, np.zeros(0, np.uint32), np.zeros(0, np.uint8), np.zeros((3, 0)), np.zeros(0), np.zeros(0), 1.0)
    liq = r.concentrations().copy()
    for _ in range(40):
        mg0 = r.gas_mass()[1].sum()
        cg_out = r.gas_concentrations()[1, 1]
        r.step_stages(T)
        np.testing.assert_allclose(r.gas_mass()[1].sum() - mg0, 0.01 * (q * cf - q * cg_out), rtol=1e-9)
        np.testing.assert_array_equal(r.concentrations(), liq)


@pytest.mark.parametrize("target,clipped", [(-1e-7, True), (-4e-7, True), (-1e-6, False), (-3e-6, False)])
def test_clear_negs_clips_small_negative_liquid_concentrations(oracle_lib, target, clipped):
    """Uptake drives the liquid glucose to about `target` in one transport
    step: a value above -1e-4*5e-3 is set to 0 (the mass keeps its negative
    value), one below stays."""
    Vl, c0, dt = 1e-3, 1e-3, 0.01

    def run(w):
        r = gl_box(oracle_lib, Vl, 2e-4, Cl=(c0, 1.0), Cg=(0.0, 0.0), dt=dt)
        r.set_kla("fixed", [0.0, 0.0])
        r.set_params([2.8e-4, 0.01, 1e-4, 0.5, 1.0, 100.0, 1e-4, 1e9, 1.4e-15, 0.2e-15])  # no division
        n = 4
        x = np.stack([np.full(n, 1e-15), np.full(n, 1e-4), np.zeros(n)])
        r.set_particles(n, np.zeros(n, np.uint32), np.zeros(n, np.uint8), x, np.zeros(n), np.zeros(n), w)
        return r

    a = run(1.0)
    a.step_stages(T | M | C_)
    s1 = a.contributions()[0, 0]  # uptake at weight 1 (linear in the weight)
    w = (target * Vl - c0 * Vl) / (dt * s1)
    r = run(w)
    r.step_stages(T | M | C_)
    r.step_stages(T | M | C_)
    m = r.mass()[0, 0]
    c = r.concentrations()[0, 0]
    assert m < 0.0
    if clipped:
        assert c == 0.0
    else:
        assert c == m * (1.0 / Vl)
