"""The full-size workloads: the 512-run ensemble (500 steps), the 50 000-
compartment lattice tank at 1e8 particles (200 steps), and the 40-request
served stream.  At these sizes the oracle checks a sample of runs; every run
is checked through properties that hold at any size:

- the Idle count: n - inactive = n0 + births - exits - deaths, and the
  repartition (Idle particles per compartment) sums to it;
- in a closed system (no feed, no outlet, no gas exchange), the conservation
  of substrate plus the biomass that consumed it, liquid mass counted with
  the contributions pending for the next transport, to 1e-10 relative.

For monod_mass (growth dm = dt*g, contribution -g/Y) the conserved quantity
is  sum(M_glucose + dt*S_glucose) + w*sum(m)/Y.  For quad_monod (glucose
uptake Us split into respiration Resp and overflow Ovf = Us - Resp; growth
Yxs*Resp + Yxa*Ua; acetate p9*Ovf - Ua; storage (1-p9)*Ovf) it is
  L_glucose + (Yxa/Yxs)*L_acetate + w*(sum(m)/Yxs + e*sum(X)),
  e = (1 - Yxa*p9/Yxs)/(1 - p9),
the combination whose rate vanishes for every Us, Resp and Ua."""
import sys

import numpy as np
import pytest

from parity import assert_parity, assert_snapshot, snapshot

pytestmark = pytest.mark.gpu

RTOL_BALANCE = 1e-10


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _liquid(run, s):
    """Liquid mass of species s with the contributions the next transport adds."""
    return run.mass()[s].sum() + run.dt * run.contributions()[s].sum()


def _monod_total(run, Y, biomass):
    return _liquid(run, 0) + run.weight * biomass / Y


def _gpu_biomass(run, rows=(0,)):
    """Sums of state variables over the Idle particles, from the per-compartment
    export sums (no per-particle copy)."""
    _, spatial = run.properties(values=False)
    return [spatial[v].sum() for v in rows]


def _idle_identity(run, n0):
    c = run.counters()
    idle = c["n"] - c["inactive"]
    assert idle == n0 + c["births"] - c["exits"] - c["deaths"], (n0, c)
    assert int(run.repartition().sum()) == idle, (idle, c)
    return c


def _oracle(case, steps, threads=16):
    from oracle.oracle import OracleRun

    o = OracleRun(case.model, case.n_c, case.n_s, case.dt, case.seed)
    case.apply(o)
    if steps:
        o.step(steps, threads=threads)
    return o


def test_ensemble512_full_size():
    """512 runs of 32..4000 compartments and 1e4..2e6 particles (1.96e8 in
    all), 500 steps in one batched engine: the Idle identity and the mass
    balance of every run; three runs (the two smallest and a median one)
    against the oracle, bit-exact in events and to 1e-12 in state."""
    import mcst
    from mcst import cases

    cs = cases.ensemble512()
    log(f"[ensemble512] defining {len(cs)} runs, {sum(c.n_particles for c in cs):.3e} particles")
    eng = mcst.Engine("monod_mass")
    runs = []
    for c in cs:
        r = eng.add_run(c.n_c, c.dt, c.seed)
        c.apply(r)
        runs.append(r)
    eng.commit()
    t0 = [_monod_total(r, c.params[2], _gpu_biomass(r)[0]) for c, r in zip(cs, runs)]
    log("[ensemble512] committed; stepping 500")
    for k in range(5):
        eng.step(100)
        eng.synchronize()
        log(f"[ensemble512] {100 * (k + 1)} steps")
    births = 0
    for c, r, a in zip(cs, runs, t0):
        cnt = _idle_identity(r, c.n_particles)
        assert cnt["steps"] == 500 and cnt["exits"] == 0 and cnt["deaths"] == 0
        births += cnt["births"]
        b = _monod_total(r, c.params[2], _gpu_biomass(r)[0])
        assert abs(b - a) <= RTOL_BALANCE * abs(a), (c.name, a, b)
    assert births > 0
    order = sorted(range(len(cs)), key=lambda i: cs[i].n_particles)
    for i in (order[0], order[1], order[len(order) // 2]):
        log(f"[ensemble512] oracle for {cs[i].name} ({cs[i].n_particles} particles)")
        assert_parity(runs[i], _oracle(cs[i], 500), probes=False)
    eng.close()


def test_lattice50k_1e8_particles_200_steps():
    """The lattice tank as specified (2500 bottom-to-top feeds, fixed-gas
    oxygen supply, quad_monod, 5 ms steps) at 1e8 particles, all 200 steps:
    divisions, exits and compactions happen, no division is refused, the
    Idle identity holds, and the liquid state stays finite."""
    import mcst
    from mcst import cases

    c = cases.lattice50k(n_particles=100_000_000)
    log(f"[lattice1e8] defining {c.n_particles:.1e} particles, capacity {c.capacity:.2e}")
    eng = mcst.Engine(c.model)
    g = eng.add_run(c.n_c, c.dt, c.seed)
    c.apply(g)
    eng.commit()
    log("[lattice1e8] committed; stepping 200")
    for k in range(4):
        eng.step(50)
        eng.synchronize()
        log(f"[lattice1e8] {50 * (k + 1)} steps, n = {g.n}")
    cnt = _idle_identity(g, c.n_particles)
    assert cnt["steps"] == 200 and cnt["overflow"] == 0
    assert cnt["births"] > 0.5 * c.n_particles and cnt["exits"] > 0 and cnt["compactions"] > 0
    assert np.isfinite(g.concentrations()).all() and np.isfinite(g.contributions()).all()
    eng.close()


def test_lattice50k_1e8_particles_closed_balance():
    """The lattice geometry at 1e8 particles with its feeds, outlet and gas
    exchange removed, 200 steps: the quad_monod substrate-plus-biomass
    combination is conserved to 1e-10 at every 50 steps."""
    import mcst
    from mcst import cases

    c = cases.lattice50k(n_particles=100_000_000)
    c.feeds = []
    c.gas = None
    p = c.params
    Yxs, Yxa, p9 = p[5], p[6], p[9]
    e = (1.0 - Yxa * p9 / Yxs) / (1.0 - p9)
    log(f"[lattice1e8 closed] defining {c.n_particles:.1e} particles")
    eng = mcst.Engine(c.model)
    g = eng.add_run(c.n_c, c.dt, c.seed)
    c.apply(g)
    eng.commit()

    def total():
        m, X = _gpu_biomass(g, rows=(0, 3))
        return _liquid(g, 0) + (Yxa / Yxs) * _liquid(g, 2) + g.weight * (m / Yxs + e * X)

    i0 = total()
    for k in range(4):
        eng.step(50)
        eng.synchronize()
        i1 = total()
        log(f"[lattice1e8 closed] {50 * (k + 1)} steps, n = {g.n}, drift {(i1 - i0) / i0:.2e}")
        assert abs(i1 - i0) <= RTOL_BALANCE * abs(i0), (k, i0, i1)
    cnt = _idle_identity(g, c.n_particles)
    assert cnt["births"] > 0.5 * c.n_particles and cnt["exits"] == 0
    eng.close()


def test_served_stream_of_40_requests():
    """The served stream (40 requests, Poisson arrivals at 20/s, 200..5000
    steps, up to 2e6 particles): every request finishes at its own step
    count with the Idle identity and its mass balance; the four requests of
    fewest particle-steps against the oracle."""
    from mcst import cases
    from mcst.serve import Request, Server

    reqs = [Request(i, t, c, c.n_steps) for i, (t, c) in enumerate(cases.served_stream(40))]
    order = sorted(reqs, key=lambda q: q.case.n_particles * q.steps)
    sample = {q.rid for q in order[:4]}
    log("[served40] initial totals from the oracle's initial state")
    init = {}
    for q in reqs:
        o = _oracle(q.case, 0)
        st = o.particles()["state"]
        init[q.rid] = _monod_total(o, q.case.params[2], st[0].sum())
        del o
    seen, snaps = {}, {}

    def on_finish(req, run):
        cnt = _idle_identity(run, req.case.n_particles)
        assert cnt["steps"] == req.steps
        seen[req.rid] = _monod_total(run, req.case.params[2], _gpu_biomass(run)[0])
        if req.rid in sample:
            snaps[req.rid] = snapshot(run)

    slots = 2 * sum(q.case.n_particles for q in reqs) + 1024 * len(reqs)
    srv = Server("monod_mass", slots=slots, max_runs=len(reqs), chunk_steps=20)
    log("[served40] serving")
    res = srv.serve(reqs, on_finish=on_finish)
    srv.close()
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    for q in reqs:
        a, b = init[q.rid], seen[q.rid]
        assert abs(b - a) <= RTOL_BALANCE * abs(a), (q.rid, a, b)
    for q in reqs:
        if q.rid in sample:
            log(f"[served40] oracle for request {q.rid} ({q.case.n_particles} particles, {q.steps} steps)")
            assert_snapshot(snaps[q.rid], _oracle(q.case, q.steps))
