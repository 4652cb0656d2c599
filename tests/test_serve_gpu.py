"""Runs that join and leave a stepping engine (serving) and large ragged
batches: every run's result is the same in the mix as alone, and matches
the oracle (events bit-exact, FP64 state within 1e-12)."""
import numpy as np
import pytest

from parity import assert_parity, assert_snapshot, close, snapshot

pytestmark = pytest.mark.gpu


def _oracle(c, steps):
    from oracle.oracle import OracleRun

    o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(o)
    o.step(steps)
    return o


def _fast(c):
    """Growth fast enough that divisions happen within the test's steps."""
    c.params = np.array([0.5, 0.01, 0.5, 2.0, 1e-15, 0.1])
    return c


def test_served_runs_match_oracle_and_runs_alone():
    import mcst
    from mcst import cases
    from mcst.serve import Request, Server

    steps = [30, 45, 12, 60, 25, 38]
    reqs = []
    for i, k in enumerate(steps):
        c = _fast(cases.ensemble_run(i, seed=21, n_steps=k, n_c_range=(8, 120), n_range=(1500, 12000)))
        reqs.append(Request(i, 0.015 * i, c, k))
    snaps = {}
    # a pool of ~2 runs: later requests wait for earlier ones to leave
    srv = Server("monod_mass", slots=2 * 24 * 1024, max_runs=16, budget_bytes=1 << 40, chunk_steps=7)
    res = srv.serve(reqs, on_finish=lambda req, run: snaps.__setitem__(req.rid, snapshot(run)))
    srv.close()
    assert [r.steps for r in res] == steps
    for req in reqs:
        assert_snapshot(snaps[req.rid], _oracle(req.case, req.steps))
    # and alone on the GPU: the same positions and events, states within 1e-12
    for req in reqs[:2]:
        eng = mcst.Engine("monod_mass")
        g = eng.add_run(req.case.n_c, req.case.dt, req.case.seed)
        req.case.apply(g)
        eng.commit()
        eng.step(req.steps)
        a = snapshot(g)
        s = snaps[req.rid]
        assert a["counters"] == s["counters"]
        np.testing.assert_array_equal(a["particles"]["pos"], s["particles"]["pos"])
        np.testing.assert_array_equal(a["particles"]["status"], s["particles"]["status"])
        close(a["particles"]["state"], s["particles"]["state"], 1e-12, "state")
        close(a["C"], s["C"], 1e-12, "concentrations")
        eng.close()


def test_admitted_run_starts_at_its_own_step_zero():
    import mcst
    from mcst import cases

    c1 = _fast(cases.ensemble_run(1, seed=4, n_c_range=(30, 30), n_range=(5000, 5000)))
    c2 = _fast(cases.ensemble_run(2, seed=4, n_c_range=(50, 50), n_range=(7000, 7000)))
    eng = mcst.Engine("monod_mass")
    eng.reserve(64 * 1024, 8)
    a = eng.add_run(c1.n_c, c1.dt, c1.seed)
    c1.apply(a)
    eng.commit()
    eng.step(10)
    b = eng.add_run(c2.n_c, c2.dt, c2.seed)
    c2.apply(b)
    assert eng.admit() == 1
    eng.step(15)
    assert a.counters()["steps"] == 25 and b.counters()["steps"] == 15
    assert_parity(a, _oracle(c1, 25), probes=False)
    assert_parity(b, _oracle(c2, 15), probes=False)
    eng.retire(a)
    eng.step(5)
    assert_parity(b, _oracle(c2, 20), probes=False)
    eng.close()


def test_retired_first_run_leaves_no_reads_behind():
    """Run 0 is large (40 000 compartments: a 1.3 MB compartment record) and
    retires first; its arrays are freed while the other runs keep stepping.
    The particle pass's phantom sub-chunks (the prefetch past the last group)
    stand at the last live tile, never at run 0, so the survivors step on
    memory they own and still match their oracles over many steps, with
    every launch shape of the pass."""
    import os

    import mcst
    from mcst import cases

    big = _fast(cases.ensemble_run(0, seed=31, n_c_range=(40000, 40000), n_range=(30000, 30000)))
    small = [_fast(cases.ensemble_run(i, seed=31, n_c_range=(40, 300), n_range=(3000, 9000))) for i in (1, 2)]
    for block in ("256", "512"):
        os.environ["MCST_PARTICLE_BLOCK"] = block
        try:
            eng = mcst.Engine("monod_mass")
            runs = []
            for c in [big] + small:
                r = eng.add_run(c.n_c, c.dt, c.seed)
                c.apply(r)
                runs.append(r)
            eng.commit()
            eng.step(3)
            eng.retire(runs[0])
            eng.step(40)
            for r, c in zip(runs[1:], small):
                assert_parity(r, _oracle(c, 43), probes=False)
            eng.close()
        finally:
            del os.environ["MCST_PARTICLE_BLOCK"]


def test_admit_reports_a_full_pool_and_retries():
    import mcst
    from mcst import cases

    cs = [_fast(cases.ensemble_run(i, seed=8, n_c_range=(20, 20), n_range=(9000, 9000))) for i in range(3)]
    eng = mcst.Engine("monod_mass")
    eng.reserve(2 * 18 * 1024, 8)  # two of the three runs fit
    runs = []

    def define(c):
        r = eng.add_run(c.n_c, c.dt, c.seed)
        c.apply(r)
        r.set_step_target(6)
        runs.append(r)

    define(cs[0])
    define(cs[1])
    eng.commit()
    define(cs[2])
    assert eng.admit() == 0  # no room: the run stays defined
    with pytest.raises(mcst.MCSTError):
        runs[2].counters()
    eng.step(6)
    eng.retire(runs[0])
    assert eng.admit() == 1
    eng.step(6)
    assert_parity(runs[1], _oracle(cs[1], 6), probes=False)
    assert_parity(runs[2], _oracle(cs[2], 6), probes=False)
    eng.close()


def test_ragged_ensemble_of_48_runs():
    """A ensemble512-shaped batch (uneven compartment and particle counts, own
    seeds and parameters), smaller sizes: every run matches its oracle."""
    import mcst
    from mcst import cases

    cs = cases.ensemble512(n_runs=48, seed=3, n_steps=8, n_c_range=(32, 600), n_range=(500, 40000))
    eng = mcst.Engine("monod_mass")
    runs = []
    for c in cs:
        r = eng.add_run(c.n_c, c.dt, c.seed)
        c.apply(r)
        runs.append(r)
    eng.commit()
    eng.capture(4)
    eng.replay(2)
    for c, r in zip(cs, runs):
        assert_parity(r, _oracle(c, 8), probes=False)
    eng.close()


def test_two_models_served_together():
    """Requests of two models (one engine each, on their own streams) served
    in one loop: every result matches its oracle run."""
    from mcst import cases
    from mcst.serve import Request, Server

    reqs = []
    for i, k in enumerate([20, 35, 15]):
        c = _fast(cases.ensemble_run(i, seed=5, n_steps=k, n_c_range=(8, 60), n_range=(1000, 6000)))
        reqs.append(Request(i, 0.01 * i, c, k))
    q = cases.small_case("quad_monod", n_c=12, n_particles=5000, seed=8)
    q.capacity = 40000
    reqs.append(Request(3, 0.005, q, 26))
    snaps = {}
    srv = Server(pools={"monod_mass": (8 * 12 * 1024, 8), "quad_monod": (48 * 1024, 2)}, budget_bytes=1 << 40,
                 chunk_steps=6)
    res = srv.serve(reqs, on_finish=lambda req, run: snaps.__setitem__(req.rid, snapshot(run)))
    srv.close()
    assert [r.steps for r in res] == [20, 35, 15, 26]
    for req in reqs:
        assert_snapshot(snaps[req.rid], _oracle(req.case, req.steps))
