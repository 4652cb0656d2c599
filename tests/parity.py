"""Helpers that build the same run on the HIP engine and on the CPU oracle
and compare them: events and integers bit-exact, FP64 state within 1e-12."""
import numpy as np

RTOL = 1e-12


def make_pair(case, division_cap=None, ratio=0.01, min_removal=0, probes=1 << 20, capacity=None):
    import mcst
    from oracle.oracle import OracleRun

    eng = mcst.Engine(case.model)
    g = eng.add_run(case.n_c, case.dt, case.seed)
    o = OracleRun(case.model, case.n_c, case.n_s, case.dt, case.seed)
    if capacity is not None:
        case.capacity = capacity
    for r in (g, o):
        case.apply(r)
        if division_cap:
            r.set_division_cap(division_cap)
        r.set_compaction(ratio, min_removal)
    g.set_probes(probes)
    eng.commit()
    return eng, g, o


def close(a, b, rtol=RTOL, what=""):
    """|a-b| <= rtol*|b| + rtol*max|b| of the row (a state variable near zero
    carries the absolute error of its larger values)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.size == 0:
        return
    scale = np.max(np.abs(b), axis=-1, keepdims=True) if b.ndim > 1 else np.max(np.abs(b))
    err = np.abs(a - b)
    lim = rtol * np.abs(b) + rtol * scale
    bad = err > lim
    assert not bad.any(), f"{what}: {bad.sum()} values differ, worst rel {np.max(err / (np.abs(b) + 1e-300))}"


def assert_parity(g, o, rtol=RTOL, probes=True):
    cg, co = g.counters(), o.counters()
    assert cg == co, (cg, co)
    pg, po = g.particles(), o.particles()
    np.testing.assert_array_equal(pg["pos"], po["pos"])
    np.testing.assert_array_equal(pg["status"], po["status"])
    close(pg["state"], po["state"], rtol, "state")
    close(pg["age_hyd"], po["age_hyd"], rtol, "age_hyd")
    close(pg["age_div"], po["age_div"], rtol, "age_div")
    close(g.concentrations(), o.concentrations(), rtol, "concentrations")
    close(g.mass(), o.mass(), rtol, "liquid mass")
    close(g.contributions(), o.contributions(), 1e-11, "contributions")
    np.testing.assert_array_equal(g.repartition(), o.repartition())
    if probes:
        close(np.sort(g.exit_ages()), np.sort(o.exit_ages()), rtol, "exit ages")
        close(np.sort(g.division_ages()), np.sort(o.division_ages()), rtol, "division ages")


def snapshot(run, probes=False):
    """Everything assert_parity compares, read from a run now (a served run
    leaves the device once it completes)."""
    s = {"counters": run.counters(), "particles": run.particles(), "C": run.concentrations(),
         "m": run.mass(), "S": run.contributions(), "rep": run.repartition()}
    if probes:
        s["exit_ages"] = run.exit_ages()
        s["division_ages"] = run.division_ages()
    return s


def assert_snapshot(s, o, rtol=RTOL):
    """A snapshot of a HIP run against an oracle run at the same step."""
    co = o.counters()
    assert s["counters"] == co, (s["counters"], co)
    pg, po = s["particles"], o.particles()
    np.testing.assert_array_equal(pg["pos"], po["pos"])
    np.testing.assert_array_equal(pg["status"], po["status"])
    close(pg["state"], po["state"], rtol, "state")
    close(pg["age_hyd"], po["age_hyd"], rtol, "age_hyd")
    close(pg["age_div"], po["age_div"], rtol, "age_div")
    close(s["C"], o.concentrations(), rtol, "concentrations")
    close(s["m"], o.mass(), rtol, "liquid mass")
    close(s["S"], o.contributions(), 1e-11, "contributions")
    np.testing.assert_array_equal(s["rep"], o.repartition())
