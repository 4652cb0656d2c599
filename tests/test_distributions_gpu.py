"""Under independent random streams (the HIP engine and the oracle seeded
differently) the populations agree in distribution: two-sample KS tests on
every state variable, division age and hydraulic age of the particles, on
residence times (exit ages) and division ages, at a family-wise level of
0.01 per model (Bonferroni over its quantities), a chi-square test of the
compartment occupancy, and event rates per particle-step.

Sister cells are identical at birth, so a population is not an independent
sample; the tests draw a fixed-size random subset of each population (few
sister pairs end up in it), which keeps the KS p-values honest."""
import math

import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

P_MIN = 0.01
SUBSET = 3000


def _pair(case, seed_gpu, seed_cpu, steps):
    import mcst
    from oracle.oracle import OracleRun

    eng = mcst.Engine(case.model)
    g = eng.add_run(case.n_c, case.dt, seed_gpu)
    o = OracleRun(case.model, case.n_c, case.n_s, case.dt, seed_cpu)
    for r in (g, o):
        case.apply(r)
        r.set_compaction(0.01, 0)
    g.set_probes(1 << 20)
    eng.commit()
    eng.step(steps)
    o.step(steps, threads=8)
    return eng, g, o


def _subset(p, seed):
    idle = np.flatnonzero(p["status"] == 0)
    if len(idle) <= SUBSET:
        return idle
    return np.sort(np.random.default_rng(seed).choice(idle, SUBSET, replace=False))


def _oracle(case, seed, steps):
    from oracle.oracle import OracleRun

    o = OracleRun(case.model, case.n_c, case.n_s, case.dt, seed)
    case.apply(o)
    o.set_compaction(0.01, 0)
    o.step(steps, threads=8)
    return o


@pytest.mark.parametrize("model, steps", [("monod_mass", 200), ("dual_monod", 200), ("quad_monod", 200),
                                          ("simple_acetate", 200), ("udf_starvation", 60)])
def test_population_distributions_match_under_independent_streams(model, steps):
    """Also the 8-state model, the model whose division redistributes its
    state with random draws, and the plugin model whose cells die.

    Some variables follow the liquid state of the realization (an uptake
    rate that tracks a nearly depleted substrate): two oracle runs with
    different seeds already differ in them beyond what KS allows.  A second
    oracle run (a third seed) calibrates: every quantity on which the two
    oracle runs agree at the corrected level must also agree between the HIP
    run and the oracle at that level, and at least three quarters of the
    quantities must be of that kind.  The level is P_MIN / (number of
    quantities), so that a model passes by chance with probability >= 0.99."""
    import mcst
    from mcst import cases

    if model == "udf_starvation":
        mcst.load_model(mcst.EXAMPLE_PLUGIN_DEATH)
    c = cases.small_case(model, n_c=12, n_particles=20000, seed=5)
    if model == "udf_starvation":
        c.C0[0, ::3] = 0.0
    c.capacity = 400_000
    eng, g, o = _pair(c, seed_gpu=1001, seed_cpu=2002, steps=steps)
    o2 = _oracle(c, 3003, steps)
    pg, po, p2 = g.particles(), o.particles(), o2.particles()
    ig, io, i2 = _subset(pg, 1), _subset(po, 2), _subset(p2, 3)
    assert len(ig) > 1000 and len(io) > 1000, (len(ig), len(io))
    samples = [(f"state{v}", pg["state"][v][ig], po["state"][v][io], p2["state"][v][i2]) for v in range(g.n_var)]

    # Ages are whole numbers of steps times dt, and many particles share one
    # (the initial ones, sisters): they are compared in steps, since the
    # engine's k*dt and the oracle's running sum of dt differ in the last bits,
    # and KS would count that as a jump of the whole tie
    def steps_of(v):
        return np.rint(np.asarray(v) / c.dt)

    samples += [(f, steps_of(pg[f][ig]), steps_of(po[f][io]), steps_of(p2[f][i2])) for f in ("age_div", "age_hyd")]
    # residence times of the particles that left through the outlet, and
    # the ages at division
    for what in ("exit_ages", "division_ages"):
        a, b, r = (steps_of(getattr(x, what)()) for x in (g, o, o2))
        if min(len(a), len(b), len(r)) > 300:
            samples.append((what, a, b, r))
    def ks(x, y):  # asymptotic p-values: samples of a few thousand
        return stats.ks_2samp(x, y, method="asymp").pvalue

    level = P_MIN / len(samples)
    checked = 0
    for name, a, b, r in samples:
        if ks(b, r) <= level:
            continue  # set by the realization's liquid state: not comparable
        checked += 1
        p = ks(a, b)
        assert p > level, (model, name, p, level)
    assert 4 * checked >= 3 * len(samples), (model, checked, len(samples))
    # event rates per particle-step: a two-proportion z-test, |z| < 5 (the
    # raw counts also carry the spread of the population size)
    cg, co = g.counters(), o.counters()
    assert cg["overflow"] == co["overflow"] == 0
    ng, no = cg["processed"], co["processed"]
    for k in ("births", "exits", "moves", "deaths"):
        if cg[k] + co[k] == 0:
            continue
        pool = (cg[k] + co[k]) / (ng + no)
        sd = math.sqrt(pool * (1 - pool) * (1 / ng + 1 / no))
        z = (cg[k] / ng - co[k] / no) / sd
        assert abs(z) < 5, (model, k, cg[k], ng, co[k], no, z)
    eng.close()


def test_compartment_occupancy_matches_under_independent_streams():
    """Particle counts per compartment after mixing (slow growth, closed
    network): a chi-square test of homogeneity of the two repartitions."""
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=12, n_particles=30000, seed=6, feeds=False, fast=False)
    eng, g, o = _pair(c, seed_gpu=11, seed_cpu=12, steps=300)
    table = np.stack([g.repartition(), o.repartition()]).astype(float)
    table = table[:, table.sum(axis=0) > 0]
    p = stats.chi2_contingency(table)[1]
    assert p > P_MIN, p
    eng.close()
