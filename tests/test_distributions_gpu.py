"""Under independent random streams (the HIP engine and the oracle seeded
differently) the populations agree in distribution: two-sample KS tests at
p > 0.01 on every state variable, division age and hydraulic age of the
particles, on residence times (exit ages) and division ages, a chi-square
test of the compartment occupancy, and event rates per particle-step.

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
        r.set_stage_cap(1024)
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


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod"])
def test_population_distributions_match_under_independent_streams(model):
    from mcst import cases

    c = cases.small_case(model, n_c=12, n_particles=20000, seed=5)
    c.capacity = 400_000
    eng, g, o = _pair(c, seed_gpu=1001, seed_cpu=2002, steps=200)
    pg, po = g.particles(), o.particles()
    ig, io = _subset(pg, 1), _subset(po, 2)
    for v in range(g.n_var):
        a, b = pg["state"][v][ig], po["state"][v][io]
        p = stats.ks_2samp(a, b).pvalue
        assert p > P_MIN, (model, "state", v, p)
    for f in ("age_div", "age_hyd"):
        p = stats.ks_2samp(pg[f][ig], po[f][io]).pvalue
        assert p > P_MIN, (model, f, p)
    # residence times of the particles that left through the outlet, and
    # the ages at division
    for what in ("exit_ages", "division_ages"):
        a, b = getattr(g, what)(), getattr(o, what)()
        assert len(a) > 1000 and len(b) > 1000, (what, len(a), len(b))
        p = stats.ks_2samp(a, b).pvalue
        assert p > P_MIN, (model, what, p)
    # event rates per particle-step: a two-proportion z-test, |z| < 5 (the
    # raw counts also carry the spread of the population size)
    cg, co = g.counters(), o.counters()
    assert cg["overflow"] == co["overflow"] == 0
    ng, no = cg["processed"], co["processed"]
    for k in ("births", "exits", "moves"):
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
