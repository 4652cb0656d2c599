"""The autotuner's trial steps are ordinary steps: after tuning, a batch of
runs still matches the oracle run for run, and the chosen shape persists
across admission of a new run."""
import pytest

from parity import assert_parity

pytestmark = pytest.mark.gpu


def test_autotuned_batch_matches_oracle():
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    specs = [(12, 3000, 1), (30, 9000, 3), (7, 500, 2)]
    eng = mcst.Engine("monod_mass")
    pairs = []
    for n_c, n, seed in specs:
        c = cases.small_case("monod_mass", n_c=n_c, n_particles=n, seed=seed)
        c.capacity = 20 * n
        g = eng.add_run(c.n_c, c.dt, c.seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        for r in (g, o):
            c.apply(r)
        g.set_probes(1 << 20)
        pairs.append((g, o))
    eng.commit()
    choice = eng.autotune(2)  # 3 shapes x 5 groupings x 2 steps at most
    assert choice["block"] in (256, 512, 1024) and choice["group"] in (0, 1, 2, 4, 8)
    assert choice["particle_ms"] > 0
    done = pairs[0][0].counters()["steps"]
    assert done % 2 == 0 and done >= 10
    eng.step(7)
    for g, o in pairs:
        o.step(done + 7)
        assert_parity(g, o)
    eng.close()
