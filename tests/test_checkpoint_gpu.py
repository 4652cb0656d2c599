"""Checkpoint and resume of device-resident runs: a run restored from a
checkpoint continues as the original (events bit for bit, FP64 within the
parity bound), and both match the oracle."""
import numpy as np
import pytest

from parity import RTOL, assert_parity, close, make_pair

pytestmark = pytest.mark.gpu


def _define(eng, case, probes=1 << 20):
    r = eng.add_run(case.n_c, case.dt, case.seed)
    case.apply(r, particles=False)
    r.set_compaction(0.01, 0)
    r.set_probes(probes)
    return r


def _same(a, b):
    """Events and integers bit for bit; FP64 within the parity bound (the
    contribution sums are FP64 atomics, whose order differs between runs)."""
    assert a.counters() == b.counters()
    pa, pb = a.particles(), b.particles()
    for k in ("pos", "status"):
        np.testing.assert_array_equal(pa[k], pb[k], err_msg=k)
    for k in ("state", "age_hyd", "age_div"):
        close(pa[k], pb[k], RTOL, k)
    np.testing.assert_array_equal(a.repartition(), b.repartition())
    for f in ("concentrations", "mass"):
        close(getattr(a, f)(), getattr(b, f)(), RTOL, f)
    close(a.contributions(), b.contributions(), 1e-11, "contributions")
    for f in ("exit_ages", "division_ages"):
        close(np.sort(getattr(a, f)()), np.sort(getattr(b, f)()), RTOL, f)


@pytest.mark.parametrize("model", ["monod_mass", "dual_monod", "simple_acetate", "udf_starvation", "timefeed"])
def test_restored_run_continues_exactly(model):
    """Also a plugin model whose cells die, and a run whose outlet is led by a
    linear feed (its flows follow the restored step count)."""
    import mcst
    from mcst import cases

    if model == "udf_starvation":
        mcst.load_model(mcst.EXAMPLE_PLUGIN_DEATH)
    c = cases.small_case("dual_monod" if model == "timefeed" else model, n_c=12, n_particles=6000, seed=21)
    if model == "udf_starvation":
        c.C0[0, ::3] = 0.0
    if model == "timefeed":
        c.feeds = [("linear", 1e-3, 0.02, 3, 9, [0, 1], [8.0, 0.5]), ("constant", 2e-3, 0.0, 0, 6, [0], [10.0])]
    c.capacity = 60000
    eng, g, o = make_pair(c)
    eng.step(31)  # mid flowmap cycle (3 steps per map, 2 maps), after divisions and exits
    blob = g.checkpoint()
    eng.step(29)
    o.step(60)
    assert_parity(g, o)

    eng2 = mcst.Engine(c.model)
    h = _define(eng2, c)
    h.restore(blob)
    eng2.commit()
    assert h.counters()["steps"] == 31
    eng2.step(29)
    _same(g, h)
    eng.close()
    eng2.close()


def test_restore_into_a_batch_and_checkpoint_again():
    """A restored run next to a fresh one; a checkpoint of the restored run
    taken later restores again."""
    import mcst
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=9, n_particles=4000, seed=5)
    c.capacity = 40000
    other = cases.small_case("monod_mass", n_c=20, n_particles=7000, seed=6)
    other.capacity = 40000
    eng, g, o = make_pair(c)
    eng.step(10)
    blob = g.checkpoint()

    eng2 = mcst.Engine("monod_mass")
    x = _define(eng2, other)
    x.init_particles(other.n_particles, other.capacity, other.x0)
    h = _define(eng2, c)
    h.restore(blob)
    eng2.commit()
    eng2.step(15)
    blob2 = h.checkpoint()
    eng2.step(5)
    eng.step(20)
    _same(g, h)

    eng3 = mcst.Engine("monod_mass")
    k = _define(eng3, c)
    k.restore(blob2)
    eng3.commit()
    eng3.step(5)
    _same(g, k)
    o.step(30)
    assert_parity(k, o)
    for e in (eng, eng2, eng3):
        e.close()


def test_restore_rejects_a_foreign_or_truncated_checkpoint():
    import mcst
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=6, n_particles=2000, seed=2)
    eng, g, _ = make_pair(c)
    eng.step(3)
    blob = g.checkpoint()
    eng2 = mcst.Engine("monod_mass")
    r = eng2.add_run(c.n_c, c.dt, c.seed + 1)  # another seed
    c.apply(r, particles=False)
    with pytest.raises(mcst.MCSTError, match="seed"):
        r.restore(blob)
    r2 = eng2.add_run(c.n_c, c.dt, c.seed)
    c.apply(r2, particles=False)
    with pytest.raises(mcst.MCSTError, match="truncated"):
        r2.restore(blob[: len(blob) // 2])
    with pytest.raises(mcst.MCSTError, match="checkpoint"):
        r2.restore(b"\0" * len(blob))
    eng3 = mcst.Engine("dual_monod")
    r3 = eng3.add_run(c.n_c, c.dt, c.seed)
    with pytest.raises(mcst.MCSTError, match="model"):
        r3.restore(blob)
    for e in (eng, eng2, eng3):
        e.close()
