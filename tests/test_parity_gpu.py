"""Parity of the HIP path with the CPU oracle on the same seeded inputs:
positions, statuses, counters and repartition bit-exact; FP64 particle and
liquid state within 1e-12 relative."""
import numpy as np
import pytest

from parity import assert_parity, close, make_pair

pytestmark = pytest.mark.gpu

MODELS = ["monod_mass", "dual_monod", "monod", "fixed_length", "simple_acetate", "quad_monod"]
TRANSPORT, MODEL, CONTRIB, MOVE, LEAVE, REMOVE = 1, 2, 4, 8, 16, 32


@pytest.mark.parametrize("model", MODELS)
def test_full_step_parity(model):
    from mcst import cases

    c = cases.small_case(model, n_particles=5000)
    eng, g, o = make_pair(c)
    for k in range(4):
        eng.step(5)
        o.step(5)
        assert_parity(g, o)
    cnt = g.counters()
    assert cnt["moves"] > 0 and cnt["exits"] > 0 and cnt["births"] > 0 and cnt["compactions"] > 0


@pytest.mark.parametrize("stages", [TRANSPORT, MOVE, LEAVE | REMOVE, MODEL, MODEL | CONTRIB,
                                    TRANSPORT | MODEL | CONTRIB, MOVE | LEAVE])
def test_single_stage_parity(stages):
    from mcst import cases

    c = cases.small_case("dual_monod", n_particles=3000)
    eng, g, o = make_pair(c)
    for _ in range(6):
        eng.step_stages(stages)
        o.step_stages(stages)
    assert_parity(g, o)


@pytest.mark.parametrize("n", [0, 1, 63, 1023, 1024, 1025, 5000])
def test_ragged_particle_counts(n):
    from mcst import cases

    c = cases.small_case("monod_mass", n_particles=n)
    eng, g, o = make_pair(c, capacity=max(2 * n, 8))
    eng.step(8)
    o.step(8)
    assert_parity(g, o)


def test_zero_d_reactor_without_move():
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=1, n_particles=4000)
    eng, g, o = make_pair(c)
    eng.step(15)
    o.step(15)
    assert_parity(g, o)
    assert g.counters()["moves"] == 0 and g.counters()["exits"] > 0


def test_closed_network_without_leave():
    from mcst import cases

    c = cases.small_case("monod", n_particles=4000, feeds=False)
    eng, g, o = make_pair(c)
    eng.step(12)
    o.step(12)
    assert_parity(g, o)
    assert g.counters()["exits"] == 0


def test_division_overflow_and_tight_capacity():
    """Requests beyond the per-step cap, then beyond the free slots, are
    refused in slot order over the whole run (Overflow) and retried."""
    from mcst import cases

    c = cases.small_case("simple_acetate", n_particles=3000)
    eng, g, o = make_pair(c, division_cap=8, capacity=3100)
    eng.step(10)
    o.step(10)
    assert_parity(g, o)
    assert g.counters()["overflow"] > 0
    eng.step(25)
    o.step(25)
    assert_parity(g, o)
    assert g.n == 3100


def test_everything_leaves():
    """A fast outlet empties the reactor: removal of all slots, then nothing."""
    from mcst import cases

    c = cases.small_case("fixed_length", n_c=1, n_particles=3000)
    c.feeds = [("constant", 0.5, 0.0, 0, 0, [0], [1.0])]
    eng, g, o = make_pair(c, ratio=0.0)
    eng.step(30)
    o.step(30)
    assert_parity(g, o)
    assert g.n < 3000


def test_removal_thresholds():
    from mcst import cases

    c = cases.small_case("monod_mass", n_particles=6000)
    for ratio, mn in ((0.0, 0), (0.05, 0), (0.0, 200), (1.0, 0)):
        eng, g, o = make_pair(c, ratio=ratio, min_removal=mn)
        eng.step(12)
        o.step(12)
        assert_parity(g, o)


def test_graph_replay_matches_eager():
    """Every replay of a captured step sequence equals eager stepping, with
    host reads between replays."""
    from mcst import cases

    c = cases.small_case("dual_monod", n_particles=4000)
    eng, g, o = make_pair(c)
    eng.capture(3)
    for _ in range(4):
        eng.replay(1)
        o.step(3)
        assert_parity(g, o)


def test_ring64_full_size():
    """The 64-compartment ring at its full 1e5 particles, 300 steps."""
    from mcst import cases

    c = cases.ring64(n_particles=100_000)
    eng, g, o = make_pair(c)
    for _ in range(3):
        eng.step(100)
        o.step(100)
    assert_parity(g, o)


def test_ring64_all_steps():
    """The 64-compartment ring as specified: 1e5 particles, all 10 000 steps,
    compared every 2500 steps (the oracle on 8 threads)."""
    from mcst import cases

    c = cases.ring64(n_particles=100_000)
    eng, g, o = make_pair(c)
    for _ in range(4):
        eng.step(2500)
        o.step(2500, threads=8)
        assert_parity(g, o)
    cnt = g.counters()
    assert cnt["steps"] == 10_000 and cnt["births"] > 0 and cnt["exits"] > 0 and cnt["compactions"] > 0
    eng.close()


def test_rgg5000_full_size_few_steps():
    """The 5000-compartment network at 1e7 particles: parity over 3 steps
    (the oracle is single-threaded here)."""
    from mcst import cases

    c = cases.rgg5000(n_particles=10_000_000)
    eng, g, o = make_pair(c)
    eng.step(3)
    o.step(3)
    assert_parity(g, o, probes=False)


def test_rgg5000_all_steps():
    """The 5000-compartment network at 1e7 particles as specified: all 1000
    steps (20 flowmaps switched every 100 steps), compared every 250 steps
    (the oracle on 16 threads)."""
    from mcst import cases

    c = cases.rgg5000(n_particles=10_000_000)
    eng, g, o = make_pair(c)
    for _ in range(4):
        eng.step(250)
        o.step(250, threads=16)
        assert_parity(g, o, probes=False)
    cnt = g.counters()
    assert cnt["steps"] == 1000 and cnt["births"] > 0 and cnt["moves"] > 0
    eng.close()


def test_batch_of_runs_equals_runs_alone():
    """A ragged batch: every run of the batch matches its own oracle."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    specs = [("monod_mass", 12, 3000, 1), ("monod_mass", 7, 500, 2), ("monod_mass", 30, 9000, 3),
             ("monod_mass", 1, 1500, 4), ("monod_mass", 20, 0, 5)]
    eng = mcst.Engine("monod_mass")
    pairs = []
    for model, n_c, n, seed in specs:
        c = cases.small_case(model, n_c=n_c, n_particles=n, seed=seed)
        g = eng.add_run(c.n_c, c.dt, c.seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        c.capacity = max(2 * n, 8)
        for r in (g, o):
            c.apply(r)
        g.set_probes(1 << 16)
        pairs.append((g, o))
    eng.commit()
    eng.step(10)
    for g, o in pairs:
        o.step(10)
        assert_parity(g, o)


def test_step_target_stops_a_run():
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    eng = mcst.Engine("monod_mass")
    c = cases.small_case("monod_mass", n_particles=2000)
    a = eng.add_run(c.n_c, c.dt, c.seed)
    b = eng.add_run(c.n_c, c.dt, c.seed)
    for r in (a, b):
        c.apply(r)
        r.set_probes(1 << 16)
    a.set_step_target(4)
    eng.commit()
    eng.step(9)
    oa = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    ob = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    for r in (oa, ob):
        c.apply(r)
    oa.step(4)
    ob.step(9)
    assert a.counters()["steps"] == 4 and b.counters()["steps"] == 9
    assert_parity(a, oa)
    assert_parity(b, ob)


@pytest.mark.parametrize("model", ["dual_monod", "quad_monod"])
def test_steps_in_one_call_equal_steps_one_by_one(model):
    """Within one step call every step but the last transports the next one
    inside its resolve kernel; one call of 13 steps, 13 calls of one step and
    calls of 5, 1 and 7 steps give the same run (events and positions
    bit-exact, FP64 within the parity bar: the contribution atomics may add
    in another order), and all match the oracle.  A run with an outlet, so
    that exits, removals and both record tables are exercised."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    c = cases.small_case(model, n_c=12, n_particles=6000, seed=8)
    c.capacity = 30000
    engines, runs = [], []
    for _ in range(3):
        eng = mcst.Engine(c.model)
        g = eng.add_run(c.n_c, c.dt, c.seed)
        c.apply(g)
        g.set_compaction(0.01, 0)
        g.set_probes(1 << 16)
        eng.commit()
        engines.append(eng)
        runs.append(g)
    engines[0].step(13)
    for _ in range(13):
        engines[1].step(1)
    for k in (5, 1, 7):
        engines[2].step(k)
    o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(o)
    o.set_compaction(0.01, 0)
    o.step(13)
    for g in runs:
        assert g.counters()["steps"] == 13
        assert_parity(g, o)
    cnt = runs[0].counters()
    assert cnt["exits"] > 0 and cnt["births"] > 0 and cnt["moves"] > 0
    for eng in engines:
        eng.close()


def test_crowded_divisions_with_refusals_in_a_batch():
    """Every particle of two runs asks to divide at once (birth mass half the
    mass): every sub-chunk holds 256 requests, so the first 16 of each are
    divided in the particle pass and the rest by the resolve kernel, and the
    runs' free slots accept only part of them, refusing staged and unstaged
    requests alike (the pass's mothers restored).  Both runs, side by side in
    one batch and over three steps (the refused ask again), match the oracle."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    c = cases.small_case("monod_mass", n_c=12, n_particles=100, seed=9)
    eng = mcst.Engine("monod_mass")
    specs = [(3000, 4000, 21), (5000, 5600, 22)]  # (particles, capacity, seed)
    pairs = []
    for n, cap, seed in specs:
        rng = np.random.default_rng(seed)
        pos = rng.integers(0, c.n_c, n).astype(np.uint32)
        m = rng.uniform(1.0e-15, 1.5e-15, n)
        state = np.stack([m, 0.5 * m])
        g = eng.add_run(c.n_c, c.dt, seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, seed)
        for r in (g, o):
            r.set_params(c.params)
            for k, (V, rows, cols, flows) in enumerate(c.maps):
                r.add_flowmap(V, rows, cols, flows)
            r.set_steps_per_map(c.steps_per_map)
            r.set_concentrations(c.C0)
            r.set_particles(cap, pos, np.zeros(n, np.uint8), state, np.zeros(n), np.zeros(n), 1e3)
            r.set_compaction(0.01, 0)
        g.set_probes(1 << 16)
        pairs.append((g, o, n, cap))
    eng.commit()
    for step in range(3):
        eng.step(1)
        for g, o, n, cap in pairs:
            o.step(1)
            assert_parity(g, o)
    for g, o, n, cap in pairs:
        cnt = g.counters()
        assert cnt["births"] == cap - n and cnt["overflow"] > 0, (cnt, n, cap)
    eng.close()


def test_batch_with_and_without_a_leaving_flow():
    """A batch where one run has an outlet and the other none takes the pass
    variant with the leaving-flow work for both (the run without an outlet
    then reads no hydraulic age of its own); a batch of runs without outlets
    takes the variant without it.  Every run matches its oracle."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    for outlets in ((True, False), (False, False)):
        eng = mcst.Engine("dual_monod")
        pairs = []
        for k, feeds in enumerate(outlets):
            c = cases.small_case("dual_monod", n_c=10, n_particles=4000, seed=30 + k, feeds=feeds)
            g = eng.add_run(c.n_c, c.dt, c.seed)
            o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
            for r in (g, o):
                c.apply(r)
            g.set_probes(1 << 16)
            pairs.append((g, o))
        eng.commit()
        eng.step(7)
        for g, o in pairs:
            o.step(7)
            assert_parity(g, o)
        assert (pairs[0][0].counters()["exits"] > 0) == outlets[0]
        eng.close()


def test_glucose_equivalent_mass_balance_rgg5000():
    """Closed dual-substrate system: liquid glucose + pending sources +
    weight * sum(m/Yxs + X) is conserved to 1e-10 relative."""
    from mcst import cases
    import mcst

    c = cases.rgg5000(n_particles=1_000_000, n_maps=3, steps_per_map=2)
    eng = mcst.Engine(c.model)
    g = eng.add_run(c.n_c, c.dt, c.seed)
    c.apply(g)
    eng.commit()
    Yxs = c.params[3]

    def total():
        st = g.particles()["state"]
        return (g.mass()[0].sum() + c.dt * g.contributions()[0].sum()
                + g.weight * (np.sum(st[0]) / Yxs + np.sum(st[2])))

    t0 = total()
    for _ in range(4):
        eng.step(5)
        assert abs(total() - t0) <= 1e-10 * abs(t0)


def test_gas_liquid_transfer_parity():
    """quad_monod on the small network with oxygen supplied through a fixed
    gas phase (kla per compartment varying), against the oracle."""
    import numpy as np

    from mcst import cases

    c = cases.small_case("quad_monod", n_c=12, n_particles=6000, seed=14)
    c.capacity = 60000
    rng = np.random.default_rng(3)
    kla = np.zeros((4, 12))
    kla[1] = rng.uniform(0.5, 5.0, 12)
    cstar = np.zeros((4, 12))
    cstar[1] = 8e-3
    c.gas = (kla, cstar)
    c.C0[1] = 1e-3
    eng, g, o = make_pair(c)
    for _ in range(4):
        eng.step(6)
        o.step(6)
        assert_parity(g, o)
    eng.close()


def test_time_dependent_feeds_parity():
    """Linear and exponential feeds (flows set per step from t = step*dt):
    the first feed of an outlet (its exit threshold follows the flow), a
    second feed of a constant feed's outlet and feeds without an outlet,
    against the oracle over 36 steps.  Capturing a graph is refused
    while such a run is on the device."""
    import mcst
    from mcst import cases

    c = cases.small_case("dual_monod", n_c=12, n_particles=6000, seed=21)
    c.capacity = 60000
    c.feeds = [
        ("constant", 2e-3, 0.0, 0, 6, [0], [10.0]),
        ("linear", 1e-3, 0.02, 3, 9, [0, 1], [8.0, 0.5]),    # first feed of outlet 9
        ("linear", 1e-3, 0.01, 5, 6, [1], [2.0]),           # second feed of outlet 6
        ("linear", 5e-4, -1e-3, 2, None, [0], [3.0]),       # no outlet
        ("exponential", 0.0, -3.0, 10, None, [1], [1e-4]),  # flow e^{-3t} m3/s, no outlet
    ]
    eng, g, o = make_pair(c)
    with pytest.raises(mcst.MCSTError):
        eng.capture(2)
    for _ in range(6):
        eng.step(6)
        o.step(6)
        assert_parity(g, o)
    cnt = g.counters()
    assert cnt["exits"] > 0 and cnt["births"] > 0
    eng.close()


def test_lattice50k_reduced():
    """The 50 000-compartment lattice tank (4 species, fixed-gas oxygen
    supply, 2500 bottom-to-top feeds, quad_monod) at 2e5 particles: parity
    over 24 steps of 5 ms, with exits, divisions and compactions."""
    from mcst import cases

    c = cases.lattice50k(n_particles=200_000)
    c.capacity = 600_000
    eng, g, o = make_pair(c)
    for _ in range(3):
        eng.step(8)
        o.step(8, threads=8)
        assert_parity(g, o, probes=False)
    cnt = g.counters()
    assert cnt["moves"] > 0 and cnt["exits"] > 0 and cnt["births"] > 0 and cnt["compactions"] > 0
    eng.close()


def test_lattice50k_1e7_all_steps():
    """The lattice tank at 1e7 particles (growing to about 2.3e7) for all 200
    steps: the windowed LDS accumulator (the run's 4 x 50 000 contributions
    do not fit in LDS) at a realistic density, compared at steps 100 and 200
    (the oracle on 16 threads)."""
    from mcst import cases

    c = cases.lattice50k(n_particles=10_000_000)
    c.capacity = 30_000_000
    eng, g, o = make_pair(c)
    for _ in range(2):
        eng.step(100)
        o.step(100, threads=16)
        assert_parity(g, o, probes=False)
    cnt = g.counters()
    assert cnt["steps"] == 200 and cnt["exits"] > 0 and cnt["births"] > 0 and cnt["compactions"] > 0
    eng.close()


@pytest.mark.parametrize("model", ["monod_mass", "quad_monod"])
def test_periodic_reorder_parity(model):
    """Reordering by compartment every 5 steps on the device (radix sort of
    the slots) matches the oracle's counting sort at the same steps; an
    explicit reorder too."""
    from mcst import cases

    c = cases.small_case(model, n_c=12, n_particles=6000, seed=31)
    c.capacity = 60000
    eng, g, o = make_pair(c)
    for r in (g, o):
        r.set_sort_interval(5)
    for _ in range(4):
        eng.step(6)
        o.step(6)
        assert_parity(g, o)
    g.sort_particles()
    o.sort_particles()
    assert_parity(g, o)
    eng.step(3)
    o.step(3)
    assert_parity(g, o)
    assert g.counters()["compactions"] > 0 and g.counters()["births"] > 0
    eng.close()


@pytest.mark.parametrize("n_c", [12, 5000, 70000])
def test_reorder_radix_passes(n_c):
    """The device reorder is a radix sort of 8-bit digits: 1, 2 and 3 passes
    for 12, 5000 and 70 000 compartments.  Reordered every 3 steps and once
    explicitly, the run matches the oracle's stable sort, and right after a
    reorder the positions are non-decreasing."""
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=n_c, n_particles=30000, seed=41)
    c.capacity = 90000
    eng, g, o = make_pair(c)
    for r in (g, o):
        r.set_sort_interval(3)
    eng.step(7)
    o.step(7)
    assert_parity(g, o)
    g.sort_particles()
    o.sort_particles()
    assert_parity(g, o)
    pos = g.particles()["pos"]
    assert np.all(np.diff(pos.astype(np.int64)) >= 0)
    eng.close()


def test_compartment_ids_past_the_mover_entry():
    """A list entry of the pass holds a compartment in 21 bits; a larger id is
    marked there and read back from the slot's position (a mover) or from a
    side list (a division request, the newborn's compartment).  On a ring of
    2^22 compartments about half the particles sit at ids of 2^21 - 1 and
    above; over 3 steps (about 8 500 moves, 570 births) the run matches the
    oracle."""
    from mcst import cases

    c = cases.small_case("monod_mass", n_c=1 << 22, n_particles=20000, seed=43)
    eng, g, o = make_pair(c)
    far = (1 << 21) - 1
    assert (g.particles()["pos"] >= far).mean() > 0.4
    eng.step(3)
    o.step(3)
    assert_parity(g, o)
    assert g.counters()["moves"] > 5000 and g.counters()["births"] > 300
    assert (g.particles()["pos"] >= far).mean() > 0.4
    eng.close()


def test_reorder_of_a_run_in_a_batch():
    """The reorder of a run that shares the engine's arrays with another
    gathers through scratch and copies back; once it is the only run left, it
    gathers into a second set of arrays that replaces the first.  Both runs
    match their oracle runs at every check, the unsorted one included."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    eng = mcst.Engine("monod_mass")
    eng.reserve(slots=80_000, runs=2)
    pairs = []
    for n_c, seed in ((12, 51), (40, 52)):
        c = cases.small_case("monod_mass", n_c=n_c, n_particles=8000, seed=seed)
        c.capacity = 30000
        g = eng.add_run(c.n_c, c.dt, c.seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        for r in (g, o):
            c.apply(r)
            r.set_compaction(0.01, 0)
        pairs.append((g, o))
    (g0, o0), (g1, o1) = pairs
    for r in (g0, o0):
        r.set_sort_interval(4)
    eng.commit()
    for _ in range(3):
        eng.step(5)
        for g, o in pairs:
            o.step(5)
            assert_parity(g, o, probes=False)
    g0.sort_particles()
    o0.sort_particles()
    for g, o in pairs:
        assert_parity(g, o, probes=False)
    eng.retire(g1)  # the sorted run is now the engine's only one
    for _ in range(2):
        eng.step(4)
        o0.step(4)
        assert_parity(g0, o0, probes=False)
    g0.sort_particles()
    o0.sort_particles()
    assert_parity(g0, o0, probes=False)
    assert np.all(np.diff(g0.particles()["pos"].astype(np.int64)) >= 0)
    eng.step(3)
    o0.step(3)
    assert_parity(g0, o0, probes=False)
    eng.close()


def test_contribution_rows_when_a_run_ends_inside_a_call(monkeypatch):
    """With contribution rows (MCST_CONTRIB_ROWS=1) the next step's transport
    takes the summed rows and S is written only on a step after which the
    host can read it: the last of a call, or the run's last, here in the
    middle of a call of 20 steps.  The contributions read afterwards are the
    last pass's, as the oracle's."""
    from mcst import cases

    import mcst
    from oracle.oracle import OracleRun

    monkeypatch.setenv("MCST_CONTRIB_ROWS", "1")
    c = cases.small_case("monod_mass", n_c=12, n_particles=6000, seed=61)
    eng = mcst.Engine(c.model)
    g = eng.add_run(c.n_c, c.dt, c.seed)
    o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    for r in (g, o):
        c.apply(r)
        r.set_compaction(0.01, 0)
    g.set_probes(1 << 20)
    g.set_step_target(13)
    eng.commit()
    eng.step(20)
    o.step(13)
    assert g.counters()["steps"] == 13
    assert_parity(g, o)
    eng.close()


def test_lattice50k_batch_of_two_window_runs():
    """Two lattice runs in one engine (their accumulators do not fit LDS, so
    the pass keeps a following window per block and switches windows between
    runs): each matches its oracle run."""
    import mcst
    from mcst import cases
    from oracle.oracle import OracleRun

    eng = mcst.Engine("quad_monod")
    pairs = []
    for n, seed in ((120_000, 3), (60_000, 4)):
        c = cases.lattice50k(n_particles=n, seed=seed)
        c.capacity = 4 * n
        g = eng.add_run(c.n_c, c.dt, c.seed)
        o = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
        for r in (g, o):
            c.apply(r)
        pairs.append((g, o))
    eng.commit()
    eng.step(16)
    for g, o in pairs:
        o.step(16, threads=8)
        assert_parity(g, o, probes=False)
        assert g.counters()["births"] > 0
    eng.close()


@pytest.mark.parametrize("env", [{"MCST_PARTICLE_BLOCK": "512"}, {"MCST_PARTICLE_BLOCK": "1024"},
                                 {"MCST_CONTRIB_LDS": "0"}, {"MCST_TILE_GROUP": "0"}, {"MCST_TILE_GROUP": "4"},
                                 {"MCST_CONTRIB_ROWS": "1"}])
def test_forced_launch_variants(env, monkeypatch):
    """Every particle-pass variant the autotuner or the overrides can select
    (block size, global-atomic contributions, tile grouping, block
    accumulators stored as rows and summed instead of added to S with
    atomics) gives the same
    events as the oracle: on the small network (whole-run LDS accumulator)
    and on the lattice (compartment window)."""
    from mcst import cases

    for k, v in env.items():
        monkeypatch.setenv(k, v)
    c = cases.small_case("dual_monod", n_particles=6000, seed=4)
    eng, g, o = make_pair(c)
    eng.step(12)
    o.step(12)
    assert_parity(g, o)
    eng.close()
    c = cases.lattice50k(n_particles=100_000)
    c.capacity = 300_000
    eng, g, o = make_pair(c)
    eng.step(8)
    o.step(8, threads=8)
    assert_parity(g, o, probes=False)
    eng.close()
