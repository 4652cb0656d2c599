"""Admission logic of the request loop, on the CPU with a stand-in engine.

The stand-in mimics the engine's serving calls: a slot pool in tiles, runs
admitted in id order while they fit, a run inactive once it reaches its step
target, and retire returning its slots.  The GPU tests check the real
engine's results against runs alone.
"""
import numpy as np

from mcst import cases
from mcst.serve import Request, Server, footprint, percentile

TILE = 1024


class FakeRun:
    def __init__(self, eng, rid, n_c):
        self.e, self.id, self.n_c = eng, rid, n_c
        self.target = None
        self.n = 0
        self.cap = 0
        self.steps = 0
        self.state = "defined"

    def set_params(self, p):
        pass

    def add_flowmap(self, *a):
        pass

    def set_steps_per_map(self, k):
        pass

    def add_feed(self, *a):
        pass

    def set_concentrations(self, c):
        pass

    def init_particles(self, n, cap, x0):
        self.n, self.cap = n, cap

    def set_step_target(self, s):
        self.target = s

    def counters(self):
        assert self.state == "live"
        return {"steps": self.steps, "processed": self.steps * self.n}


class FakeEngine:
    n_var = 2

    def __init__(self, tiles):
        self.free = tiles
        self.runs = []
        self.next_admit = 0
        self.log = []

    def add_run(self, n_c, dt, seed):
        r = FakeRun(self, len(self.runs), n_c)
        self.runs.append(r)
        return r

    def admit(self):
        k = 0
        while self.next_admit < len(self.runs):
            r = self.runs[self.next_admit]
            need = -(-r.cap // TILE)
            if need > self.free:
                break
            self.free -= need
            r.state = "live"
            self.log.append(("admit", r.id))
            self.next_admit += 1
            k += 1
        return k

    def step(self, n):
        for r in self.runs:
            if r.state == "live":
                r.steps = min(r.target, r.steps + n)

    def synchronize(self):
        pass

    def run_steps(self):
        return [r.steps if r.state == "live" else 0 for r in self.runs]

    def retire(self, r):
        assert r.state == "live" and r.steps == r.target
        r.state = "retired"
        self.free += -(-r.cap // TILE)
        self.log.append(("retire", r.id))


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t

    def sleep(self, d):
        self.t += max(d, 0.001)


def make_requests(n, steps):
    reqs = []
    for i in range(n):
        c = cases.ensemble_run(i, seed=5, n_steps=steps[i], n_c_range=(8, 16), n_range=(3000, 3000))
        reqs.append(Request(i, 0.01 * i, c, steps[i]))
    return reqs


def test_footprint_counts_slots_maps_and_liquid():
    c = cases.ensemble_run(0, seed=5, n_c_range=(50, 50), n_range=(10000, 10000))
    fp = footprint(c, 2)
    slots = -(-2 * 10000 // TILE) * TILE
    assert fp > slots * (8 * 2 + 21)
    c2 = cases.ensemble_run(0, seed=5, n_c_range=(50, 50), n_range=(40000, 40000))
    assert footprint(c2, 2) > fp


def test_every_request_completes_and_pool_is_reused():
    steps = [40, 10, 25, 5, 30, 15]
    reqs = make_requests(6, steps)
    eng = FakeEngine(tiles=12)  # each request needs 6 tiles: two fit at once
    clk = Clock()
    srv = Server("monod_mass", 0, 0, budget_bytes=10**12, chunk_steps=5, engine=eng)
    res = srv.serve(reqs, clock=clk, sleep=clk.sleep)
    assert [r.rid for r in res] == list(range(6))
    assert [r.steps for r in res] == steps
    assert all(r.particle_steps == r.steps * 3000 for r in res)
    assert all(r.finished >= r.admitted >= r.arrival for r in res)
    # admission in arrival order, never more than two resident
    admits = [i for a, i in eng.log if a == "admit"]
    assert admits == sorted(admits)
    live = 0
    for a, _ in eng.log:
        live += 1 if a == "admit" else -1
        assert live <= 2
    assert eng.free == 12


def test_budget_limits_residents():
    reqs = make_requests(4, [20, 20, 20, 20])
    fp = footprint(reqs[0].case, 2)
    eng = FakeEngine(tiles=1000)
    clk = Clock()
    srv = Server("monod_mass", 0, 0, budget_bytes=int(1.5 * fp), chunk_steps=5, engine=eng)
    srv.serve(reqs, clock=clk, sleep=clk.sleep)
    live = 0
    for a, _ in eng.log:
        live += 1 if a == "admit" else -1
        assert live <= 1


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert percentile(v, 50) == 50
    assert percentile(v, 99) == 99
    assert percentile([3.0], 99) == 3.0
