"""Serving: a request loop that keeps simulations resident on one GPU.

Requests (a Case with its step count, arriving at given times) are admitted
in arrival order while the resident runs' device footprint stays under a
budget (particles x state bytes + flowmaps + liquid).  Admitted runs join the
engine's batch at the next step; every step advances all resident runs
together; a run leaves the batch when it reaches its step count and its
slots return to the pool for the next requests.  So requests that arrive
while others run are folded into the running batch as room frees up
(continuous batching), not queued behind them.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time

from . import Engine, model_info

TILE = 1024


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float       # seconds after the start of serving
    case: object         # cases.Case
    steps: int


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    particle_steps: int

    @property
    def latency(self):
        return self.finished - self.arrival


def footprint(case, n_var):
    """Device bytes a run holds while resident: its slot range (state, ages,
    position, status and the pass's lists), its flowmaps and liquid arrays."""
    cap = case.capacity or 2 * case.n_particles
    slots = -(-cap // TILE) * TILE
    per_slot = 8 * n_var + 8 + 8 + 4 + 1 + 4 + 12
    maps = 0
    for V, rows, cols, flows in case.maps:
        deg = max(1, int(max(1, (len(flows) + case.n_c - 1) // case.n_c)))
        maps += case.n_c * (8 * 4 + 12 * 2 * deg) + 20 * (len(flows) + case.n_c)
    liquid = 8 * 6 * case.n_s * case.n_c
    return slots * per_slot + maps + liquid


class Server:
    """Continuous batching over one engine with a reserved slot pool."""

    def __init__(self, model, slots, max_runs, budget_bytes, chunk_steps=20, device=0, engine=None):
        self.model = model
        if engine is None:
            _, self.n_var, _, _ = model_info(model)
            engine = Engine(model, device)
            engine.reserve(slots, max_runs)
            engine.commit()
        else:
            self.n_var = engine.n_var
        self.eng = engine
        self.budget = budget_bytes
        self.chunk = chunk_steps

    def close(self):
        self.eng.close()

    def serve(self, requests, on_finish=None, clock=time.perf_counter, sleep=time.sleep):
        """Serve `requests` (in arrival order); returns one Result each.
        `on_finish(request, run)` may read a run's results before it leaves."""
        eng = self.eng
        queue = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()   # arrived, not yet defined on the engine
        defined = collections.deque()   # defined, waiting for room in the pool
        live = {}                       # run id -> (request, run, admitted time, bytes)
        used = 0
        results = []
        t0 = clock()
        while queue or waiting or defined or live:
            now = clock() - t0
            while queue and queue[0].arrival <= now:
                waiting.append(queue.popleft())
            while waiting:  # define arrivals while the footprint budget allows
                req = waiting[0]
                fp = footprint(req.case, self.n_var)
                if live and used + fp > self.budget:
                    break
                waiting.popleft()
                run = eng.add_run(req.case.n_c, req.case.dt, req.case.seed)
                req.case.apply(run)
                run.set_step_target(req.steps)
                defined.append((req, run, fp))
                used += fp
            if defined:
                k = eng.admit()
                t_adm = clock() - t0
                for _ in range(k):
                    req, run, fp = defined.popleft()
                    live[run.id] = (req, run, t_adm, fp)
                if defined and not live:
                    raise RuntimeError(f"request {defined[0][0].rid} does not fit the engine's pool")
            if not live:
                if queue:
                    sleep(max(0.0, min(0.05, queue[0].arrival - (clock() - t0))))
                continue
            eng.step(self.chunk)
            eng.synchronize()
            t_now = clock() - t0
            for rid in list(live):
                req, run, t_adm, fp = live[rid]
                c = run.counters()
                if c["steps"] < req.steps:
                    continue
                if on_finish is not None:
                    on_finish(req, run)
                results.append(Result(req.rid, req.arrival, t_adm, t_now, int(c["steps"]), int(c["processed"])))
                eng.retire(run)
                used -= fp
                del live[rid]
        return sorted(results, key=lambda r: r.rid)


def run_alone(req, model, device=0):
    """The same request on its own engine, stepped to completion; returns
    (seconds, particle-steps)."""
    t0 = time.perf_counter()
    eng = Engine(model, device)
    run = eng.add_run(req.case.n_c, req.case.dt, req.case.seed)
    req.case.apply(run)
    eng.commit()
    eng.step(req.steps)
    eng.synchronize()
    c = run.counters()
    eng.close()
    return time.perf_counter() - t0, int(c["processed"])


def percentile(values, q):
    v = sorted(values)
    if not v:
        return math.nan
    k = min(len(v) - 1, max(0, int(math.ceil(q / 100.0 * len(v))) - 1))
    return v[k]
