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
    """Device bytes a run holds while resident: its slot range (state, the
    two base ages and their clock origins, position, status, the removal's
    filler list and the pass's mover and division lists), its flowmaps and
    liquid arrays."""
    cap = case.capacity or 2 * case.n_particles
    slots = -(-cap // TILE) * TILE
    per_slot = 8 * n_var + 2 * 8 + 2 * 4 + 4 + 1 + 4 + 3 * 4
    maps = 0
    for V, rows, cols, flows in case.maps:
        deg = max(1, int(max(1, (len(flows) + case.n_c - 1) // case.n_c)))
        maps += case.n_c * (8 * 4 + 12 * 2 * deg) + 20 * (len(flows) + case.n_c)
    liquid = 8 * 6 * case.n_s * case.n_c
    return slots * per_slot + maps + liquid


class Server:
    """Continuous batching with a reserved slot pool: one engine per model
    (an engine batches runs of one model), all stepping on the same GPU on
    their own streams.

    Server(model, slots, max_runs, ...) serves one model; `pools` maps more
    models to their (slots, max_runs).  A request's model is its case's."""

    def __init__(self, model=None, slots=0, max_runs=0, budget_bytes=200 * 10**9, chunk_steps=20, device=0,
                 engine=None, pools=None):
        self.engines = {}
        self.n_var = {}
        if engine is not None:
            self.engines[model] = engine
            self.n_var[model] = engine.n_var
        pools = dict(pools or {})
        if model is not None and engine is None:
            pools.setdefault(model, (slots, max_runs))
        for m, (sl, mr) in pools.items():
            if m in self.engines:
                continue
            _, self.n_var[m], _, _ = model_info(m)
            e = Engine(m, device)
            e.reserve(sl, mr)
            e.commit()
            self.engines[m] = e
        self.model = model if model is not None else next(iter(self.engines))
        self.eng = self.engines[self.model]
        self.budget = budget_bytes
        self.chunk = chunk_steps

    def close(self):
        for e in self.engines.values():
            e.close()

    def _engine(self, req):
        m = getattr(req.case, "model", self.model)
        if m not in self.engines:
            raise KeyError(f"request {req.rid}: no engine serves model {m!r}")
        return m, self.engines[m]

    def serve(self, requests, on_finish=None, clock=time.perf_counter, sleep=time.sleep):
        """Serve `requests` (in arrival order); returns one Result each.
        `on_finish(request, run)` may read a run's results before it leaves.

        Each turn launches a chunk of steps on every engine with live runs,
        defines the requests that have arrived (building their particles on
        the host) while the GPU steps, then retires the runs that are done
        and admits waiting ones, which join at the next chunk."""
        queue = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()   # arrived, not yet defined on an engine
        defined = {m: collections.deque() for m in self.engines}  # defined, waiting for room in the pool
        live = {}                       # (model, run id) -> (request, run, admitted time, bytes)
        used = 0
        results = []
        t0 = clock()
        while queue or waiting or any(defined.values()) or live:
            stepping = {key[0] for key in live}
            for m in stepping:  # launches are asynchronous: the engines' streams overlap
                self.engines[m].step(self.chunk)
            now = clock() - t0
            while queue and queue[0].arrival <= now:
                waiting.append(queue.popleft())
            while waiting:  # define arrivals while the footprint budget allows
                req = waiting[0]
                m, eng = self._engine(req)
                fp = footprint(req.case, self.n_var[m])
                if (live or any(defined.values())) and used + fp > self.budget:
                    break
                waiting.popleft()
                run = eng.add_run(req.case.n_c, req.case.dt, req.case.seed)
                req.case.apply(run)
                run.set_step_target(req.steps)
                defined[m].append((req, run, fp))
                used += fp
            for m in stepping:
                self.engines[m].synchronize()
            if stepping:
                t_now = clock() - t0
                steps = {m: self.engines[m].run_steps() for m in stepping}  # one device read per engine
                for key in list(live):
                    req, run, t_adm, fp = live[key]
                    if steps[key[0]][run.id] < req.steps:
                        continue
                    if on_finish is not None:
                        on_finish(req, run)
                    c = run.counters()
                    results.append(Result(req.rid, req.arrival, t_adm, t_now, int(c["steps"]), int(c["processed"])))
                    self.engines[key[0]].retire(run)
                    used -= fp
                    del live[key]
            for m, dq in defined.items():
                if not dq:
                    continue
                k = self.engines[m].admit()
                t_adm = clock() - t0
                for _ in range(k):
                    req, run, fp = dq.popleft()
                    live[(m, run.id)] = (req, run, t_adm, fp)
                if dq and not any(key[0] == m for key in live):
                    raise RuntimeError(f"request {dq[0][0].rid} does not fit the engine's pool")
            if not live and queue and not waiting:
                sleep(max(0.0, min(0.05, queue[0].arrival - (clock() - t0))))
        return sorted(results, key=lambda r: r.rid)


class OneAtATime:
    """The same requests served one at a time: one engine, created and
    committed beforehand with a slot pool (reserve/admit/retire, as the
    Server's engines are), on which each request is defined, admitted,
    stepped to its step count alone and retired before the next one starts.
    Engine creation is outside the timing here as in Server.serve."""

    def __init__(self, model, slots, max_runs, device=0):
        self.eng = Engine(model, device)
        self.eng.reserve(slots, max_runs)
        self.eng.commit()

    def close(self):
        self.eng.close()

    def run(self, req):
        """One request on the idle engine; returns (seconds, particle-steps)."""
        t0 = time.perf_counter()
        run = self.eng.add_run(req.case.n_c, req.case.dt, req.case.seed)
        req.case.apply(run)
        run.set_step_target(req.steps)
        if self.eng.admit() != 1:
            raise RuntimeError(f"request {req.rid} does not fit the engine's pool")
        self.eng.step(req.steps)
        self.eng.synchronize()
        c = run.counters()
        self.eng.retire(run)
        return time.perf_counter() - t0, int(c["processed"])


def percentile(values, q):
    v = sorted(values)
    if not v:
        return math.nan
    k = min(len(v) - 1, max(0, int(math.ceil(q / 100.0 * len(v))) - 1))
    return v[k]
