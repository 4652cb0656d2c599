"""Synthetic compartment networks and run set-ups for the benchmark cases.

Every generator is deterministic in its seed and returns a `Case`: raw
volumes and COO flow matrices (one per flowmap), feeds, initial liquid
concentrations and particle settings.  `Case.apply(run)` configures any object
with the run API (the HIP engine's `Run`, or the CPU oracle's run), so the same
inputs reach both.  Units are SI: m3, m3/s, kg/m3 (= g/L), kg, s.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

L = 1e-3  # one litre in m3


@dataclasses.dataclass
class Case:
    name: str
    model: str
    n_c: int
    n_s: int
    dt: float
    seed: int
    maps: list                      # [(V[n_c], rows, cols, flows)]
    steps_per_map: int = 1
    feeds: list = dataclasses.field(default_factory=list)  # (kind, f0, param, in, out|None, species, conc)
    C0: np.ndarray = None           # [n_s][n_c]
    params: np.ndarray = None
    n_particles: int = 0
    capacity: int = 0
    x0: float = 1.0
    n_steps: int = 1
    gas: tuple = None               # (kla[n_s], C*[n_s]) gas-liquid transfer, or None
    # two-phase run: {"henry": [n_s], "c_gas": [n_s][n_c], "kla": ("fixed", [n_s]) or
    # ("turbulence", species, d_b), "gas_maps": [(Vg, rows, cols, flows, eps)] (one per
    # flowmap), "gas_feeds": [(kind, f0, param, in, out|None, species, conc)]}
    two_phase: dict = None

    def apply(self, run, particles=True):
        if self.params is not None:
            run.set_params(self.params)
        for V, r, c, q in self.maps:
            run.add_flowmap(V, r, c, q)
        run.set_steps_per_map(self.steps_per_map)
        for f in self.feeds:
            run.add_feed(*f)
        run.set_concentrations(self.C0)
        if self.gas is not None:
            run.set_gas_transfer(*self.gas)
        if self.two_phase is not None:
            tp = self.two_phase
            run.set_gas_phase(tp["henry"], tp["c_gas"])
            if tp["kla"][0] == "fixed":
                run.set_kla("fixed", tp["kla"][1])
            else:
                run.set_kla("turbulence", species=tp["kla"][1], bubble_diameter=tp["kla"][2])
            for Vg, r, c, q, eps in tp["gas_maps"]:
                run.add_gas_flowmap(Vg, r, c, q, eps)
            for f in tp.get("gas_feeds", []):
                run.add_gas_feed(*f)
        if particles:
            run.init_particles(self.n_particles, self.capacity or 2 * self.n_particles, self.x0)
        return run

    @property
    def nnz(self):
        return len(self.maps[0][3])


def _symmetric(edges, q):
    """Bidirectional exchange: F(i,j) = F(j,i) = q_edge."""
    i, j = edges[:, 0].astype(np.uint32), edges[:, 1].astype(np.uint32)
    return np.concatenate([i, j]), np.concatenate([j, i]), np.concatenate([q, q])


def ring_with_chords(n=64, n_chords=16, rng=None):
    rng = rng or np.random.default_rng(0)
    ring = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1)
    chords = set()
    while len(chords) < n_chords:
        a, b = rng.integers(0, n, 2)
        if a == b or abs(a - b) in (1, n - 1):
            continue
        chords.add((min(a, b), max(a, b)))
    return np.concatenate([ring, np.array(sorted(chords))])


def geometric_graph_cylinder(n, mean_degree=6.0, rng=None, radius=1.0, height=1.0):
    """Random geometric graph in a cylinder: the n*mean_degree/2 closest pairs."""
    from scipy.spatial import cKDTree

    rng = rng or np.random.default_rng(0)
    r = radius * np.sqrt(rng.random(n))
    th = 2 * math.pi * rng.random(n)
    pts = np.stack([r * np.cos(th), r * np.sin(th), height * rng.random(n)], axis=1)
    n_edges = int(round(n * mean_degree / 2))
    tree = cKDTree(pts)
    vol = math.pi * radius * radius * height
    d = (3 * vol * mean_degree / (4 * math.pi * n)) ** (1 / 3)
    while True:
        pairs = tree.query_pairs(d, output_type="ndarray")
        if len(pairs) >= n_edges or d > 10:
            break
        d *= 1.25
    dist = np.linalg.norm(pts[pairs[:, 0]] - pts[pairs[:, 1]], axis=1)
    keep = np.argsort(dist, kind="stable")[:n_edges]
    e = pairs[keep]
    return e[np.lexsort((e[:, 1], e[:, 0]))], pts


def ring64(n_particles=100_000, n_steps=10_000, seed=1):
    """64-compartment ring + 16 chords, 1 species, mass-based Monod, static flowmap."""
    rng = np.random.default_rng(seed)
    n = 64
    edges = ring_with_chords(n, 16, rng)
    V = rng.uniform(0.5, 2.0, n) * L
    q = np.exp(rng.uniform(math.log(1e-5), math.log(1e-3), len(edges)))
    rows, cols, flows = _symmetric(edges, q)
    C0 = np.full((1, n), 5.0)
    p = np.array([0.5 / 3600.0, 0.01, 0.5, 2.0, 1e-15, 0.1])
    return Case("ring64", "monod_mass", n, 1, 0.01, seed, [(V, rows, cols, flows)],
                feeds=[("constant", 1e-5, 0.0, 0, 32, [0], [10.0])], C0=C0, params=p,
                n_particles=n_particles, x0=1.0, n_steps=n_steps)


def rgg5000(n_particles=10_000_000, n_steps=1000, seed=7, n_c=5000, n_maps=20, steps_per_map=100):
    """5000-compartment geometric graph, 20 cyclic flowmaps, 2 species, dual-substrate Monod."""
    rng = np.random.default_rng(seed)
    edges, _ = geometric_graph_cylinder(n_c, 6.0, rng)
    V = rng.uniform(0.1, 1.0, n_c) * L
    maps = []
    for _ in range(n_maps):
        q = rng.lognormal(math.log(1e-4), 1.0, len(edges))
        rows, cols, flows = _symmetric(edges, q)
        maps.append((V, rows, cols, flows))
    C0 = np.stack([np.full(n_c, 5.0), np.full(n_c, 8e-3)])
    p = np.array([2.8e-4, 0.01, 1e-4, 0.5, 1.0, 100.0, 1e-4, 2e-15, 1.4e-15, 0.2e-15])
    return Case("rgg5000", "dual_monod", n_c, 2, 0.01, seed, maps, steps_per_map=steps_per_map,
                C0=C0, params=p, n_particles=n_particles, x0=1.0, n_steps=n_steps)


def lattice_tank(dims=(50, 50, 20), loop_flow=1e-3, exchange=(1e-6, 1e-4), rng=None):
    """A stirred-tank-like lattice: compartment (x, y, z) is z + NZ*(y + NY*x)
    (vertical index fastest, so vertical neighbours are adjacent ids).
    Recirculation: every column and its outward neighbour column (one cell
    further from the axis) carry two closed loops of `loop_flow` m3/s: out
    at impeller height, down the outer column, in at the bottom, up the
    inner column (lower loop), and out at impeller height, up the outer
    column, in at the top, down the inner column (upper loop).  Loops are
    closed, so every compartment's inflow equals its outflow.  On top, every
    lattice edge exchanges a log-uniform flow in both directions.
    Returns (rows, cols, flows) summed per directed pair."""
    rng = rng or np.random.default_rng(0)
    NX, NY, NZ = dims
    n = NX * NY * NZ

    def cid(x, y, z):
        return z + NZ * (y + NY * x)

    pairs = {}

    def add(a, b, q):
        pairs[(a, b)] = pairs.get((a, b), 0.0) + q

    cx, cy, zm = (NX - 1) / 2.0, (NY - 1) / 2.0, NZ // 2
    for x in range(NX):
        for y in range(NY):
            dx, dy = x - cx, y - cy
            if abs(dx) >= abs(dy):
                ox, oy = x + (1 if dx > 0 else -1), y
            else:
                ox, oy = x, y + (1 if dy > 0 else -1)
            if not (0 <= ox < NX and 0 <= oy < NY):
                continue
            inner = [cid(x, y, z) for z in range(NZ)]
            outer = [cid(ox, oy, z) for z in range(NZ)]
            # lower loop: out at zm-1, down the outer column, in at 0, up the inner column
            add(inner[zm - 1], outer[zm - 1], loop_flow)
            for z in range(zm - 1, 0, -1):
                add(outer[z], outer[z - 1], loop_flow)
            add(outer[0], inner[0], loop_flow)
            for z in range(0, zm - 1):
                add(inner[z], inner[z + 1], loop_flow)
            # upper loop: out at zm, up the outer column, in at the top, down the inner column
            add(inner[zm], outer[zm], loop_flow)
            for z in range(zm, NZ - 1):
                add(outer[z], outer[z + 1], loop_flow)
            add(outer[NZ - 1], inner[NZ - 1], loop_flow)
            for z in range(NZ - 1, zm, -1):
                add(inner[z], inner[z - 1], loop_flow)
    keys = np.array(list(pairs.keys()), dtype=np.int64).reshape(-1, 2)
    vals = np.array(list(pairs.values()))
    # exchanges on every lattice edge
    g = np.arange(n).reshape(NX, NY, NZ)
    e = np.concatenate([np.stack([g[:-1].ravel(), g[1:].ravel()], 1), np.stack([g[:, :-1].ravel(), g[:, 1:].ravel()], 1),
                        np.stack([g[:, :, :-1].ravel(), g[:, :, 1:].ravel()], 1)])
    q = np.exp(rng.uniform(math.log(exchange[0]), math.log(exchange[1]), len(e)))
    rows = np.concatenate([keys[:, 0], e[:, 0], e[:, 1]])
    cols = np.concatenate([keys[:, 1], e[:, 1], e[:, 0]])
    flows = np.concatenate([vals, q, q])
    # sum duplicates (a loop pair that is also an exchange pair)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key, rows, cols, flows = key[order], rows[order], cols[order], flows[order]
    first = np.concatenate([[True], key[1:] != key[:-1]])
    idx = np.cumsum(first) - 1
    summed = np.zeros(first.sum())
    np.add.at(summed, idx, flows)
    return rows[first].astype(np.uint32), cols[first].astype(np.uint32), summed


QUAD_MONOD_LATTICE = np.array([2.0, 0.01, 1e-4, 0.01, 0.5, 0.5, 0.3, 0.3, 0.4, 0.3, 0.2, 0.4, 2e-15, 1.5e-15, 0.2e-15])


def lattice50k(n_particles=100_000_000, n_steps=200, seed=17, dims=(50, 50, 20), feed_flow=2e-3, kla_o2=1.0):
    """50x50x20 lattice tank (50 000 compartments, 1 L each +-20 %), loop flows
    1e-3 m3/s plus log-uniform 1e-6..1e-4 m3/s exchanges; feeds enter through
    the bottom layer and leave through the top layer (one feed per column,
    `feed_flow` m3/s each); 4 species (glucose 5 g/L, O2 8 mg/L with a fixed
    gas phase supplying it at kla = `kla_o2` 1/s, acetate, CO2); quad_monod
    (8 FP64 states) growing about two-fold over the 200 steps of 5 ms while
    the outlet washes particles out."""
    rng = np.random.default_rng(seed)
    NX, NY, NZ = dims
    n = NX * NY * NZ
    rows, cols, flows = lattice_tank(dims, rng=rng)
    V = rng.uniform(0.8, 1.2, n) * L
    feeds = []
    for x in range(NX):
        for y in range(NY):
            bottom = 0 + NZ * (y + NY * x)
            top = NZ - 1 + NZ * (y + NY * x)
            feeds.append(("constant", feed_flow, 0.0, bottom, top, [0, 1], [10.0, 8e-3]))
    C0 = np.stack([np.full(n, 5.0), np.full(n, 8e-3), np.zeros(n), np.zeros(n)])
    gas = (np.array([0.0, kla_o2, 0.0, 0.0]), np.array([0.0, 8e-3, 0.0, 0.0]))
    return Case("lattice50k", "quad_monod", n, 4, 0.005, seed, [(V, rows, cols, flows)], feeds=feeds,
                C0=C0, params=QUAD_MONOD_LATTICE.copy(), n_particles=n_particles, capacity=int(2.5 * n_particles),
                x0=0.005, n_steps=n_steps, gas=gas)


def small_case(model="monod_mass", n_c=12, n_particles=5000, seed=3, n_maps=2, steps_per_map=3,
               feeds=True, fast=True):
    """A small network for parity tests: chords + feed/outlet, fast growth so
    that divisions, moves and exits all happen within a few steps."""
    rng = np.random.default_rng(seed)
    n_s = {"monod_mass": 1, "dual_monod": 2, "monod": 1, "fixed_length": 1, "simple_acetate": 2,
           "quad_monod": 4, "udf_minimal": 1, "udf_starvation": 1}[model]
    if n_c > 1:
        edges = ring_with_chords(n_c, min(4, max(0, n_c * (n_c - 3) // 2)), rng) if n_c > 3 else \
            np.array([[i, (i + 1) % n_c] for i in range(n_c)])
        edges = np.unique(np.sort(edges, axis=1), axis=0)
        edges = edges[edges[:, 0] != edges[:, 1]]
    else:
        edges = np.zeros((0, 2), np.int64)
    V = rng.uniform(0.5, 2.0, n_c) * L
    maps = []
    for _ in range(n_maps):
        q = np.exp(rng.uniform(math.log(2e-3), math.log(2e-2), len(edges)))
        rows, cols, flows = _symmetric(edges, q) if len(edges) else (np.zeros(0, np.uint32),) * 2 + (np.zeros(0),)
        maps.append((V, rows, cols, flows))
    C0 = np.stack([np.full(n_c, 5.0 if s == 0 else 1.0) for s in range(n_s)])
    fd = []
    if feeds:
        out = n_c // 2
        fd.append(("constant", 2e-3, 0.0, 0, out, [0], [10.0]))
    if model == "monod_mass":
        p = np.array([0.5, 0.01, 0.5, 2.0, 1e-15, 0.1]) if fast else None
    elif model == "dual_monod":
        p = np.array([0.3, 0.01, 1e-4, 0.5, 1.0, 10.0, 0.05, 2e-15, 1.9e-15, 0.05e-15]) if fast else None
    elif model == "monod":
        p = np.array([0.3, 1e-3, 2.0, 2e-6, 0.6e-6, 3.0]) if fast else None
    elif model == "fixed_length":
        p = np.array([2e-7, 2e-6, 1e-3, 0.6e-6]) if fast else None
    elif model == "udf_minimal":  # the example model plugins
        p = np.array([0.5, 2e-7, 2e-6, 1e-3, 0.6e-6]) if fast else None
    elif model == "udf_starvation":
        p = np.array([0.5, 1.0, 0.5, 2.0, 10.0, 1e-15, 0.1]) if fast else None
    elif model == "quad_monod":
        p = np.array([20.0, 0.01, 1e-4, 0.01, 5.0, 0.5, 0.3, 0.3, 0.4, 0.3, 0.05, 0.4, 2e-15, 1.4e-15,
                      0.2e-15]) if fast else None
    else:
        p = None
    return Case(f"small_{model}", model, n_c, n_s, 0.01, seed, maps, steps_per_map=steps_per_map, feeds=fd,
                C0=C0, params=p, n_particles=n_particles, x0=1.0)


MONOD_MASS_RING64 = np.array([0.5 / 3600.0, 0.01, 0.5, 2.0, 1e-15, 0.1])


def bubble_column(n_particles=20000, levels=10, columns=3, seed=4, n_maps=2, steps_per_map=5, kla="turbulence",
                  model="dual_monod", holdup=0.05, air_o2=0.28, dt=0.01):
    """A two-phase bubble column: `columns` side-by-side stacks of `levels`
    compartments (compartment c*levels + z).  Liquid: an up-flow loop in the
    first stack and down-flow in the others, exchanges between neighbouring
    compartments, a glucose feed at the bottom of the first stack and an
    outlet at the top of the last.  Gas: air (oxygen `air_o2` kg/m3) fed at
    the bottom of every stack and rising level by level to a gas outlet at
    the top; gas volume `holdup` of the compartment; energy dissipation
    highest at the sparger level.  Species: glucose (0) and oxygen (1, Henry
    constant 3.181e-2 as the reference's mass transfer model sets,
    mass_transfer.cpp:125).  kLa from the turbulence correlation for oxygen
    (or fixed per species with kla="fixed")."""
    rng = np.random.default_rng(seed)
    n = levels * columns

    def cid(c, z):
        return c * levels + z

    V = rng.uniform(0.8, 1.2, n) * L
    Vg = V * holdup / (1.0 - holdup)
    maps, gas_maps = [], []
    for _ in range(n_maps):
        pairs = {}

        def add(a, b, q):
            pairs[(a, b)] = pairs.get((a, b), 0.0) + q

        loop = rng.uniform(2e-5, 5e-5)
        for z in range(levels - 1):  # up the first stack, down the others
            add(cid(0, z), cid(0, z + 1), loop)
            for c in range(1, columns):
                add(cid(c, z + 1), cid(c, z), loop / (columns - 1))
        for c in range(1, columns):  # closing the loop at the ends
            add(cid(0, levels - 1), cid(c, levels - 1), loop / (columns - 1))
            add(cid(c, 0), cid(0, 0), loop / (columns - 1))
        for c in range(columns):
            for z in range(levels):
                if c + 1 < columns:
                    q = rng.uniform(1e-6, 1e-5)
                    add(cid(c, z), cid(c + 1, z), q)
                    add(cid(c + 1, z), cid(c, z), q)
        keys = sorted(pairs)
        maps.append((V, np.array([k[0] for k in keys], np.uint32), np.array([k[1] for k in keys], np.uint32),
                     np.array([pairs[k] for k in keys])))
        gq = rng.uniform(1e-5, 2e-5)  # gas rising through each stack
        g_rows, g_cols, g_flows = [], [], []
        for c in range(columns):
            for z in range(levels - 1):
                g_rows.append(cid(c, z))
                g_cols.append(cid(c, z + 1))
                g_flows.append(gq * (1.0 + 0.1 * z))
        eps = np.array([2.0 if z == 0 else 0.2 + 0.05 * rng.random() for c in range(columns) for z in range(levels)])
        gas_maps.append((Vg, np.array(g_rows, np.uint32), np.array(g_cols, np.uint32), np.array(g_flows), eps))
    C0 = np.stack([np.full(n, 5.0), np.full(n, 8e-3)])
    feeds = [("constant", 1e-6, 0.0, cid(0, 0), cid(columns - 1, levels - 1), [0], [10.0])]
    gas_feeds = [("constant", 2e-5, 0.0, cid(c, 0), cid(c, levels - 1), [1], [air_o2]) for c in range(columns)]
    tp = {"henry": [0.0, 3.181e-2], "c_gas": np.stack([np.zeros(n), np.full(n, air_o2)]),
          "kla": ("turbulence", 1, 5e-3) if kla == "turbulence" else ("fixed", [0.0, 0.05]),
          "gas_maps": gas_maps, "gas_feeds": gas_feeds}
    p = np.array([2.8e-3, 0.01, 1e-4, 0.5, 1.0, 10.0, 1e-3, 2e-15, 1.9e-15, 0.05e-15])
    return Case("bubble_column", model, n, 2, dt, seed, maps, steps_per_map=steps_per_map, feeds=feeds, C0=C0,
                params=p, n_particles=n_particles, x0=5e-3, n_steps=200, two_phase=tp)


def ensemble_run(i, seed=11, n_steps=500, n_c_range=(32, 4000), n_range=(1e4, 2e6)):
    """Run i of the batched ensemble: a geometric-graph network as in rgg5000
    with a log-uniform compartment count, a log-uniform particle count, its own
    seed and Monod parameters (mu_max, K_s, yield) drawn uniform +-30 % around
    ring64's.  One species (glucose 5 g/L), one static flowmap, no feed."""
    rng = np.random.default_rng([seed, i])
    n_c = int(round(math.exp(rng.uniform(math.log(n_c_range[0]), math.log(n_c_range[1])))))
    n_p = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    edges, _ = geometric_graph_cylinder(n_c, 6.0, rng)
    V = rng.uniform(0.1, 1.0, n_c) * L
    q = rng.lognormal(math.log(1e-4), 1.0, len(edges))
    rows, cols, flows = _symmetric(edges, q)
    p = MONOD_MASS_RING64.copy()
    p[:3] *= rng.uniform(0.7, 1.3, 3)
    C0 = np.full((1, n_c), 5.0)
    return Case(f"ensemble_{i}", "monod_mass", n_c, 1, 0.01, int(seed) * 1_000_003 + i, [(V, rows, cols, flows)],
                C0=C0, params=p, n_particles=n_p, x0=1.0, n_steps=n_steps)


def ensemble512(n_runs=512, seed=11, n_steps=500, **kw):
    """512 independent runs of uneven size (the batched ensemble)."""
    return [ensemble_run(i, seed, n_steps, **kw) for i in range(n_runs)]


def served_stream(n_requests, rate=20.0, seed=13, steps_range=(200, 5000), **kw):
    """The served stream: Poisson arrivals at `rate` per second, each request
    an ensemble run with a log-uniform step count.  Returns
    [(arrival_seconds, case)] in arrival order."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    out = []
    for i in range(n_requests):
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        out.append((float(t[i]), ensemble_run(i, seed, steps, **kw)))
    return out
