#!/usr/bin/env python3
"""Particle-steps/s of the MI355X particle Monte Carlo step.

Workload: 5000-compartment random geometric network (mean degree 6) with 20
cyclic flowmaps swapped every 100 steps, 2 liquid species, 1e7 particles of
the dual-substrate Monod model (3 FP64 states + the FP64 division age; the
network has no outlet, so no hydraulic age), dt = 0.01 s.
A step is one pass of the whole pipeline (transport, fused particle pass,
mover/division resolution with the bookkeeping, removal) over all particles, with every input resident
in HBM.  particle-steps = sum over steps of the Idle particles processed.

Prints one JSON line.  --gpus above 1 is refused: this build targets one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "biocma-mcst_amd"))

HBM_PEAK_GBS = 8000.0
KERNELS = ["transport_layout", "particle", "resolve_bookkeep", "removal"]  # mcst_engine_profile's order
METRIC = "particle-steps/s at 1e7 particles (% HBM peak); served p50/p99 run latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per particle-kernel launch measured with rocprofv3 PMC
    counters (profiles/particle_rgg5000_traffic.txt), if present."""
    path = os.path.join(ROOT, "profiles", "particle_rgg5000_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# Threads of the CPU baseline: the host-core share of one GPU on the
# benchmark hosts (16), or fewer when the process may run on fewer cores.
CPU_SHARE = 16


def host_info():
    """nproc, the cores this process may run on, and lscpu's model/topology."""
    info = {"nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0))}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline_child(n_particles, seconds, threads):
    """Runs in a child process started before the parent touches the GPU:
    the oracle (C/OpenMP restatement of the step) on the full
    5000-compartment workload, timed over the step loop only."""
    from mcst import cases
    from oracle.oracle import OracleRun

    t_init = time.perf_counter()
    c = cases.rgg5000(n_particles=n_particles)
    r = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    t_init = time.perf_counter() - t_init
    r.step(2, threads=threads)  # warm: first touch of every page of the particle arrays
    done, t0 = 0, time.perf_counter()
    p0 = r.counters()["processed"]
    while time.perf_counter() - t0 < seconds:
        r.step(5, threads=threads)
        done += 5
    el = time.perf_counter() - t0
    ps = r.counters()["processed"] - p0
    print(json.dumps({"value": ps / el, "steps": done, "seconds": el, "init_s": t_init}), flush=True)


def cpu_baseline(n_particles=10_000_000, seconds=15.0):
    """The CPU baseline in a child process (so that the OpenMP placement of
    the reference's bench, OMP_PROC_BIND=spread OMP_PLACES=threads
    (tools/benchs/benchmark.py:147-153), pins the child's threads and never
    this process), run to completion before any HIP call here."""
    import subprocess

    host = host_info()
    threads = max(1, min(CPU_SHARE, host["affinity_cores"]))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread", OMP_PLACES="threads")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", "--particles", str(n_particles),
           "--cpu-seconds", str(seconds), "--cpu-threads", str(threads)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError("CPU baseline child failed:\n" + out.stderr[-2000:])
    res = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": res["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C/OpenMP restatement of the reference step) on the bench workload itself: "
                      f"5000-compartment network, {n_particles:.0e} dual-substrate Monod particles, "
                      f"{res['steps']} steps in {res['seconds']:.1f} s, {threads} threads "
                      f"(OMP_PROC_BIND=spread, OMP_PLACES=threads)",
            "host": host}


def served_with_lattice(n_requests=40, rate=20.0, seed=13, chunk=20, lattice_particles=10_000_000):
    """The served stream with the lattice tank (quad_monod, its own engine)
    arriving first and running its 200 steps alongside: the small requests'
    latencies with a large run in the mix."""
    from mcst import cases
    from mcst.serve import Request, Server, percentile

    reqs = [Request(i + 1, t, c, c.n_steps) for i, (t, c) in enumerate(cases.served_stream(n_requests, rate, seed))]
    big = cases.lattice50k(n_particles=lattice_particles)
    big.capacity = int(2.5 * lattice_particles)
    reqs.append(Request(0, 0.0, big, 200))
    slots = 2 * sum(r.case.n_particles for r in reqs[:-1]) + 1024 * n_requests
    srv = Server(pools={"monod_mass": (slots, n_requests), "quad_monod": (big.capacity + 1024, 1)},
                 budget_bytes=200 * 10**9, chunk_steps=chunk)
    res = srv.serve(reqs)
    srv.close()
    small = [r for r in res if r.rid != 0]
    lat = [r.latency for r in small]
    ps = sum(r.particle_steps for r in res)
    span = max(r.finished for r in res) - min(r.arrival for r in reqs)
    big_res = [r for r in res if r.rid == 0][0]
    return {"requests": n_requests, "lattice_particles": lattice_particles,
            "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99),
            "lattice_latency_s": big_res.latency, "aggregate_particle_steps_per_s": ps / span,
            "workload": "the served stream plus the 50 000-compartment lattice run (200 steps) arriving first"}


def served_measurement(n_requests=40, rate=20.0, seed=13, chunk=20):
    """The served stream (Poisson arrivals at `rate`/s, ensemble-sized runs with
    200..5000 steps) through the request loop, and the same requests one at a
    time on the same GPU.  Both legs run on engines created and committed
    before their clocks start (Server and OneAtATime: reserve/admit/retire);
    both define each request's particles inside their timing.

    Two comparisons, each against the one-at-a-time leg under the same
    arrivals: the stream at `rate`/s (arrival-bound: the ratio grows with the
    offered load), and the saturated case, every request arriving at once
    (the batched engine's capacity against back-to-back single runs)."""
    from mcst import cases
    from mcst.serve import OneAtATime, Request, Server, percentile

    reqs = [Request(i, t, c, c.n_steps) for i, (t, c) in enumerate(cases.served_stream(n_requests, rate, seed))]
    slots = 2 * sum(r.case.n_particles for r in reqs) + 1024 * n_requests

    def serve(rs):
        srv = Server("monod_mass", slots=slots, max_runs=n_requests, budget_bytes=200 * 10**9, chunk_steps=chunk)
        out = srv.serve(rs)
        srv.close()
        return out

    res = serve(reqs)
    burst = serve([Request(r.rid, 0.0, r.case, r.steps) for r in reqs])
    # one at a time: each request's duration on the idle pre-created engine
    one = OneAtATime("monod_mass", max(r.case.capacity or 2 * r.case.n_particles for r in reqs) + 2048, n_requests)
    durations, ps_seq = {}, 0
    for r in sorted(reqs, key=lambda q: q.arrival):
        durations[r.rid], p_steps = one.run(r)
        ps_seq += p_steps
    one.close()

    def sequential(arrivals):
        """The measured durations scheduled one after another against the
        arrivals: (latencies, span from the first arrival)."""
        t_free, lat = 0.0, []
        for rid, t in sorted(arrivals.items(), key=lambda kv: (kv[1], kv[0])):
            start = max(t, t_free)
            t_free = start + durations[rid]
            lat.append(t_free - t)
        return lat, t_free - min(arrivals.values())

    def leg(results, arrivals):
        ps = sum(r.particle_steps for r in results)
        span = max(r.finished for r in results) - min(arrivals.values())
        lat_seq, span_seq = sequential(arrivals)
        agg, agg_seq = ps / span, ps_seq / span_seq
        return {"p50_s": percentile([r.latency for r in results], 50),
                "p99_s": percentile([r.latency for r in results], 99),
                "aggregate_particle_steps_per_s": agg,
                "one_at_a_time": {"p50_s": percentile(lat_seq, 50), "p99_s": percentile(lat_seq, 99),
                                  "aggregate_particle_steps_per_s": agg_seq},
                "speedup_vs_one_at_a_time": agg / agg_seq}

    stream = leg(res, {r.rid: r.arrival for r in reqs})
    saturated = leg(burst, {r.rid: 0.0 for r in reqs})
    return {
        "requests": n_requests, "arrival_rate_per_s": rate, "chunk_steps": chunk,
        "particle_steps": sum(r.particle_steps for r in res),
        "p50_s": stream["p50_s"], "p99_s": stream["p99_s"],
        "aggregate_particle_steps_per_s": stream["aggregate_particle_steps_per_s"],
        f"stream_{rate:g}_per_s": stream,
        "saturated_all_at_once": saturated,
        "one_at_a_time_total_s": sum(durations.values()),
        "workload": "served stream: monod_mass runs, 32..4000 compartments, 1e4..2e6 particles, "
                    "200..5000 steps, Poisson arrivals; both legs on engines created before their clocks",
    }


def ensemble_measurement(n_runs=512, steps=500, graph=50, seed=11):
    """The batched ensemble: n_runs uneven runs in one engine; the particle
    pass's launch shape is chosen by the autotuner over the first steps, the
    rest are captured in a graph of `graph` steps and replayed."""
    import mcst
    from mcst import cases

    cs = cases.ensemble512(n_runs, seed, steps)
    eng = mcst.Engine("monod_mass")
    runs = []
    for c in cs:
        r = eng.add_run(c.n_c, c.dt, c.seed)
        c.apply(r)
        runs.append(r)
    eng.commit()
    tuned = eng.autotune(2)  # its trial steps are steps of the runs
    done = runs[0].counters()["steps"]
    eng.capture(graph)
    eng.replay(1)
    eng.synchronize()
    p0 = sum(r.counters()["processed"] for r in runs)
    reps = max(1, (steps - done - graph) // graph)
    t0 = time.perf_counter()
    eng.replay(reps)
    eng.synchronize()
    el = time.perf_counter() - t0
    ps = sum(r.counters()["processed"] for r in runs) - p0
    eng.close()
    return {"runs": n_runs, "particles": int(sum(c.n_particles for c in cs)),
            "compartments": int(sum(c.n_c for c in cs)), "steps_timed": reps * graph,
            "particle_steps_per_s": ps / el, "ms_per_step": el * 1e3 / (reps * graph),
            "autotuned": {"block": tuned["block"], "tile_group": tuned["group"], "trial_steps": done}}


def lattice_measurement(n_particles=100_000_000, steps=200, profile_steps=5, sort_interval=0):
    """The 50 000-compartment lattice tank at full size: one run of
    quad_monod (8 FP64 states, division age, hydraulic age), 4 species,
    outlet washout with device compaction, 200 steps of 5 ms, all timed."""
    import mcst
    from mcst import cases

    log(f"building the 50x50x20 lattice case with {n_particles:.0e} particles")
    case = cases.lattice50k(n_particles=n_particles)
    eng = mcst.Engine(case.model)
    run = eng.add_run(case.n_c, case.dt, case.seed)
    t0 = time.perf_counter()
    case.apply(run)
    run.set_sort_interval(sort_interval)
    eng.commit()
    eng.synchronize()
    log(f"initialised and uploaded in {time.perf_counter() - t0:.1f} s")
    c0 = run.counters()
    t0 = time.perf_counter()
    eng.step(steps)
    eng.synchronize()
    el = time.perf_counter() - t0
    c1 = run.counters()
    ps = c1["processed"] - c0["processed"]
    p0 = c1["processed"]
    _, per = eng.profile(profile_steps)
    per_launch = (run.counters()["processed"] - p0) / profile_steps
    _, nv, ns, _ = mcst.model_info(case.model)
    b_ps = 2 * (8 * (nv + 2) + 4 + 1)  # states, division age, hydraulic age (the top layer is an outlet)
    achieved = per_launch * b_ps / (per[1] * 1e-3) / 1e9
    eng.close()
    log(f"lattice50k: {ps / el:.3e} particle-steps/s, {el * 1e3 / steps:.2f} ms/step, particle pass {achieved:.0f} GB/s")
    return {"workload": "50x50x20 lattice tank, 50 000 compartments, 4 species, quad_monod (8 states), "
                        "outlet washout, dt=0.005 s", "particles_initial": int(c0["n"]), "particles_final": int(c1["n"]),
            "steps": steps, "particle_steps_per_s": ps / el, "ms_per_step": el * 1e3 / steps,
            "sort_interval": sort_interval,
            "births": int(c1["births"]), "exits": int(c1["exits"]), "compactions": int(c1["compactions"]),
            "kernel_ms_per_step": dict(zip(KERNELS,
                                           map(float, per))),
            "particle_pass": {"B_ps": b_ps, "achieved_GBps": achieved, "frac_of_8TBps": achieved / HBM_PEAK_GBS}}


def copy_calibration(mib, reps=5):
    """Plain device-to-device copies of a known size: what FETCH_SIZE and
    WRITE_SIZE read for them calibrates the particle kernel's counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_per_dispatch": n, "dispatches": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=10_000_000)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--graph", type=int, default=0, help="replay a captured graph of this many steps")
    ap.add_argument("--sort", type=int, default=0, metavar="K",
                    help="reorder the 1e7-particle run by compartment on the device every K steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--served-requests", type=int, default=40)
    ap.add_argument("--ensemble", action="store_true", help="also measure the 512-run batched ensemble")
    ap.add_argument("--lattice50k", type=float, default=0, metavar="N",
                    help="also run the 50 000-compartment lattice case with N particles (e.g. 1e8)")
    ap.add_argument("--served-lattice", type=float, default=0, metavar="N",
                    help="also serve the stream with an N-particle lattice run in the mix")
    ap.add_argument("--lattice50k-sort", type=int, default=0, metavar="K",
                    help="reorder the lattice case's particles by compartment every K steps (0: never)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of this many MiB (PMC counter calibration)")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=1, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.cpu_child:
        cpu_baseline_child(a.particles, a.cpu_seconds, a.cpu_threads)
        return
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the step would shard by particles with one all-reduce of the contribution matrix per step",
              file=sys.stderr)
        sys.exit(2)

    cpu = None
    if not a.no_cpu:  # first, before this process makes any HIP call
        log(f"timing the CPU baseline on {a.particles:.0e} particles (child process)")
        cpu = cpu_baseline(a.particles)
        log(f"cpu baseline {cpu['value']:.3e} particle-steps/s on {cpu['cores']} threads")

    import mcst
    from mcst import cases

    log("building the 5000-compartment case")
    case = cases.rgg5000(n_particles=a.particles)
    eng = mcst.Engine(case.model)
    run = eng.add_run(case.n_c, case.dt, case.seed)
    case.capacity = 2 * a.particles
    log(f"initialising {a.particles:.0e} particles")
    case.apply(run)
    if a.sort:
        run.set_sort_interval(a.sort)
    t_c = time.perf_counter()
    eng.commit()  # device allocation and the one host-to-device upload of particles, flowmaps and liquid
    t_commit = time.perf_counter() - t_c
    _, nv, ns, _ = mcst.model_info(case.model)
    # FP64 fields a step reads and writes per particle: the model state, the
    # division age, and the hydraulic age only when a leaving flow exists
    # (the leave pass is not run otherwise, kernels.hpp:56)
    leave = any(f[4] is not None for f in case.feeds)
    S = nv + 1 + (1 if leave else 0)
    b_ps = 2 * (8 * S + 4 + 1)
    import numpy as np

    m_width = int(np.bincount(case.maps[0][1], minlength=case.n_c).max())  # padded CDF row width

    log(f"warmup {a.warmup} steps")
    eng.step(a.warmup)
    eng.synchronize()
    c0 = run.counters()
    if a.graph:
        eng.capture(a.graph)
        eng.replay(1)
        eng.synchronize()
        c0 = run.counters()
        reps = max(1, a.steps // a.graph)
        steps = reps * a.graph
        log(f"timing {steps} steps as {reps} graph replays")
        t0 = time.perf_counter()
        eng.replay(reps)
        eng.synchronize()
        t1 = time.perf_counter()
    else:
        steps = a.steps
        log(f"timing {steps} steps")
        t0 = time.perf_counter()
        eng.step(steps)
        eng.synchronize()
        t1 = time.perf_counter()
    c1 = run.counters()
    elapsed = t1 - t0
    ps = c1["processed"] - c0["processed"]
    value = ps / elapsed
    log(f"{ps:.3e} particle-steps in {elapsed:.3f} s -> {value:.3e}/s")
    # the host-buffer view: the same steps with the upload at commit and one
    # read-back of every particle field counted in (never the value)
    t_r = time.perf_counter()
    run.particles()
    t_readback = time.perf_counter() - t_r
    pcie = {"value": ps / (elapsed + t_commit + t_readback), "unit": "particle-steps/s",
            "commit_s": t_commit, "readback_s": t_readback, "steps": steps}
    log(f"with the commit upload ({t_commit:.3f} s) and a read-back of the particles ({t_readback:.3f} s): "
        f"{pcie['value']:.3e}/s")

    log(f"profiling {a.profile_steps} steps with HIP events")
    p0 = run.counters()["processed"]
    _, per = eng.profile(a.profile_steps)
    per_launch = (run.counters()["processed"] - p0) / a.profile_steps
    t_particle = per[1] * 1e-3
    achieved = per_launch * b_ps / t_particle / 1e9
    names = KERNELS
    log("kernel ms/step: " + ", ".join(f"{n}={v:.4f}" for n, v in zip(names, per)))

    traffic = read_traffic()
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_particle<DualMonod>", "bytes_per_particle_step": b_ps,
            "particles_per_launch": per_launch, "kernel_ms": per[1]}
    served = None
    if not a.no_served:
        log(f"serving {a.served_requests} requests of the served stream, then one at a time")
        served = served_measurement(a.served_requests)
        st, sat = served[f"stream_{20.0:g}_per_s"], served["saturated_all_at_once"]
        log(f"served p50 {served['p50_s']:.3f} s p99 {served['p99_s']:.3f} s; aggregate vs one at a time: "
            f"{st['speedup_vs_one_at_a_time']:.2f}x at 20 req/s, {sat['speedup_vs_one_at_a_time']:.2f}x saturated")
    ensemble = None
    if a.ensemble:
        log("batched ensemble of 512 runs")
        ensemble = ensemble_measurement()
    lattice = lattice_measurement(int(a.lattice50k), sort_interval=a.lattice50k_sort) if a.lattice50k else None
    mixed = served_with_lattice(a.served_requests, lattice_particles=int(a.served_lattice)) if a.served_lattice else None
    nnz = case.nnz + case.n_c
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "5000-compartment random geometric network, 20 flowmaps x 100 steps, 2 species, "
                        f"{a.particles:.0e} dual-substrate Monod particles, dt=0.01 s",
            "particles": a.particles,
            "compartments": case.n_c,
            "species": ns,
            "fp64_fields_per_particle": S,
            "hydraulic_age_tracked": leave,
            "B_ps": b_ps,
            "B_fix": case.n_c * (8 * ns * 4 + 16 + 12 * m_width) + 12 * nnz,
            "graph_steps": a.graph,
            "kernel_ms_per_step": dict(zip(names, [float(x) for x in per])),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "served": served,
        # events of the timed steps: every division request is accepted while
        # the run has free slots (overflow counts refusals)
        "events": {k: int(c1[k] - c0[k]) for k in ("births", "overflow", "moves", "exits")},
        # the timed steps' particle-steps over their time plus the commit
        # (the one upload of the host buffers) and one read-back of the particles
        "host_buffers_inclusive": pcie,
    }
    if ensemble is not None:
        out["ensemble"] = ensemble
    if lattice is not None:
        out["lattice50k"] = lattice
    if mixed is not None:
        out["served_with_lattice"] = mixed
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
