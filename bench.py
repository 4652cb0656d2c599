#!/usr/bin/env python3
"""Particle-steps/s of the MI355X particle Monte Carlo step.

Workload: 5000-compartment random geometric network (mean degree 6) with 20
cyclic flowmaps swapped every 100 steps, 2 liquid species, 1e7 particles of
the dual-substrate Monod model (3 FP64 states + the FP64 division age; the
network has no outlet, so no hydraulic age), dt = 0.01 s.
A step is one pass of the whole pipeline (transport, fused particle pass,
bookkeeping, removal, merge) over all particles, with every input resident
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
METRIC = "particle-steps/s at 1e7 particles (% HBM peak); served p50/p99 run latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per particle-kernel launch measured with rocprofv3 PMC
    counters (profiles/particle_config1_traffic.txt), if present."""
    path = os.path.join(ROOT, "profiles", "particle_config1_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(case_fn, seconds=12.0, max_steps=20000, n_particles=1_000_000):
    from oracle.oracle import OracleRun

    threads = max(1, min(16, os.cpu_count() or 1))
    c = case_fn(n_particles=n_particles)
    r = OracleRun(c.model, c.n_c, c.n_s, c.dt, c.seed)
    c.apply(r)
    r.step(2, threads=threads)  # warm
    done, t0 = 0, time.perf_counter()
    p0 = r.counters()["processed"]
    while done < max_steps and time.perf_counter() - t0 < seconds:
        r.step(10, threads=threads)
        done += 10
    el = time.perf_counter() - t0
    ps = r.counters()["processed"] - p0
    return {"value": ps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C/OpenMP restatement) on the same 5000-compartment network, "
                      f"{n_particles:.0e} particles, {done} steps, {el:.1f} s"}


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
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of this many MiB (PMC counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "see DESIGN.md for how the step shards across GPUs", file=sys.stderr)
        sys.exit(2)

    import mcst
    from mcst import cases

    log("building the 5000-compartment case")
    case = cases.config1(n_particles=a.particles)
    eng = mcst.Engine(case.model)
    run = eng.add_run(case.n_c, case.dt, case.seed)
    case.capacity = 2 * a.particles
    log(f"initialising {a.particles:.0e} particles")
    case.apply(run)
    eng.commit()
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

    log(f"profiling {a.profile_steps} steps with HIP events")
    p0 = run.counters()["processed"]
    _, per = eng.profile(a.profile_steps)
    per_launch = (run.counters()["processed"] - p0) / a.profile_steps
    t_particle = per[1] * 1e-3
    achieved = per_launch * b_ps / t_particle / 1e9
    names = ["transport_layout", "particle", "resolve", "bookkeep", "removal", "merge"]
    log("kernel ms/step: " + ", ".join(f"{n}={v:.4f}" for n, v in zip(names, per)))

    traffic = read_traffic()
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_particle<DualMonod>", "bytes_per_particle_step": b_ps,
            "particles_per_launch": per_launch, "kernel_ms": per[1]}
    cpu = None
    if not a.no_cpu:
        log("timing the CPU baseline")
        cpu = cpu_baseline(cases.config1)
        log(f"cpu baseline {cpu['value']:.3e} particle-steps/s on {cpu['cores']} threads")
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
    }
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
