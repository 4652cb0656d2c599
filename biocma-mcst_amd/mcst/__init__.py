"""Python binding of libmcst.so, the MI355X particle Monte Carlo step.

Thin ctypes layer over the C ABI in include/mcst.h.  There is no CPU
fallback: if the HIP library is missing or no device is present, the calls
raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libmcst.so")

FEED_KINDS = {"constant": 0, "linear": 1, "exponential": 2}
KLA_MODES = {"fixed": 0, "turbulence": 1}
COUNTERS = ["steps", "births", "exits", "deaths", "moves", "overflow", "compactions", "n", "inactive", "processed"]
STAGES = {"transport": 1, "model": 2, "contrib": 4, "move": 8, "leave": 16, "remove": 32}
STAGE_ALL = 63
IDLE, DIVISION, EXIT, DEAD = 0, 1, 2, 3

# every symbol include/mcst.h declares
EXPORTS = [
    "mcst_version", "mcst_last_error", "mcst_model_lookup", "mcst_model_default_params",
    "mcst_engine_create", "mcst_engine_destroy", "mcst_run_create", "mcst_run_set_model_params",
    "mcst_run_add_flowmap", "mcst_run_set_steps_per_flowmap", "mcst_run_add_feed",
    "mcst_run_set_concentrations", "mcst_run_set_removal", "mcst_run_set_division_cap",
    "mcst_run_set_probes", "mcst_run_set_step_target", "mcst_run_init_particles", "mcst_run_set_particles",
    "mcst_engine_commit", "mcst_engine_step", "mcst_engine_step_stages", "mcst_engine_capture",
    "mcst_engine_replay", "mcst_engine_synchronize", "mcst_engine_profile", "mcst_run_get_counters",
    "mcst_run_get_n", "mcst_run_get_weight", "mcst_run_get_particles", "mcst_run_get_concentrations",
    "mcst_run_get_mass", "mcst_run_get_contributions", "mcst_run_get_repartition",
    "mcst_run_get_exit_ages", "mcst_run_get_division_ages",
    "mcst_engine_reserve", "mcst_engine_admit", "mcst_engine_retire",
    "mcst_run_checkpoint", "mcst_run_restore", "mcst_run_get_properties", "mcst_engine_autotune",
    "mcst_run_set_gas_transfer", "mcst_run_set_sort_interval", "mcst_run_sort_particles",
    "mcst_engine_get_steps", "mcst_model_load", "mcst_run_set_growth", "mcst_run_add_flowmap_tables",
    "mcst_run_get_repartition_noalloc", "mcst_run_set_gas_phase", "mcst_run_set_kla", "mcst_run_add_gas_flowmap",
    "mcst_run_add_gas_feed", "mcst_run_get_gas_concentrations", "mcst_run_get_gas_mass",
]

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


class MCSTError(RuntimeError):
    pass


def lib():
    """Load libmcst.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MCSTError(f"{LIB_PATH} is missing: run `python biocma-mcst_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        i32, u32, u64, i64, dbl = C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_double
        sig = {
            "mcst_version": [C.POINTER(C.c_int)] * 3,
            "mcst_model_lookup": [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)],
            "mcst_model_default_params": [C.c_int, _f64p, C.c_int],
            "mcst_model_load": [C.c_char_p, C.POINTER(C.c_int)],
            "mcst_run_set_growth": [_vp, i32, u32, dbl],
            "mcst_engine_create": [C.c_char_p, C.c_int, C.POINTER(_vp)],
            "mcst_engine_destroy": [C.POINTER(_vp)],
            "mcst_run_create": [_vp, u32, u32, dbl, u64, C.POINTER(i32)],
            "mcst_run_set_model_params": [_vp, i32, _f64p, C.c_int],
            "mcst_run_add_flowmap": [_vp, i32, _f64p, u64, _u32p, _u32p, _f64p],
            "mcst_run_add_flowmap_tables": [_vp, i32, _f64p, u32, _u64p, _f64p, _f64p, u64, _u64p, _u64p, _f64p],
            "mcst_run_set_steps_per_flowmap": [_vp, i32, u32],
            "mcst_run_add_feed": [_vp, i32, C.c_int, dbl, dbl, u32, i64, u32, _u32p, _f64p],
            "mcst_run_set_concentrations": [_vp, i32, _f64p],
            "mcst_run_set_removal": [_vp, i32, dbl, u64],
            "mcst_run_set_division_cap": [_vp, i32, u32],
            "mcst_run_set_probes": [_vp, i32, u32],
            "mcst_run_set_step_target": [_vp, i32, u32],
            "mcst_run_init_particles": [_vp, i32, u64, u64, dbl],
            "mcst_run_set_particles": [_vp, i32, u64, u64, _u32p, _u8p, _f64p, _f64p, _f64p, dbl],
            "mcst_engine_commit": [_vp],
            "mcst_engine_reserve": [_vp, u64, u32],
            "mcst_engine_admit": [_vp, C.POINTER(u32)],
            "mcst_engine_retire": [_vp, i32],
            "mcst_engine_step": [_vp, u32],
            "mcst_engine_step_stages": [_vp, u32],
            "mcst_engine_capture": [_vp, u32],
            "mcst_engine_replay": [_vp, u32],
            "mcst_engine_synchronize": [_vp],
            "mcst_engine_profile": [_vp, u32, C.POINTER(dbl), _f64p],
            "mcst_run_get_counters": [_vp, i32, _u64p],
            "mcst_run_get_n": [_vp, i32, C.POINTER(u64)],
            "mcst_run_get_weight": [_vp, i32, C.POINTER(dbl)],
            "mcst_run_get_particles": [_vp, i32, _vp, _vp, _vp, _vp, _vp],
            "mcst_run_get_concentrations": [_vp, i32, _f64p],
            "mcst_run_get_mass": [_vp, i32, _f64p],
            "mcst_run_get_contributions": [_vp, i32, _f64p],
            "mcst_run_get_repartition": [_vp, i32, _u64p],
            "mcst_run_get_repartition_noalloc": [_vp, i32, _u64p],
            "mcst_run_get_exit_ages": [_vp, i32, _vp, u64, C.POINTER(u64)],
            "mcst_run_get_division_ages": [_vp, i32, _vp, u64, C.POINTER(u64)],
            "mcst_run_checkpoint": [_vp, i32, _vp, u64, C.POINTER(u64)],
            "mcst_run_restore": [_vp, i32, _vp, u64],
            "mcst_run_get_properties": [_vp, i32, _vp, _vp],
            "mcst_engine_autotune": [_vp, u32, C.POINTER(i32), C.POINTER(i32), C.POINTER(dbl)],
            "mcst_run_set_gas_transfer": [_vp, i32, _f64p, _f64p],
            "mcst_run_set_gas_phase": [_vp, i32, _f64p, _f64p],
            "mcst_run_set_kla": [_vp, i32, C.c_int, _vp, u32, dbl],
            "mcst_run_add_gas_flowmap": [_vp, i32, _f64p, u64, _u32p, _u32p, _f64p, _vp],
            "mcst_run_add_gas_feed": [_vp, i32, C.c_int, dbl, dbl, u32, i64, u32, _u32p, _f64p],
            "mcst_run_get_gas_concentrations": [_vp, i32, _f64p],
            "mcst_run_get_gas_mass": [_vp, i32, _f64p],
            "mcst_run_set_sort_interval": [_vp, i32, u32],
            "mcst_run_sort_particles": [_vp, i32],
            "mcst_engine_get_steps": [_vp, _u32p, u32],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.mcst_last_error.argtypes = []
        L.mcst_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _ck(rc, what):
    if rc != 0:
        msg = lib().mcst_last_error().decode(errors="replace")
        raise MCSTError(f"{what} failed ({rc}): {msg}")


def model_info(name):
    mid, nv, ns, npar = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _ck(lib().mcst_model_lookup(name.encode(), C.byref(mid), C.byref(nv), C.byref(ns), C.byref(npar)),
        "model_lookup")
    return mid.value, nv.value, ns.value, npar.value


def load_model(path):
    """Load a model plugin (built by `build.py --udf`); returns its model id.
    Its model is then known by the name its struct declares."""
    mid = C.c_int()
    _ck(lib().mcst_model_load(os.fsencode(os.path.abspath(path)), C.byref(mid)), "model_load")
    return mid.value


EXAMPLE_PLUGIN = os.path.join(_PKG, "examples", "libudf_minimal.so")
EXAMPLE_PLUGIN_DEATH = os.path.join(_PKG, "examples", "libudf_starvation.so")


def default_params(name):
    mid, nv, ns, npar = model_info(name)
    p = np.zeros(16)
    _ck(lib().mcst_model_default_params(mid, p, 16), "default_params")
    return p[:npar].copy()


class Engine:
    """A device-resident batch of runs sharing one particle model."""

    def __init__(self, model, device=0):
        self.model = model
        _, self.n_var, self.n_s, self.n_params = model_info(model)
        h = _vp()
        _ck(lib().mcst_engine_create(model.encode(), device, C.byref(h)), "engine_create")
        self.h = h
        self.runs = []

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().mcst_engine_destroy(C.byref(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def add_run(self, n_c, dt, seed):
        rid = C.c_int32()
        _ck(lib().mcst_run_create(self.h, n_c, self.n_s, dt, seed, C.byref(rid)), "run_create")
        run = Run(self, rid.value, n_c)
        run.dt = float(dt)
        self.runs.append(run)
        return run

    def commit(self):
        _ck(lib().mcst_engine_commit(self.h), "commit")

    def reserve(self, slots, runs):
        """Before commit: a particle-slot pool and run entries for runs added later."""
        _ck(lib().mcst_engine_reserve(self.h, int(slots), int(runs)), "reserve")

    def admit(self):
        """After commit: put the runs defined since on the device, in id order.
        Returns how many were admitted; stops quietly at the first that does
        not fit (it stays defined)."""
        n = C.c_uint32()
        rc = lib().mcst_engine_admit(self.h, C.byref(n))
        if rc not in (0, -4):  # MCST_ERR_MEMORY: no room for the next run yet
            _ck(rc, "admit")
        return n.value

    def retire(self, run):
        _ck(lib().mcst_engine_retire(self.h, run.id), "retire")

    def step(self, n=1):
        _ck(lib().mcst_engine_step(self.h, n), "step")

    def step_stages(self, stages):
        _ck(lib().mcst_engine_step_stages(self.h, stages), "step_stages")

    def capture(self, n_steps):
        _ck(lib().mcst_engine_capture(self.h, n_steps), "capture")

    def replay(self, times=1):
        _ck(lib().mcst_engine_replay(self.h, times), "replay")

    def synchronize(self):
        _ck(lib().mcst_engine_synchronize(self.h), "synchronize")

    def run_steps(self):
        """Step count of every run defined on this engine (0 off the device),
        read from the device at once."""
        out = np.zeros(max(len(self.runs), 1), np.uint32)
        _ck(lib().mcst_engine_get_steps(self.h, out, len(self.runs)), "get_steps")
        return out[: len(self.runs)]

    def autotune(self, steps=3):
        """Pick the particle pass's launch shape by timing each candidate over
        `steps` real steps (mcst_engine_autotune); returns the choice."""
        blk, grp, ms = C.c_int32(), C.c_int32(), C.c_double()
        _ck(lib().mcst_engine_autotune(self.h, steps, C.byref(blk), C.byref(grp), C.byref(ms)), "autotune")
        return {"block": blk.value, "group": grp.value, "particle_ms": ms.value}

    def profile(self, n_steps):
        """(total ms, per-kernel mean ms: transport+layout, particle, resolve+bookkeeping, removal)."""
        tot = C.c_double()
        per = np.zeros(4)
        _ck(lib().mcst_engine_profile(self.h, n_steps, C.byref(tot), per), "profile")
        return tot.value, per


class Run:
    """One run of an Engine; method names mirror oracle.OracleRun."""

    def __init__(self, engine, rid, n_c):
        self.e, self.id, self.n_c = engine, rid, n_c
        self.n_s, self.n_var = engine.n_s, engine.n_var

    @property
    def _h(self):
        return self.e.h

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float64)
        _ck(lib().mcst_run_set_model_params(self._h, self.id, p, len(p)), "set_params")

    def add_flowmap(self, V, rows, cols, flows):
        _ck(lib().mcst_run_add_flowmap(self._h, self.id, np.ascontiguousarray(V, np.float64), len(flows),
                                       np.ascontiguousarray(rows, np.uint32), np.ascontiguousarray(cols, np.uint32),
                                       np.ascontiguousarray(flows, np.float64)), "add_flowmap")

    def set_steps_per_map(self, k):
        _ck(lib().mcst_run_set_steps_per_flowmap(self._h, self.id, k), "steps_per_map")

    def add_flowmap_tables(self, V, nbr, cdf, out_flow, rows, cols, vals):
        """A flowmap from neighbour/cumulative-probability rows [n][m] as given,
        out-flows [n] and the transition matrix as COO triplets
        (mcst_run_add_flowmap_tables)."""
        nbr = np.ascontiguousarray(nbr, np.uint64)
        _ck(lib().mcst_run_add_flowmap_tables(self._h, self.id, np.ascontiguousarray(V, np.float64), nbr.shape[1],
                                              nbr, np.ascontiguousarray(cdf, np.float64),
                                              np.ascontiguousarray(out_flow, np.float64), len(vals),
                                              np.ascontiguousarray(rows, np.uint64),
                                              np.ascontiguousarray(cols, np.uint64),
                                              np.ascontiguousarray(vals, np.float64)), "add_flowmap_tables")

    def add_feed(self, kind, f0, param, inp, out, species, conc):
        sp = np.ascontiguousarray(species, np.uint32)
        cc = np.ascontiguousarray(conc, np.float64)
        _ck(lib().mcst_run_add_feed(self._h, self.id, FEED_KINDS[kind], f0, param, inp, -1 if out is None else out,
                                    len(sp), sp, cc), "add_feed")

    def set_concentrations(self, Cmat):
        _ck(lib().mcst_run_set_concentrations(self._h, self.id, np.ascontiguousarray(Cmat, np.float64)),
            "set_concentrations")

    def set_compaction(self, ratio, min_removal):
        _ck(lib().mcst_run_set_removal(self._h, self.id, ratio, min_removal), "set_removal")

    def set_division_cap(self, cap):
        """At most `cap` newborns per step (the reference's division buffer)."""
        _ck(lib().mcst_run_set_division_cap(self._h, self.id, cap), "set_division_cap")

    def set_sort_interval(self, steps):
        _ck(lib().mcst_run_set_sort_interval(self._h, self.id, steps), "set_sort_interval")

    def set_growth(self, every, factor=2.5):
        """Grow the capacity to ceil(n*factor) when fewer slots than particles
        are free, checked before every `every`-th step (0: fixed capacity)."""
        _ck(lib().mcst_run_set_growth(self._h, self.id, every, factor), "set_growth")

    def sort_particles(self):
        _ck(lib().mcst_run_sort_particles(self._h, self.id), "sort_particles")

    def set_gas_transfer(self, kla, cstar):
        """kla [1/s] and saturation C* [kg/m3], per species (broadcast over
        compartments) or [n_s][n_c]."""
        kla = np.ascontiguousarray(np.broadcast_to(np.asarray(kla, float).reshape(self.n_s, -1),
                                                   (self.n_s, self.n_c)), np.float64)
        cstar = np.ascontiguousarray(np.broadcast_to(np.asarray(cstar, float).reshape(self.n_s, -1),
                                                     (self.n_s, self.n_c)), np.float64)
        _ck(lib().mcst_run_set_gas_transfer(self._h, self.id, kla, cstar), "set_gas_transfer")

    def set_gas_phase(self, henry, c_gas):
        """Two-phase run: Henry constants [n_s], initial gas concentrations
        [n_s][n_c] (or per species)."""
        c_gas = np.ascontiguousarray(np.broadcast_to(np.asarray(c_gas, float).reshape(self.n_s, -1),
                                                     (self.n_s, self.n_c)), np.float64)
        _ck(lib().mcst_run_set_gas_phase(self._h, self.id, np.ascontiguousarray(henry, np.float64), c_gas),
            "set_gas_phase")

    def set_kla(self, mode, kla=None, species=0, bubble_diameter=5e-3):
        """mode "fixed" (kla per species) or "turbulence" (species, bubble diameter)."""
        k = None if kla is None else np.ascontiguousarray(kla, np.float64)
        _ck(lib().mcst_run_set_kla(self._h, self.id, KLA_MODES[mode], None if k is None else k.ctypes.data,
                                   species, bubble_diameter), "set_kla")

    def add_gas_flowmap(self, Vg, rows, cols, flows, eps=None):
        e = None if eps is None else np.ascontiguousarray(eps, np.float64)
        _ck(lib().mcst_run_add_gas_flowmap(self._h, self.id, np.ascontiguousarray(Vg, np.float64), len(flows),
                                           np.ascontiguousarray(rows, np.uint32),
                                           np.ascontiguousarray(cols, np.uint32),
                                           np.ascontiguousarray(flows, np.float64),
                                           None if e is None else e.ctypes.data), "add_gas_flowmap")

    def add_gas_feed(self, kind, f0, param, inp, out, species, conc):
        sp = np.ascontiguousarray(species, np.uint32)
        cc = np.ascontiguousarray(conc, np.float64)
        _ck(lib().mcst_run_add_gas_feed(self._h, self.id, FEED_KINDS[kind], f0, param, inp,
                                        -1 if out is None else out, len(sp), sp, cc), "add_gas_feed")

    def gas_concentrations(self):
        return self._mat("mcst_run_get_gas_concentrations")

    def gas_mass(self):
        return self._mat("mcst_run_get_gas_mass")

    def set_probes(self, cap):
        _ck(lib().mcst_run_set_probes(self._h, self.id, cap), "set_probes")

    def set_step_target(self, steps):
        _ck(lib().mcst_run_set_step_target(self._h, self.id, steps), "set_step_target")

    def init_particles(self, n, capacity, x0):
        _ck(lib().mcst_run_init_particles(self._h, self.id, n, capacity, x0), "init_particles")

    def set_particles(self, capacity, pos, status, state, age_hyd, age_div, weight):
        n = len(pos)
        _ck(lib().mcst_run_set_particles(self._h, self.id, n, capacity, np.ascontiguousarray(pos, np.uint32),
                                         np.ascontiguousarray(status, np.uint8),
                                         np.ascontiguousarray(state, np.float64).reshape(-1),
                                         np.ascontiguousarray(age_hyd, np.float64),
                                         np.ascontiguousarray(age_div, np.float64), weight), "set_particles")

    @property
    def n(self):
        v = C.c_uint64()
        _ck(lib().mcst_run_get_n(self._h, self.id, C.byref(v)), "get_n")
        return int(v.value)

    @property
    def weight(self):
        v = C.c_double()
        _ck(lib().mcst_run_get_weight(self._h, self.id, C.byref(v)), "get_weight")
        return float(v.value)

    def particles(self):
        n = self.n
        pos = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint8)
        state = np.zeros((self.n_var, n))
        ah = np.zeros(n)
        ad = np.zeros(n)
        if n:
            _ck(lib().mcst_run_get_particles(self._h, self.id, pos.ctypes.data, st.ctypes.data, state.ctypes.data,
                                             ah.ctypes.data, ad.ctypes.data), "get_particles")
        return {"pos": pos, "status": st, "state": state, "age_hyd": ah, "age_div": ad}

    def _mat(self, fn):
        out = np.zeros((self.n_s, self.n_c))
        _ck(getattr(lib(), fn)(self._h, self.id, out), fn)
        return out

    def concentrations(self):
        return self._mat("mcst_run_get_concentrations")

    def mass(self):
        return self._mat("mcst_run_get_mass")

    def contributions(self):
        return self._mat("mcst_run_get_contributions")

    def counters(self):
        c = np.zeros(len(COUNTERS), np.uint64)
        _ck(lib().mcst_run_get_counters(self._h, self.id, c), "get_counters")
        return {k: int(v) for k, v in zip(COUNTERS, c)}

    def repartition(self):
        h = np.zeros(self.n_c, np.uint64)
        _ck(lib().mcst_run_get_repartition(self._h, self.id, h), "get_repartition")
        return h

    def repartition_noalloc(self):
        """repartition() through the entry point that allocates no device
        memory (mcst_run_get_repartition_noalloc)."""
        h = np.zeros(self.n_c, np.uint64)
        _ck(lib().mcst_run_get_repartition_noalloc(self._h, self.id, h), "get_repartition_noalloc")
        return h

    def _probe(self, fn):
        n = C.c_uint64()
        _ck(getattr(lib(), fn)(self._h, self.id, None, 0, C.byref(n)), fn)
        out = np.zeros(max(n.value, 1))
        _ck(getattr(lib(), fn)(self._h, self.id, out.ctypes.data, n.value, C.byref(n)), fn)
        return out[: n.value]

    def exit_ages(self):
        return self._probe("mcst_run_get_exit_ages")

    def division_ages(self):
        return self._probe("mcst_run_get_division_ages")

    def properties(self, values=True):
        """(values [n_var+1][n], spatial [n_var+1][n_c]): each Idle particle's
        state and mass, and their per-compartment sums (mcst_run_get_properties).
        values=False reads only the sums (values is then None)."""
        n = self.n
        vals = np.zeros((self.n_var + 1, n)) if values else None
        spatial = np.zeros((self.n_var + 1, self.n_c))
        _ck(lib().mcst_run_get_properties(self._h, self.id, vals.ctypes.data if values and n else None,
                                          spatial.ctypes.data), "get_properties")
        return vals, spatial

    def checkpoint(self):
        """Everything the next step needs, as bytes (mcst_run_checkpoint)."""
        n = C.c_uint64()
        _ck(lib().mcst_run_checkpoint(self._h, self.id, None, 0, C.byref(n)), "checkpoint")
        buf = np.zeros(n.value, np.uint8)
        _ck(lib().mcst_run_checkpoint(self._h, self.id, buf.ctypes.data, n.value, C.byref(n)), "checkpoint")
        return buf.tobytes()

    def restore(self, blob):
        """Load a checkpoint into this defined run (same n_c, n_s, dt, seed)."""
        buf = np.frombuffer(blob, np.uint8)
        _ck(lib().mcst_run_restore(self._h, self.id, buf.ctypes.data, len(buf)), "restore")
