"""ctypes binding of the CPU oracle (oracle/libmcst_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, where it is the checker or the timed CPU
baseline -- never the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmcst_oracle.so")

MODEL_IDS = {
    "monod_mass": 0,
    "dual_monod": 1,
    "monod": 2,
    "fixed_length": 3,
    "simple_acetate": 4,
    "quad_monod": 5,
    "udf_minimal": 6,
    "udf_starvation": 7,
}
FEED_TYPES = {"constant": 0, "linear": 1, "exponential": 2}
KLA_MODES = {"fixed": 0, "turbulence": 1}
COUNTERS = ["steps", "births", "exits", "deaths", "moves", "overflow", "compactions", "n", "inactive", "processed"]

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc, no fast-math)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u01.restype = C.c_double
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_u01_52.restype = C.c_double
        L.orc_u01_52.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_norm_inv.restype = C.c_double
        L.orc_norm_inv.argtypes = [C.c_double]
        L.orc_norm_cdf.restype = C.c_double
        L.orc_norm_cdf.argtypes = [C.c_double]
        L.orc_feed_flow.restype = C.c_double
        L.orc_feed_flow.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_step_stages.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32]
        L.orc_model_info.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_model_default_params.argtypes = [C.c_int, _f64p, C.c_int]
        L.orc_transition_dense.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, _f64p]
        L.orc_build_move_tables.argtypes = [C.c_uint32, _f64p, C.c_uint64, _u32p, _u32p, _f64p, C.c_double,
                                            C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_transport_once.argtypes = [C.c_uint32, C.c_uint32, _f64p, C.c_uint64, _u32p, _u32p, _f64p, _f64p,
                                         _f64p, C.c_double, _f64p, _f64p, _f64p]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_model_params.argtypes = [C.c_void_p, _f64p, C.c_int]
        L.orc_add_flowmap.argtypes = [C.c_void_p, _f64p, C.c_uint64, _u32p, _u32p, _f64p]
        L.orc_set_steps_per_map.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_add_flowmap_tables.argtypes = [C.c_void_p, _f64p, C.c_uint32, _u64p, _f64p, _f64p, C.c_uint64,
                                             _u64p, _u64p, _f64p]
        L.orc_add_feed.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_int64,
                                   C.c_uint32, _u32p, _f64p]
        L.orc_set_concentrations.argtypes = [C.c_void_p, _f64p]
        L.orc_set_compaction.argtypes = [C.c_void_p, C.c_double, C.c_uint64]
        L.orc_set_division_cap.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_init_particles.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double]
        L.orc_set_particles.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, _u32p, _u8p, _f64p, _f64p, _f64p,
                                        C.c_double]
        L.orc_step.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
        L.orc_n.restype = C.c_uint64
        L.orc_n.argtypes = [C.c_void_p]
        L.orc_weight.restype = C.c_double
        L.orc_weight.argtypes = [C.c_void_p]
        L.orc_get_particles.argtypes = [C.c_void_p, _u32p, _u8p, _f64p, _f64p, _f64p]
        for name in ("orc_get_concentrations", "orc_get_mass", "orc_get_contributions"):
            getattr(L, name).argtypes = [C.c_void_p, _f64p]
        L.orc_get_counters.argtypes = [C.c_void_p, _u64p]
        L.orc_get_repartition.argtypes = [C.c_void_p, _u64p]
        L.orc_get_properties.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_set_gas_transfer.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_set_gas_phase.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_set_kla.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32, C.c_double]
        L.orc_add_gas_flowmap.argtypes = [C.c_void_p, _f64p, C.c_uint64, _u32p, _u32p, _f64p, C.c_void_p]
        L.orc_add_gas_feed.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_int64,
                                       C.c_uint32, _u32p, _f64p]
        L.orc_get_gas_concentrations.argtypes = [C.c_void_p, _f64p]
        L.orc_get_gas_mass.argtypes = [C.c_void_p, _f64p]
        L.orc_set_sort_interval.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_sort_particles.argtypes = [C.c_void_p]
        L.orc_set_growth.argtypes = [C.c_void_p, C.c_uint32, C.c_double]
        for name in ("orc_get_exit_ages", "orc_get_division_ages"):
            getattr(L, name).restype = C.c_uint64
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def model_info(model):
    nv, ns, npar = C.c_int(), C.c_int(), C.c_int()
    _check(lib().orc_model_info(MODEL_IDS[model], C.byref(nv), C.byref(ns), C.byref(npar)), "model_info")
    return nv.value, ns.value, npar.value


def default_params(model):
    nv, ns, npar = model_info(model)
    p = np.zeros(16, np.float64)
    _check(lib().orc_model_default_params(MODEL_IDS[model], p, 16), "default_params")
    return p[:npar].copy()


def transition_dense(n, rows, cols, flows):
    out = np.zeros((n, n), np.float64)
    _check(lib().orc_transition_dense(n, len(flows), np.ascontiguousarray(rows, np.uint32),
                                      np.ascontiguousarray(cols, np.uint32),
                                      np.ascontiguousarray(flows, np.float64), out), "transition_dense")
    return out


def move_tables(V, rows, cols, flows, dt):
    n = len(V)
    L = lib()
    m = C.c_uint32()
    args = (n, np.ascontiguousarray(V, np.float64), len(flows), np.ascontiguousarray(rows, np.uint32),
            np.ascontiguousarray(cols, np.uint32), np.ascontiguousarray(flows, np.float64), dt)
    _check(L.orc_build_move_tables(*args, C.byref(m), None, None, None, None), "move_tables")
    mm = m.value
    out_flow = np.zeros(n)
    p_move = np.zeros(n)
    cdf = np.zeros((n, mm))
    nbr = np.zeros((n, mm), np.uint32)
    _check(L.orc_build_move_tables(*args, C.byref(m), out_flow.ctypes.data, p_move.ctypes.data,
                                   cdf.ctypes.data, nbr.ctypes.data), "move_tables")
    return out_flow, p_move, cdf, nbr


def transport_once(V, rows, cols, flows, sink, S, dt, C_in, m):
    n_s, n = C_in.shape
    m = np.ascontiguousarray(m, np.float64).copy()
    C_out = np.zeros_like(m)
    _check(lib().orc_transport_once(n, n_s, np.ascontiguousarray(V, np.float64), len(flows),
                                    np.ascontiguousarray(rows, np.uint32), np.ascontiguousarray(cols, np.uint32),
                                    np.ascontiguousarray(flows, np.float64), np.ascontiguousarray(sink, np.float64),
                                    np.ascontiguousarray(S, np.float64), dt,
                                    np.ascontiguousarray(C_in, np.float64), m, C_out), "transport_once")
    return C_out, m


class OracleRun:
    """One simulation run on the CPU oracle; mirrors the product's run API."""

    def __init__(self, model, n_c, n_s, dt, seed):
        self.model = model
        self.n_var, ns, _ = model_info(model)
        if ns != n_s:
            raise ValueError(f"model {model} has {ns} species, got {n_s}")
        self.n_c, self.n_s, self.dt = n_c, n_s, dt
        self.h = lib().orc_create(MODEL_IDS[model], n_c, n_s, dt, seed)
        if not self.h:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float64)
        _check(lib().orc_set_model_params(self.h, p, len(p)), "set_params")

    def add_flowmap(self, V, rows, cols, flows):
        _check(lib().orc_add_flowmap(self.h, np.ascontiguousarray(V, np.float64), len(flows),
                                     np.ascontiguousarray(rows, np.uint32), np.ascontiguousarray(cols, np.uint32),
                                     np.ascontiguousarray(flows, np.float64)), "add_flowmap")

    def set_steps_per_map(self, k):
        _check(lib().orc_set_steps_per_map(self.h, k), "steps_per_map")

    def add_flowmap_tables(self, V, nbr, cdf, out_flow, rows, cols, vals):
        """A flowmap from neighbour/cumulative-probability rows [n][m] as given,
        out-flows [n] and the transition matrix as COO triplets."""
        nbr = np.ascontiguousarray(nbr, np.uint64)
        _check(lib().orc_add_flowmap_tables(self.h, np.ascontiguousarray(V, np.float64), nbr.shape[1], nbr,
                                            np.ascontiguousarray(cdf, np.float64),
                                            np.ascontiguousarray(out_flow, np.float64), len(vals),
                                            np.ascontiguousarray(rows, np.uint64),
                                            np.ascontiguousarray(cols, np.uint64),
                                            np.ascontiguousarray(vals, np.float64)), "add_flowmap_tables")

    def add_feed(self, kind, f0, param, inp, out, species, conc):
        sp = np.ascontiguousarray(species, np.uint32)
        cc = np.ascontiguousarray(conc, np.float64)
        _check(lib().orc_add_feed(self.h, FEED_TYPES[kind], f0, param, inp, -1 if out is None else out,
                                  len(sp), sp, cc), "add_feed")

    def set_concentrations(self, Cmat):
        _check(lib().orc_set_concentrations(self.h, np.ascontiguousarray(Cmat, np.float64)), "set_concentrations")

    def set_compaction(self, ratio, min_removal):
        _check(lib().orc_set_compaction(self.h, ratio, min_removal), "set_compaction")

    def set_division_cap(self, cap):
        _check(lib().orc_set_division_cap(self.h, cap), "set_division_cap")

    def set_sort_interval(self, steps):
        _check(lib().orc_set_sort_interval(self.h, steps), "set_sort_interval")

    def set_growth(self, every, factor=2.5):
        _check(lib().orc_set_growth(self.h, every, factor), "set_growth")

    def sort_particles(self):
        _check(lib().orc_sort_particles(self.h), "sort_particles")

    def set_gas_transfer(self, kla, cstar):
        kla = np.ascontiguousarray(np.broadcast_to(np.asarray(kla, float).reshape(self.n_s, -1),
                                                   (self.n_s, self.n_c)), np.float64)
        cstar = np.ascontiguousarray(np.broadcast_to(np.asarray(cstar, float).reshape(self.n_s, -1),
                                                     (self.n_s, self.n_c)), np.float64)
        _check(lib().orc_set_gas_transfer(self.h, kla, cstar), "set_gas_transfer")

    def set_gas_phase(self, henry, c_gas):
        """Two-phase run: Henry constants [n_s], initial gas concentrations
        [n_s][n_c] (or per species)."""
        c_gas = np.ascontiguousarray(np.broadcast_to(np.asarray(c_gas, float).reshape(self.n_s, -1),
                                                     (self.n_s, self.n_c)), np.float64)
        _check(lib().orc_set_gas_phase(self.h, np.ascontiguousarray(henry, np.float64), c_gas), "set_gas_phase")

    def set_kla(self, mode, kla=None, species=0, bubble_diameter=5e-3):
        """mode "fixed" (kla per species) or "turbulence" (species, bubble diameter)."""
        k = None if kla is None else np.ascontiguousarray(kla, np.float64)
        _check(lib().orc_set_kla(self.h, KLA_MODES[mode], None if k is None else k.ctypes.data, species,
                                 bubble_diameter), "set_kla")
        self._kla_keep = k

    def add_gas_flowmap(self, Vg, rows, cols, flows, eps=None):
        e = None if eps is None else np.ascontiguousarray(eps, np.float64)
        _check(lib().orc_add_gas_flowmap(self.h, np.ascontiguousarray(Vg, np.float64), len(flows),
                                         np.ascontiguousarray(rows, np.uint32), np.ascontiguousarray(cols, np.uint32),
                                         np.ascontiguousarray(flows, np.float64),
                                         None if e is None else e.ctypes.data), "add_gas_flowmap")

    def add_gas_feed(self, kind, f0, param, inp, out, species, conc):
        sp = np.ascontiguousarray(species, np.uint32)
        cc = np.ascontiguousarray(conc, np.float64)
        _check(lib().orc_add_gas_feed(self.h, FEED_TYPES[kind], f0, param, inp, -1 if out is None else out,
                                      len(sp), sp, cc), "add_gas_feed")

    def gas_concentrations(self):
        return self._mat("orc_get_gas_concentrations")

    def gas_mass(self):
        return self._mat("orc_get_gas_mass")

    def set_probes(self, cap):
        """The oracle always records exit and division ages."""

    def init_particles(self, n, capacity, x0):
        _check(lib().orc_init_particles(self.h, n, capacity, x0), "init_particles")

    def set_particles(self, capacity, pos, status, state, age_hyd, age_div, weight):
        n = len(pos)
        _check(lib().orc_set_particles(self.h, n, capacity, np.ascontiguousarray(pos, np.uint32),
                                       np.ascontiguousarray(status, np.uint8),
                                       np.ascontiguousarray(state, np.float64).reshape(-1),
                                       np.ascontiguousarray(age_hyd, np.float64),
                                       np.ascontiguousarray(age_div, np.float64), weight), "set_particles")

    def step(self, n_steps=1, threads=1):
        _check(lib().orc_step(self.h, n_steps, threads), "step")

    def step_stages(self, stages, threads=1):
        _check(lib().orc_step_stages(self.h, 1, threads, stages), "step_stages")

    @property
    def n(self):
        return int(lib().orc_n(self.h))

    @property
    def weight(self):
        return float(lib().orc_weight(self.h))

    def particles(self):
        n = self.n
        pos = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint8)
        state = np.zeros((self.n_var, n))
        ah = np.zeros(n)
        ad = np.zeros(n)
        _check(lib().orc_get_particles(self.h, pos, st, state.reshape(-1) if n else np.zeros(1), ah, ad),
               "get_particles")
        return {"pos": pos, "status": st, "state": state, "age_hyd": ah, "age_div": ad}

    def _mat(self, fn):
        out = np.zeros((self.n_s, self.n_c))
        _check(getattr(lib(), fn)(self.h, out), fn)
        return out

    def concentrations(self):
        return self._mat("orc_get_concentrations")

    def mass(self):
        return self._mat("orc_get_mass")

    def contributions(self):
        return self._mat("orc_get_contributions")

    def counters(self):
        c = np.zeros(len(COUNTERS), np.uint64)
        _check(lib().orc_get_counters(self.h, c), "counters")
        return {k: int(v) for k, v in zip(COUNTERS, c)}

    def repartition(self):
        h = np.zeros(self.n_c, np.uint64)
        _check(lib().orc_get_repartition(self.h, h), "repartition")
        return h

    def properties(self):
        n = self.n
        values = np.zeros((self.n_var + 1, max(n, 1)))
        spatial = np.zeros((self.n_var + 1, self.n_c))
        _check(lib().orc_get_properties(self.h, values, spatial), "properties")
        return values[:, :n], spatial

    def exit_ages(self):
        n = lib().orc_get_exit_ages(self.h, None, 0)
        out = np.zeros(max(n, 1))
        lib().orc_get_exit_ages(self.h, out.ctypes.data, n)
        return out[:n]

    def division_ages(self):
        n = lib().orc_get_division_ages(self.h, None, 0)
        out = np.zeros(max(n, 1))
        lib().orc_get_division_ages(self.h, out.ctypes.data, n)
        return out[:n]
