"""CMA case input: the flowmaps a run is built from, read from disk.

The reference reads cases through rcmtool (a Rust library that is not part of
the reference tree, `subprojects/rcmtool.wrap`), called from
`GlobalInitialiser::init_transitionner` (apps/core/src/global_initaliser.cpp:
200-244,521-568).  Its on-disk format is documented nowhere in the reference;
the one case it ships, apps/api/tests/data/0d, fixes what is read here:

  * a vector file (`vofL.raw`, liquid volumes): u32 count, then count f64
    values, little-endian;
  * a matrix file (`flowL.raw`, liquid flows F(i, j) from i to j): u32 rows,
    u32 cols, then rows*cols f64 values, row-major (bytes past them are
    ignored: the fixture carries 16);
  * the case index (`cma_case`): a binary header followed by u32
    length-prefixed relative paths of the flowmap files.

Cases with several flowmaps are read from numbered subdirectories
`i_0, i_1, ...` of the case directory, each holding one `vofL.raw` and one
`flowL.raw` (the layout `write_case` produces); a directory holding the files
itself is a single flowmap.  Large flow matrices may instead be stored sparse
in `flowL.coo`: u64 nnz, u32 rows[nnz], u32 cols[nnz], f64 flows[nnz].

`time_step` and `steps_per_flowmap` restate the reference's choice of dt and
of the flowmap period (global_initaliser.cpp:52-114).
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np


def read_raw_vector(path: str) -> np.ndarray:
    b = open(path, "rb").read()
    if len(b) < 4:
        raise ValueError(f"{path}: shorter than its count")
    (n,) = struct.unpack("<I", b[:4])
    if len(b) < 4 + 8 * n:
        raise ValueError(f"{path}: {n} values announced, {(len(b) - 4) // 8} present")
    return np.frombuffer(b[4:4 + 8 * n], dtype="<f8").astype(np.float64)


def write_raw_vector(path: str, v) -> None:
    v = np.ascontiguousarray(v, dtype="<f8")
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(v)))
        f.write(v.tobytes())


def read_raw_matrix(path: str) -> np.ndarray:
    b = open(path, "rb").read()
    if len(b) < 8:
        raise ValueError(f"{path}: shorter than its shape")
    rows, cols = struct.unpack("<II", b[:8])
    if len(b) < 8 + 8 * rows * cols:
        raise ValueError(f"{path}: {rows}x{cols} values announced, {(len(b) - 8) // 8} present")
    return np.frombuffer(b[8:8 + 8 * rows * cols], dtype="<f8").reshape(rows, cols).astype(np.float64)


def write_raw_matrix(path: str, M) -> None:
    M = np.ascontiguousarray(M, dtype="<f8")
    with open(path, "wb") as f:
        f.write(struct.pack("<II", *M.shape))
        f.write(M.tobytes())


def read_coo(path: str):
    b = open(path, "rb").read()
    (nnz,) = struct.unpack("<Q", b[:8])
    need = 8 + nnz * (4 + 4 + 8)
    if len(b) < need:
        raise ValueError(f"{path}: {nnz} entries announced, file too short")
    o = 8
    rows = np.frombuffer(b[o:o + 4 * nnz], "<u4").astype(np.uint32)
    o += 4 * nnz
    cols = np.frombuffer(b[o:o + 4 * nnz], "<u4").astype(np.uint32)
    o += 4 * nnz
    flows = np.frombuffer(b[o:o + 8 * nnz], "<f8").astype(np.float64)
    return rows, cols, flows


def write_coo(path: str, rows, cols, flows) -> None:
    rows = np.ascontiguousarray(rows, "<u4")
    cols = np.ascontiguousarray(cols, "<u4")
    flows = np.ascontiguousarray(flows, "<f8")
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(flows)))
        f.write(rows.tobytes())
        f.write(cols.tobytes())
        f.write(flows.tobytes())


def read_case_index(path: str) -> list[str]:
    """The relative file paths listed in a binary case index: every u32
    length followed by that many printable bytes ending in '.raw'."""
    b = open(path, "rb").read()
    out, i = [], 0
    while i + 4 <= len(b):
        (L,) = struct.unpack("<I", b[i:i + 4])
        if 4 < L <= 4096 and i + 4 + L <= len(b):
            s = b[i + 4:i + 4 + L]
            if all(32 <= c < 127 for c in s) and s.endswith(b".raw"):
                out.append(s.decode())
                i += 4 + L
                continue
        i += 1
    return out


@dataclasses.dataclass
class CmaCase:
    """Flowmaps of a case: per map the liquid volumes and the COO flows."""
    volumes: list          # [n_maps] arrays [n_c]
    flows: list            # [n_maps] (rows, cols, flows)

    @property
    def n_c(self) -> int:
        return len(self.volumes[0])

    @property
    def n_maps(self) -> int:
        return len(self.volumes)


def _read_map(d: str):
    V = read_raw_vector(os.path.join(d, "vofL.raw"))
    if os.path.exists(os.path.join(d, "flowL.coo")):
        rows, cols, flows = read_coo(os.path.join(d, "flowL.coo"))
    else:
        F = read_raw_matrix(os.path.join(d, "flowL.raw"))
        if F.shape != (len(V), len(V)):
            raise ValueError(f"{d}: flow matrix {F.shape} for {len(V)} compartments")
        r, c = np.nonzero(F)
        keep = r != c  # the diagonal carries no exchange
        rows, cols, flows = r[keep].astype(np.uint32), c[keep].astype(np.uint32), F[r[keep], c[keep]]
    if len(flows) and (rows.max() >= len(V) or cols.max() >= len(V)):
        raise ValueError(f"{d}: flow entry outside the {len(V)} compartments")
    return V, (rows, cols, flows)


def load_case(directory: str) -> CmaCase:
    """Read a case directory (one flowmap, or numbered i_k subdirectories)."""
    subs = []
    k = 0
    while os.path.isdir(os.path.join(directory, f"i_{k}")):
        subs.append(os.path.join(directory, f"i_{k}"))
        k += 1
    if not subs:
        idx = os.path.join(directory, "cma_case")
        if os.path.exists(idx):
            listed = {os.path.basename(p) for p in read_case_index(idx)}
            if "vofL.raw" not in listed:
                raise ValueError(f"{idx}: lists no liquid volumes")
        subs = [directory]
    volumes, flows = [], []
    for d in subs:
        V, F = _read_map(d)
        volumes.append(V)
        flows.append(F)
    n = {len(v) for v in volumes}
    if len(n) != 1:
        raise ValueError(f"{directory}: flowmaps with different compartment counts {sorted(n)}")
    return CmaCase(volumes, flows)


def write_case(directory: str, volumes, flows, sparse: bool | None = None) -> None:
    """Write flowmaps in the layout load_case reads: i_k/vofL.raw plus a
    dense i_k/flowL.raw (or sparse flowL.coo, the default above 2048
    compartments)."""
    os.makedirs(directory, exist_ok=True)
    for k, (V, (rows, cols, q)) in enumerate(zip(volumes, flows)):
        d = os.path.join(directory, f"i_{k}")
        os.makedirs(d, exist_ok=True)
        write_raw_vector(os.path.join(d, "vofL.raw"), V)
        if sparse if sparse is not None else len(V) > 2048:
            write_coo(os.path.join(d, "flowL.coo"), rows, cols, q)
        else:
            F = np.zeros((len(V), len(V)))
            np.add.at(F, (np.asarray(rows, np.int64), np.asarray(cols, np.int64)), q)
            write_raw_matrix(os.path.join(d, "flowL.raw"), F)


def min_residence_time(case: CmaCase) -> float:
    """Smallest V_i / Q_out(i) over compartments with an outflow and over
    flowmaps (the hydrodynamic time scale of get_time_step)."""
    best = np.inf
    for V, (rows, _, q) in zip(case.volumes, case.flows):
        qout = np.bincount(np.asarray(rows, np.int64), weights=q, minlength=len(V))
        m = qout > 0
        if m.any():
            best = min(best, float(np.min(V[m] / qout[m])))
    return best


def time_step(case: CmaCase, user_dt: float) -> float:
    """get_time_step (global_initaliser.cpp:52-85): the user's dt if given
    (> 0) or for a 0-D case; else min residence time / 100."""
    if user_dt > 0 or case.n_c == 1:
        return user_dt
    tau = min_residence_time(case)
    if not np.isfinite(tau) or tau == 0.0:
        raise ValueError("no time step given and none can be estimated from the flowmaps")
    return tau / 100.0


def steps_per_flowmap(t_per_flowmap: float, n_maps: int, dt: float) -> int:
    """compute_n_per_flowmap (global_initaliser.cpp:104-114)."""
    if t_per_flowmap == 0 or n_maps == 1:
        return 1
    return int(t_per_flowmap / dt) + 1


def to_case(cma: CmaCase, model: str, C0, dt: float = 0.0, t_per_flowmap: float = 0.0, feeds=(),
            params=None, n_particles: int = 0, capacity: int = 0, x0: float = 1.0, seed: int = 1,
            name: str = "cma_case"):
    """A run set-up (`cases.Case`) over the case's flowmaps."""
    from . import cases

    C0 = np.atleast_2d(np.asarray(C0, np.float64))
    if C0.shape[1] == 1 and cma.n_c > 1:
        C0 = np.repeat(C0, cma.n_c, axis=1)
    d_t = time_step(cma, dt)
    maps = [(V, r, c, q) for V, (r, c, q) in zip(cma.volumes, cma.flows)]
    return cases.Case(name, model, cma.n_c, C0.shape[0], d_t, seed, maps,
                      steps_per_map=steps_per_flowmap(t_per_flowmap, cma.n_maps, d_t), feeds=list(feeds),
                      C0=C0, params=params, n_particles=n_particles, capacity=capacity, x0=x0)
