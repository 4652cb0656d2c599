"""Result export in the reference's dataset layout.

The reference writes one HDF5 file per rank through `PartialExporter`
(apps/core/src/dataexporter/partial_exporter.cpp:85-163) from the values of
`PostProcessing::get_properties` (apps/core/public/core/post_process.hpp:
54-98,173-190).  h5py is not part of this environment, so `Exporter` keeps the
same dataset names and array shapes and writes them into one `.npz` archive:

    biological_model/<k>/<name>           per-particle values of export k
    biological_model/<k>/spatial/<name>   their sums per compartment
    biological_model/<k>/age, .../age_hydro
    records/number_particle               Idle particles per compartment, one row per export
    records/tallies                       event counters, one row per export
    records/concentration_liquid          liquid concentrations, [export][n_s][n_c]
    records/time                          simulated time of each export
    probes/exit_ages, probes/division_ages

Names are the model's variable names followed by "mass"; for the models the
reference also has, the exported variables are the reference's
(`names()`/`get_number()` of monod.hpp, fixed_length.hpp, simple_acetate.hpp).
Works with any run object of this package's run API (the HIP engine's `Run`
or the oracle's run).
"""
from __future__ import annotations

import numpy as np

# (index in the state, exported name) per model
EXPORTED = {
    "monod_mass": [(0, "m"), (1, "m_birth")],
    "dual_monod": [(0, "m"), (1, "uptake_rate"), (2, "storage")],
    "monod": [(0, "length"), (2, "mu"), (3, "mu_eff")],                     # monod.hpp:203-215
    "fixed_length": [(0, "length")],                                         # fixed_length.hpp:85-98
    "simple_acetate": [(i, n) for i, n in enumerate(                         # simple_acetate.hpp:115-121
        ["length", "l_max", "a_p", "a_max", "a_e", "a_e_s", "a_e_a", "phi_s", "phi_a"])],
    "quad_monod": [(i, n) for i, n in enumerate(
        ["m", "qs", "qa", "storage", "o2_rate", "overflow", "m_birth", "generation"])],
}

COUNTER_NAMES = ["steps", "births", "exits", "deaths", "moves", "overflow", "compactions", "n", "inactive",
                 "processed"]


class Exporter:
    def __init__(self, path: str, model: str):
        if model not in EXPORTED:
            raise ValueError(f"no export layout for model {model!r}")
        self.path = path
        self.vars = EXPORTED[model]
        self.data: dict[str, np.ndarray] = {}
        self.rows: dict[str, list] = {"records/number_particle": [], "records/tallies": [],
                                      "records/concentration_liquid": [], "records/time": []}
        self.k = 0

    def write(self, run, time: float | None = None) -> int:
        """Export number k of `run`: particle values, their spatial sums,
        ages, and one row of each record.  Returns k."""
        k = self.k
        values, spatial = run.properties()
        p = run.particles()
        idle = p["status"] == 0
        base = f"biological_model/{k}/"
        for i, name in self.vars:
            self.data[base + name] = values[i].copy()
            self.data[base + "spatial/" + name] = spatial[i].copy()
        self.data[base + "mass"] = values[-1].copy()
        self.data[base + "spatial/mass"] = spatial[-1].copy()
        self.data[base + "age_hydro"] = np.where(idle, p["age_hyd"], 0.0)
        self.data[base + "age"] = np.where(idle, p["age_div"], 0.0)
        cnt = run.counters()
        self.rows["records/number_particle"].append(np.asarray(run.repartition(), np.uint64))
        self.rows["records/tallies"].append(np.array([cnt[c] for c in COUNTER_NAMES], np.uint64))
        self.rows["records/concentration_liquid"].append(np.asarray(run.concentrations()))
        self.rows["records/time"].append(float(time) if time is not None else cnt["steps"] * run.dt)
        self.k += 1
        return k

    def write_probes(self, run):
        self.data["probes/exit_ages"] = np.asarray(run.exit_ages())
        self.data["probes/division_ages"] = np.asarray(run.division_ages())

    def close(self):
        out = dict(self.data)
        for name, rows in self.rows.items():
            if rows:
                out[name] = np.stack(rows)
        out["records/tally_names"] = np.array(COUNTER_NAMES)
        np.savez_compressed(self.path, **out)
        return self.path


def read(path: str) -> dict[str, np.ndarray]:
    """All datasets of an export, by their reference names."""
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}
