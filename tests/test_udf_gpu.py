"""Models loaded from plugins (the example udf_minimal, the reference's
minimal udf_model, and udf_starvation, whose cells die) run the same kernels
as built-in ones: parity with the oracle's restatements of the models, and
the same results when a plugin is named through BIOMC_LIB_UDF as
"udf_model"."""
import os
import subprocess
import sys

import numpy as np
import pytest

from parity import assert_parity, make_pair

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_plugin_model_parity():
    import mcst
    from mcst import cases

    mcst.load_model(mcst.EXAMPLE_PLUGIN)
    c = cases.small_case("udf_minimal", n_particles=5000)
    eng, g, o = make_pair(c)
    for _ in range(4):
        eng.step(5)
        o.step(5)
        assert_parity(g, o)
    cnt = g.counters()
    assert cnt["moves"] > 0 and cnt["exits"] > 0 and cnt["births"] > 0
    eng.close()


def test_plugin_model_with_deaths_parity():
    """The dying example plugin (udf_starvation) on the small network with a
    third of the compartments starved: deaths, removal of dead and exited
    cells, divisions and moves, against the oracle."""
    import mcst
    from mcst import cases

    mcst.load_model(mcst.EXAMPLE_PLUGIN_DEATH)
    c = cases.small_case("udf_starvation", n_particles=6000, seed=5)
    c.C0[0, ::3] = 0.0
    eng, g, o = make_pair(c)
    for _ in range(6):
        eng.step(5)
        o.step(5)
        assert_parity(g, o)
    cnt = g.counters()
    assert cnt["deaths"] > 0 and cnt["compactions"] > 0 and cnt["births"] > 0 and cnt["moves"] > 0
    eng.close()


_ALIAS_RUN = """
import sys, numpy as np, mcst
from mcst import cases
c = cases.small_case("udf_minimal", n_particles=3000)
c.model = sys.argv[1]
e = mcst.Engine(c.model)
r = e.add_run(c.n_c, c.dt, c.seed)
c.apply(r)
e.commit()
e.step(12)
p = r.particles()
np.savez(sys.argv[2], pos=p["pos"], state=p["state"], C=r.concentrations(),
         counters=np.array([r.counters()[k] for k in sorted(r.counters())], np.uint64))
e.close()
"""


def test_udf_model_alias_gives_the_same_results(tmp_path):
    import mcst

    env = dict(os.environ, BIOMC_LIB_UDF=mcst.EXAMPLE_PLUGIN,
               PYTHONPATH=os.pathsep.join([ROOT, os.path.join(ROOT, "biocma-mcst_amd")]))
    out = {}
    for name in ("udf_model", "udf_minimal"):
        f = str(tmp_path / f"{name}.npz")
        code = _ALIAS_RUN if name == "udf_model" else "import mcst; mcst.load_model(mcst.EXAMPLE_PLUGIN)\n" + _ALIAS_RUN
        r = subprocess.run([sys.executable, "-c", code, name, f], env=env, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr[-2000:]
        out[name] = np.load(f)
    for k in ("pos", "state", "C", "counters"):
        np.testing.assert_array_equal(out["udf_model"][k], out["udf_minimal"][k])
