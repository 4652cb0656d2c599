"""Model plugins (mcst_model_load, the reference's udf_model): loading,
naming and refusals, on the host (no device call)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_example_plugin_registers_its_model():
    import mcst

    mid = mcst.load_model(mcst.EXAMPLE_PLUGIN)
    assert mid >= 100
    assert mcst.load_model(mcst.EXAMPLE_PLUGIN) == mid  # same path: same model
    assert mcst.model_info("udf_minimal") == (mid, 2, 1, 5)
    assert list(mcst.default_params("udf_minimal")) == [0.5, 2e-6 / 3600.0, 2e-6, 1e-3, 0.6e-6]


def test_a_library_that_is_no_plugin_is_refused():
    import mcst

    with pytest.raises(mcst.MCSTError, match="not a model plugin"):
        mcst.load_model(os.path.join(ROOT, "oracle", "libmcst_oracle.so"))
    with pytest.raises(mcst.MCSTError, match="cannot load"):
        mcst.load_model(os.path.join(ROOT, "no_such_plugin.so"))


def _in_subprocess(code, env_extra):
    env = dict(os.environ, **env_extra)
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "biocma-mcst_amd")])
    return subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)


def test_udf_model_name_loads_the_plugin_of_biomc_lib_udf():
    import mcst

    r = _in_subprocess("import mcst; print(mcst.model_info('udf_model')); print(mcst.model_info('udf_minimal'))",
                       {"BIOMC_LIB_UDF": mcst.EXAMPLE_PLUGIN})
    assert r.returncode == 0, r.stderr
    assert r.stdout.split("\n")[:2] == ["(100, 2, 1, 5)", "(100, 2, 1, 5)"]
    r = _in_subprocess("import mcst; mcst.model_info('udf_model')", {"BIOMC_LIB_UDF": ""})
    assert r.returncode != 0 and "BIOMC_LIB_UDF" in r.stderr


def test_a_plugin_may_not_take_a_taken_name(tmp_path):
    """A plugin whose struct is named like a built-in model is refused."""
    import mcst
    from build import build_udf

    hdr = tmp_path / "clash.hpp"
    src = open(os.path.join(ROOT, "biocma-mcst_amd", "examples", "udf_minimal.hpp")).read()
    hdr.write_text(src.replace('"udf_minimal"', '"monod_mass"'))
    out = build_udf(str(hdr), "UdfMinimal", str(tmp_path / "libclash.so"), verbose=False)
    with pytest.raises(mcst.MCSTError, match="built-in model"):
        mcst.load_model(out)
