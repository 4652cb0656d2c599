"""The C-ABI library loads and exports every symbol include/mcst.h declares
(no device needed: only host-side registry calls are made)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcst.h")
LIB = os.path.join(ROOT, "biocma-mcst_amd", "libmcst.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcst_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["python", os.path.join(ROOT, "biocma-mcst_amd", "build.py")], check=True)
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import mcst

    assert sorted(mcst.EXPORTS) == declared()


def test_model_registry_matches_oracle(lib, oracle_lib):
    import mcst

    mcst.load_model(mcst.EXAMPLE_PLUGIN)  # the example plugins' models: built into the oracle
    mcst.load_model(mcst.EXAMPLE_PLUGIN_DEATH)
    for name in oracle_lib.MODEL_IDS:
        mid, nv, ns, npar = mcst.model_info(name)
        assert (nv, ns, npar) == oracle_lib.model_info(name)
        if not name.startswith("udf_"):
            assert mid == oracle_lib.MODEL_IDS[name]
        np.testing.assert_array_equal(mcst.default_params(name), oracle_lib.default_params(name))


def test_unknown_model_is_an_error(lib):
    import mcst

    with pytest.raises(mcst.MCSTError):
        mcst.model_info("no_such_model")
