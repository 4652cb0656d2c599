"""Build libmcst.so (the HIP kernels and the C ABI) in-tree for gfx950, and
model plugins for it.

    python biocma-mcst_amd/build.py
    python biocma-mcst_amd/build.py --udf model.hpp --model Struct -o plugin.so

The first builds the library, the C++ example drivers (examples/ring_driver,
and examples/case_driver with the C++ case loader of host/) and the example
model plugins (examples/libudf_*.so); the second builds a plugin from a model
struct (the contract of csrc/models.hpp), for mcst_model_load.

Flags: -ffp-contract=off keeps every FP64 expression rounded as written (the
CPU oracle uses the same rule), packed FP32 instructions are disabled, and
-munsafe-fp-atomics lets FP64 atomic adds use the hardware instruction.
"""
from __future__ import annotations

import argparse
import os
import shlex
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmcst.so")
SOURCES = ["kernels.hip", "sort.hip", "engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _run_hipcc(cmd: list, what: str) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know the packed-fp32 feature; harmless
    noise = [ln for ln in r.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
             and "is not a recognized feature" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {what}:\n" + "\n".join(noise))
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def _compile(src: str, obj: str) -> None:
    _run_hipcc([HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", obj], src)


def build_udf(header: str, model: str, out: str, verbose: bool = True) -> str:
    """A model plugin: csrc/udf_plugin.hip instantiated for the struct `model`
    defined in `header`, linked into the shared library `out`."""
    header = os.path.abspath(header)
    if not os.path.isfile(header):
        raise FileNotFoundError(header)
    if not model.replace("::", "_").replace("_", "a").isalnum():
        raise ValueError(f"not a C++ type name: {model!r}")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", "-fvisibility=hidden", "-I", SRC, "-I", os.path.dirname(header),
           f"-DMCST_UDF_HEADER=\"{header}\"", f"-DMCST_UDF_MODEL={model}",
           os.path.join(SRC, "udf_plugin.hip"), "-o", tmp]
    _run_hipcc(cmd, f"the plugin of {model} ({shlex.quote(header)})")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build(verbose: bool = True) -> str:
    with tempfile.TemporaryDirectory(prefix="mcst_build_") as tmp:
        objs = [os.path.join(tmp, s.replace(".hip", ".o")) for s in SOURCES]
        with ThreadPoolExecutor(len(SOURCES)) as ex:
            list(ex.map(lambda a: _compile(*a), zip(SOURCES, objs)))
        out = LIB + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs, "-lpthread", "-ldl"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
        os.replace(out, LIB)
    # the C++ host drivers over the C ABI (host code only: g++): the ring
    # example, and the case driver with the C++ case loader (host/)
    for name, extra in (("ring_driver", []), ("case_driver", [os.path.join(HERE, "host", "case_loader.cpp")])):
        ex_src = os.path.join(HERE, "examples", name + ".cpp")
        ex_out = os.path.join(HERE, "examples", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"), ex_src,
               *extra, "-o", ex_out, "-L", HERE, "-lmcst", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example {name} failed to build:\n" + r.stderr)
    # the example model plugins
    for hdr, model in (("udf_minimal.hpp", "UdfMinimal"), ("udf_starvation.hpp", "UdfStarvation")):
        build_udf(os.path.join(HERE, "examples", hdr), model,
                  os.path.join(HERE, "examples", "lib" + hdr.replace(".hpp", ".so")), verbose=False)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--udf", help="header that defines a model struct: build a plugin of it")
    ap.add_argument("--model", help="the model struct's name (with --udf)")
    ap.add_argument("-o", "--output", help="the plugin's file (with --udf)")
    a = ap.parse_args()
    if a.udf:
        if not a.model or not a.output:
            ap.error("--udf needs --model and -o")
        build_udf(a.udf, a.model, a.output)
    else:
        build()
