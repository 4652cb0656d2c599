"""Build libmcst.so (the HIP kernels and the C ABI) in-tree for gfx950.

    python biocma-mcst_amd/build.py

Flags: -ffp-contract=off keeps every FP64 expression rounded as written (the
CPU oracle uses the same rule), packed FP32 instructions are disabled, and
-munsafe-fp-atomics lets FP64 atomic adds use the hardware instruction.
"""
from __future__ import annotations

import os
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


def _compile(src: str, obj: str) -> None:
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass warns that it does not know the packed-fp32 feature; harmless
    noise = [ln for ln in r.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
             and "is not a recognized feature" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(noise))
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = True) -> str:
    with tempfile.TemporaryDirectory(prefix="mcst_build_") as tmp:
        objs = [os.path.join(tmp, s.replace(".hip", ".o")) for s in SOURCES]
        with ThreadPoolExecutor(len(SOURCES)) as ex:
            list(ex.map(lambda a: _compile(*a), zip(SOURCES, objs)))
        out = LIB + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs, "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr)
        os.replace(out, LIB)
    # the C++ host driver example over the C ABI (host code only: g++)
    ex_src = os.path.join(HERE, "examples", "ring_driver.cpp")
    ex_out = os.path.join(HERE, "examples", "ring_driver")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"), ex_src,
           "-o", ex_out, "-L", HERE, "-lmcst", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("example driver failed to build:\n" + r.stderr)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
