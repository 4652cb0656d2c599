"""The division-age origin word (div_age in csrc/mcst_common.hpp), compiled
for the host with g++: the origin in the low 31 bits with a zero-base flag,
so that a restart from 0 is one store.  Checked against clock_age, the
expression every other age uses, and against step counts across the 32-bit
wrap of a run's clock."""
import math
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "biocma-mcst_amd", "csrc")
FRESH = 1 << 31

PROGRAM = r"""
#include "mcst_common.hpp"
#include <cstdio>
int main() {
  double base, dt;
  unsigned clock, word;
  while (std::scanf("%lf %u %u %lf", &base, &clock, &word, &dt) == 4)
    std::printf("%a %a\n", mcst::div_age(base, clock, word, dt),
                mcst::clock_age(base, clock, word & ~mcst::CLK_FRESH, dt));
  return 0;
}
"""


@pytest.fixture(scope="module")
def ages(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the HIP headers are not installed")
    d = tmp_path_factory.mktemp("age_words")
    src, exe = d / "ages.cpp", d / "ages"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}",
                    str(src), "-o", str(exe)], check=True, capture_output=True, timeout=120)

    def run(rows):
        text = "".join(f"{b!r} {c} {w} {dt!r}\n" for b, c, w, dt in rows)
        out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True, timeout=60)
        return [tuple(float.fromhex(x) for x in line.split()) for line in out.stdout.splitlines()]

    return run


def test_plain_word_is_the_clock_age(ages):
    """Without the flag the word is an origin and the age is clock_age,
    bit for bit, for any age of fewer than 2^31 steps."""
    rows = [(0.0, 7, 3, 0.01), (12.5, 1000, 0, 0.005), (3.25, 2**31 - 1, 0, 0.01), (1e-3, 17, 17, 0.1),
            (0.75, 2**31 + 5, 6, 0.01)]
    for (b, c, w, dt), (got, ref) in zip(rows, ages(rows)):
        assert got == ref and got == b + (c - w) * dt


def test_fresh_word_ignores_the_base(ages):
    """With the flag the base age is 0 whatever age_div holds (a newborn's
    slot may hold an old particle's base, or anything)."""
    rows = [(5.0, 40, 30 | FRESH, 0.01), (float("nan"), 40, 30 | FRESH, 0.01), (float("inf"), 9, 9 | FRESH, 0.5)]
    out = ages(rows)
    assert out[0][0] == 10 * 0.01 and out[1][0] == 10 * 0.01 and out[2][0] == 0.0
    assert math.isnan(out[1][1])  # the plain expression would have used the base


def test_ages_across_the_clock_wrap(ages):
    """A run's clock may wrap past 2^32 and a restart may happen at a clock
    of 2^31 or more, where the flag takes the origin's top bit: the age still
    counts the steps since the restart."""
    dt = 0.01
    rows, want = [], []
    for start in (2**31 - 2, 2**31, 2**31 + 12345, 2**32 - 3, 2**32 - 1):
        for k in (0, 1, 5, 1000, 2**31 - 1):
            clock = (start + k) % 2**32
            rows.append((99.0, clock, (start | FRESH) % 2**32, dt))
            want.append(k * dt)
    for (got, _), w in zip(ages(rows), want):
        assert got == w


def test_frozen_age_uses_the_stop_clock(ages):
    """A slot that stops being Idle keeps the division clock it stopped at
    (clk_hyd) and its division age is read at that clock, flag or not."""
    dt = 0.02
    rows = [(0.0, 50, 20 | FRESH, dt), (1.5, 50, 20, dt), (0.0, 3, (2**32 - 4) | FRESH, dt)]
    out = ages(rows)
    assert out[0][0] == 30 * dt and out[1][0] == 1.5 + 30 * dt and out[2][0] == 7 * dt
