"""The library's host code and the C oracle under AddressSanitizer + UndefinedBehaviorSanitizer (CPU).

tests/native/Makefile builds tests/native/sanitize_host from the host sources (DAG compiler,
schedule and reverse-plan builders, statistics and serving bookkeeping) and oracle/mcdp_oracle.c,
all instrumented, linked with the device objects that mc_dagprop_amd/build.py leaves in
mc_dagprop_amd/_build (uninstrumented; nothing runs on a device here).  The program compiles random
DAGs with repeated precedence targets, index gaps, edges outside the link table, cycles, events out
of range and reserved types, builds every schedule shape and the reverse plan, and runs the
oracle's propagation and gradient; any out-of-bounds access or undefined operation aborts it.
A planted out-of-bounds read, built with the same flags, must be reported: the sanitizer is live.
"""

from __future__ import annotations

import os
import re
import subprocess
from pathlib import Path

import pytest

NATIVE = Path(__file__).resolve().parent / "native"
DEV_OBJ = Path(__file__).resolve().parents[1] / "mc_dagprop_amd" / "_build" / "mcdp_kernels_dev.o"
ENV = {**os.environ, "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


@pytest.fixture(scope="module")
def driver() -> Path:
    if not DEV_OBJ.exists():
        pytest.fail("device objects missing: run __graft_entry__.build() first")
    res = subprocess.run(["make", "-C", str(NATIVE), "-j8"], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    return NATIVE / "_build" / "sanitize_host"


def test_host_code_is_clean_under_sanitizers(driver: Path) -> None:
    res = subprocess.run([str(driver), "2000"], capture_output=True, text=True, env=ENV, timeout=600)
    assert res.returncode == 0 and "SANITIZE_OK" in res.stdout, res.stdout[-2000:] + res.stderr[-6000:]
    m = re.search(r"(\d+) compiled", res.stdout)
    assert m and int(m.group(1)) >= 1000, res.stdout


def test_planted_out_of_bounds_read_is_reported(tmp_path: Path) -> None:
    src = tmp_path / "oob.cpp"
    src.write_text("#include <vector>\nint main(int argc, char **) { std::vector<int> v(4, 1); "
                   "return v.data()[3 + argc]; }\n")
    exe = tmp_path / "oob"
    subprocess.run(["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", str(src), "-o",
                    str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, env=ENV)
    assert res.returncode != 0 and "AddressSanitizer" in res.stderr
