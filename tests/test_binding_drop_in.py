"""The binding as a drop-in for the reference's own extension module.

A maintainer replaces the extension built from src/mc_dagprop/monte_carlo/_core.cpp with
mc_dagprop_amd/csrc/pybind_core.cpp, compiled as ``mc_dagprop.monte_carlo._core`` with
``-DMCDP_TYPES_MODULE="mc_dagprop.types"`` (the module the reference imports at _core.cpp:453).
This test does exactly that in a temporary package tree that holds only an ``mc_dagprop`` package
(its own ``types`` module and a ``monte_carlo/__init__.py`` re-exporting ``_core``, as
src/mc_dagprop/monte_carlo/__init__.py:3-5 does), imports it in a fresh interpreter that cannot see
``mc_dagprop_amd``, and checks construction, reprs, the frozen-dataclass decoration and that the
annotations are the objects of ``mc_dagprop.types``.  CPU only: running needs no device here, and
a run without one must raise RuntimeError (no CPU fallback).
"""

from __future__ import annotations

import subprocess
import sys
import sysconfig
import textwrap
from pathlib import Path

import pybind11
import pytest

from mc_dagprop_amd import types as amd_types

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "mc_dagprop_amd"
LIB = PKG / "libmcdp.so"

_BASES = {"Second": "float", "ProbabilityMass": "float", "ActivityIndex": "int", "EventIndex": "int",
          "ActivityType": "int", "EventId": "str"}

_CHECK = r"""
import sys
import dataclasses
import mc_dagprop.types as T
from mc_dagprop.monte_carlo import (Activity, DagContext, Event, EventTimestamp, GenericDelayGenerator,
                                    MonteCarloPropagator, Simulator, SimResult)
from mc_dagprop.monte_carlo import _core

assert "mc_dagprop_amd" not in sys.modules, "the drop-in module pulled in mc_dagprop_amd"
assert _core.__name__ == "mc_dagprop.monte_carlo._core", _core.__name__
assert Simulator is MonteCarloPropagator

ts = EventTimestamp(1.0, 2.0, 3.0)
assert repr(ts) == "EventTimestamp(earliest=1.0, latest=2.0, actual=3.0)", repr(ts)
ev = Event("a", ts)
act = Activity(idx=0, minimal_duration=5.0, activity_type=1)
for cls in (EventTimestamp, Event, Activity, DagContext):
    assert dataclasses.is_dataclass(cls), cls
assert EventTimestamp.__annotations__ == {"earliest": T.Second, "latest": T.Second, "actual": T.Second}
assert Event.__annotations__["event_id"] is T.EventId and Event.__annotations__["timestamp"] is EventTimestamp
assert Activity.__annotations__ == {"idx": T.ActivityIndex, "minimal_duration": T.Second,
                                    "activity_type": T.ActivityType}
try:
    act.idx = 3
except (AttributeError, dataclasses.FrozenInstanceError):
    pass
else:
    raise AssertionError("Activity is not frozen")

events = [Event("0", EventTimestamp(0.0, 100.0, 0.0)), Event("1", EventTimestamp(5.0, 100.0, 0.0))]
ctx = DagContext(events=events, activities={(0, 1): act}, precedence_list=[(1, [(0, 0)])], max_delay=10.0)
assert ctx.max_delay == 10.0 and len(ctx.events) == 2 and ctx.precedence_list == [(1, [(0, 0)])]
gen = GenericDelayGenerator()
gen.add_constant(activity_type=1, factor=0.5)
sim = Simulator(ctx, gen)
assert (sim.node_count(), sim.activity_count()) == (2, 1)
try:
    GenericDelayGenerator().add_empirical_absolute(activity_type=0, values=[1.0, 2.0], weights=[1.0])
except RuntimeError:
    pass
else:
    raise AssertionError("length mismatch accepted")
if not HAS_GPU:
    try:
        sim.run(seed=1)
    except RuntimeError:
        pass
    else:
        raise AssertionError("a run without a device did not raise")
print("DROP_IN_OK")
"""


def _build(tmp: Path) -> Path:
    pkg = tmp / "mc_dagprop"
    (pkg / "monte_carlo").mkdir(parents=True)
    (pkg / "__init__.py").write_text("")
    lines = ["from typing import NewType", ""]
    lines += [f'{name} = NewType("{name}", {_BASES[name]})' for name in amd_types.__all__]
    (pkg / "types.py").write_text("\n".join(lines) + "\n")
    (pkg / "monte_carlo" / "__init__.py").write_text(textwrap.dedent("""\
        from ._core import (Activity, DagContext, Event, EventTimestamp, GenericDelayGenerator,
                            MonteCarloPropagator, SimResult)

        Simulator = MonteCarloPropagator
        """))
    out = pkg / "monte_carlo" / ("_core" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["g++", "-O0", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           '-DMCDP_TYPES_MODULE="mc_dagprop.types"', f"-I{pybind11.get_include()}",
           f"-I{sysconfig.get_paths()['include']}", f"-I{ROOT / 'include'}",
           str(PKG / "csrc" / "pybind_core.cpp"), f"-L{PKG}", "-lmcdp", f"-Wl,-rpath,{PKG}", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return out


def test_binding_builds_and_imports_as_the_reference_module(tmp_path: Path, has_gpu: bool) -> None:
    if not LIB.exists():
        pytest.fail("libmcdp.so is not built (run __graft_entry__.build())")
    _build(tmp_path)
    code = f"HAS_GPU = {bool(has_gpu)}\n" + _CHECK
    env = {"PATH": "/usr/bin:/bin", "PYTHONPATH": str(tmp_path), "HOME": str(tmp_path)}
    res = subprocess.run([sys.executable, "-I", "-c", f"import sys; sys.path.insert(0, {str(tmp_path)!r})\n" + code],
                         capture_output=True, text=True, cwd=str(tmp_path), env=env, timeout=300)
    assert res.returncode == 0 and "DROP_IN_OK" in res.stdout, res.stdout[-2000:] + res.stderr[-4000:]
