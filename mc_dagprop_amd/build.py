"""Builds the native parts of mc_dagprop_amd in place.

* ``libmcdp.so``: the gfx950 kernels (``csrc/mcdp_kernels.hip``, ``csrc/mcdp_stats.hip``,
  ``csrc/mcdp_grad.hip``, ``csrc/mcdp_noalloc.hip``; hipcc) and the host compiler + C ABI (``csrc/mcdp_host.cpp``, ``csrc/mcdp_stats.cpp``,
  ``csrc/mcdp_batch.cpp``, ``csrc/mcdp_grad.cpp``, ``csrc/mcdp_serve.cpp``; g++),
  declared in
  ``include/mcdp.h``.
* ``monte_carlo/_core*.so``: the pybind11 module over that C ABI.

Kernel flags:
  -ffp-contract=off       every FP64 multiply and add rounds separately, as the
                          reference's duration and propagation arithmetic does;
  -mllvm -disable-machine-licm
                          keeps FP64 polynomial constants and Philox round keys
                          from being hoisted out of the predecessor loop, which
                          otherwise costs ~60 VGPRs and spills;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                          no packed FP32 instructions.

Run ``python mc_dagprop_amd/build.py`` (a script: importing the package needs the built files).
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

LIB = PKG / "libmcdp.so"
EXT = PKG / "monte_carlo" / ("_core" + sysconfig.get_config_var("EXT_SUFFIX"))

KERNEL_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-mllvm",
    "-disable-machine-licm",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(str(c) for c in cmd)}")


def _pybind_includes() -> list[str]:
    import pybind11

    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def build(verbose: bool = False) -> None:
    """Compiles libmcdp.so and the _core extension (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    hipcc = str(ROCM / "bin" / "hipcc")
    kernels = ["mcdp_kernels.hip", "mcdp_stats.hip", "mcdp_grad.hip", "mcdp_noalloc.hip"]
    hosts = ["mcdp_host.cpp", "mcdp_stats.cpp", "mcdp_batch.cpp", "mcdp_grad.cpp", "mcdp_serve.cpp"]
    objs = [BUILD / (Path(f).stem + ("_dev.o" if f.endswith(".hip") else "_host.o")) for f in kernels + hosts]
    steps = [[hipcc, *KERNEL_FLAGS, "-c", CSRC / f, "-o", o] for f, o in zip(kernels, objs)]
    steps += [["g++", *HOST_FLAGS, "-c", CSRC / f, "-o", o] for f, o in zip(hosts, objs[len(kernels):])]
    with ThreadPoolExecutor(max_workers=len(steps)) as pool:
        list(pool.map(lambda c: _run(c, verbose), steps))
    tmp_lib = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib], verbose)
    os.replace(tmp_lib, LIB)
    tmp_ext = EXT.with_name(EXT.name + ".tmp")
    _run(
        [
            "g++",
            "-O2",
            "-std=c++17",
            "-shared",
            "-fPIC",
            "-fvisibility=hidden",
            *_pybind_includes(),
            CSRC / "pybind_core.cpp",
            f"-L{PKG}",
            "-lmcdp",
            "-Wl,-rpath,$ORIGIN/..",
            "-o",
            tmp_ext,
        ],
        verbose,
    )
    os.replace(tmp_ext, EXT)


if __name__ == "__main__":
    build(verbose=True)
