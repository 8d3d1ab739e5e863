"""Writes the golden fixtures of this directory from the compiled reference engine.

Run in a checkout where the reference sources exist and ``make -C oracle ref`` has
built ``oracle/_ref/libref_mcdp.so`` (the reference's own _core.cpp, compiled in
place).  Each ``<case>.npz`` holds one DAG (flat arrays + distributions), the
seeds, and what the reference's Simulator::run returned for them, seed by seed in
order on one Simulator (as run_many does): realized (n, E), durations (n, L),
cause (n, E).  ``MANIFEST.txt`` lists the sha256 of every fixture.

    python tests/golden/make_golden.py            # every case
    python tests/golden/make_golden.py NAME ...   # only these cases (their manifest lines are replaced)
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import dags as D  # noqa: E402
from mc_dagprop_amd.workloads import (  # noqa: E402
    CONSTANT,
    EMPIRICAL_ABSOLUTE,
    EMPIRICAL_RELATIVE,
    EXPONENTIAL,
    GAMMA,
    DistSpec,
    layered_network,
    layered_small,
)
from oracle import RefSim  # noqa: E402

INF = float("inf")


def cases():
    c = D.DistSpec
    yield "sim6_constant", D.sim6([c(1, CONSTANT, (1.0, 0, INF))]), [7, 1, 2, 3]
    yield "sim6_propagation", D.sim6([c(1, CONSTANT, (1.0, 0, INF)), c(3, CONSTANT, (3.0, 0, INF))]), list(range(5))
    yield "sim6_unsorted", D.sim6_reversed([c(1, CONSTANT, (1.0, 0, INF))]), [7]
    yield "sim6_exponential", D.sim6([c(1, EXPONENTIAL, (1000.0, 1.0, INF))]), list(range(3))
    yield "sim6_empirical_absolute", D.sim6([c(1, EMPIRICAL_ABSOLUTE, values=np.array([10.0, 20, 40, 50]),
                                               weights=np.array([0.1, 0.2, 0.3, 0.4]))]), [7]
    yield "sim6_empirical_relative", D.sim6([c(1, EMPIRICAL_RELATIVE, values=np.array([1.2, 1.3, 1.35, 4.5]),
                                               weights=np.array([0.1, 0.2, 0.3, 0.4]))]), [7]
    centers, hist = D.emp_hist_weights_seed7()
    yield "sim6_empirical_relative_hist", D.sim6([c(1, EMPIRICAL_RELATIVE, values=centers, weights=hist)]), [7]
    yield "chain3_constant", D.chain3([c(1, CONSTANT, (1.0, 0, INF))]), [42]
    yield "chain3_exponential", D.chain3([c(1, EXPONENTIAL, (2.0, 0.5, INF))]), [0]
    yield "chain3_gamma", D.chain3([c(1, GAMMA, (2.0, 1.0, 0.5))]), [1]
    yield "chain3_run_many", D.chain3([c(1, EXPONENTIAL, (1.0, 0.5, INF))]), list(range(5))
    yield "parity_chain", D.parity_chain(), list(range(200))
    yield "parity_branching", D.parity_branching(), list(range(200))
    yield "parity_long_chain", D.parity_long_chain(), list(range(32))
    yield "parity_cap", D.parity_cap(), list(range(200))
    yield "layered_small", layered_small(), list(range(64))
    # 2,000-event slices of the network workload's generator (20 of its 200 layers): the literal eight
    # activity types, whose extras multiply minimal durations of 30-600 s and saturate most events at their
    # cap, and a calibrated variant (exponential means, gamma scales and relative factors divided by 315,
    # the mean minimal duration) where caps bind less often
    yield "network_slice", layered_network(layers=20), list(range(8))
    calib = layered_network(layers=20)
    scaled = []
    for d in calib.dists:
        if d.kind in (EXPONENTIAL,):
            d = DistSpec(d.activity_type, d.kind, (d.params[0] / 315.0, d.params[1], d.params[2]))
        elif d.kind == GAMMA:
            d = DistSpec(d.activity_type, d.kind, (d.params[0], d.params[1] / 315.0, d.params[2]))
        elif d.kind == EMPIRICAL_RELATIVE:
            d = DistSpec(d.activity_type, d.kind, d.params, values=np.asarray(d.values) / 315.0, weights=d.weights)
        scaled.append(d)
    calib.dists = scaled
    yield "network_slice_calibrated", calib, list(range(100, 108))
    rng = np.random.default_rng(20260)
    for i, (n, kw) in enumerate([(40, {}), (60, {"duplicates": True}), (50, {"gaps": True, "duplicates": True}),
                                 (80, {"max_delay": 0.0}), (30, {"max_delay": 7.0, "gaps": True})]):
        yield f"zoo_{i}", D.random_dag(rng, n, dists=D.STD_DISTS, **kw), list(range(-8, 8))


def main() -> None:
    only = set(sys.argv[1:])
    manifest = HERE / "MANIFEST.txt"
    kept = {}
    if only and manifest.exists():
        for line in manifest.read_text().splitlines():
            if line.strip():
                kept[line.split()[1]] = line
    lines = []
    for name, dag, seeds in cases():
        if only and name not in only:
            if f"{name}.npz" in kept:
                lines.append(kept[f"{name}.npz"])
            continue
        ref = RefSim(dag)
        r, d, cz = ref.run(seeds)
        path = HERE / f"{name}.npz"
        np.savez_compressed(path, seeds=np.asarray(seeds, dtype=np.int32), realized=r, durations=d, cause=cz,
                            **D.dag_to_npz_dict(dag))
        lines.append(f"{hashlib.sha256(path.read_bytes()).hexdigest()}  {path.name}")
        print(name, r.shape, d.shape, flush=True)
    (HERE / "MANIFEST.txt").write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
