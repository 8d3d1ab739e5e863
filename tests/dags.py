"""Test DAGs: the reference test suite's DAGs and a generator of edge-case DAGs.

Reference DAGs are transcribed from WonJayne/mc_dagprop's tests (file:line on each).
"""

from __future__ import annotations

import numpy as np

from mc_dagprop_amd.workloads import (
    CONSTANT,
    EMPIRICAL_ABSOLUTE,
    EMPIRICAL_RELATIVE,
    EXPONENTIAL,
    GAMMA,
    DistSpec,
    FlatDag,
)

INF = float("inf")


def sim6(dists=()) -> FlatDag:
    """test/test_simulator.py:11-35 (6 events, 5 links, event 5 isolated, type 2 unregistered)."""
    events = [(0.0, 100.0, 0.0), (5.0, 100.0, 0.0), (10.0, 100.0, 0.0), (22.0, 100.0, 0.0), (20.0, 100.0, 0.0),
              (100.0, 100.0, 0.0)]
    acts = {(0, 1): (0, 3.0, 1), (1, 2): (1, 5.0, 1), (1, 3): (2, 5.0, 1), (2, 4): (3, 15.0, 2), (3, 4): (4, 10.0, 3)}
    prec = [(1, [(0, 0)]), (2, [(1, 1)]), (3, [(1, 2)]), (4, [(2, 3), (3, 4)])]
    return FlatDag.from_lists(events, acts, prec, 1e6, dists)


def sim6_reversed(dists=()) -> FlatDag:
    """test/test_simulator.py:64-68: the same DAG with the precedence list reversed."""
    d = sim6(dists)
    events = list(zip(d.earliest, d.latest, d.actual))
    acts = {(0, 1): (0, 3.0, 1), (1, 2): (1, 5.0, 1), (1, 3): (2, 5.0, 1), (2, 4): (3, 15.0, 2), (3, 4): (4, 10.0, 3)}
    prec = list(reversed([(1, [(0, 0)]), (2, [(1, 1)]), (3, [(1, 2)]), (4, [(2, 3), (3, 4)])]))
    return FlatDag.from_lists(events, acts, prec, 1e6, dists)


def chain3(dists=(), max_delay=1e6) -> FlatDag:
    """test/test_monte_carlo_extra.py:9-20."""
    events = [(0.0, 100.0, 0.0)] * 3
    acts = {(0, 1): (0, 1.0, 1), (1, 2): (1, 2.0, 1)}
    prec = [(1, [(0, 0)]), (2, [(1, 1)])]
    return FlatDag.from_lists(events, acts, prec, max_delay, dists)


def large_chain(dists=()) -> FlatDag:
    """test/test_simulator.py:175-181: 10,000-event chain whose last entry names edge 9999 (one past the links)."""
    n = 10_000
    events = [(float(i), 100.0 + i, 0.0) for i in range(n)]
    acts = {(i, i + 1): (i, 3.0, 1) for i in range(n - 1)}
    prec = [(i, [(i - 1, i)]) for i in range(1, n)]
    return FlatDag.from_lists(events, acts, prec, 1e6, dists)


def emp_hist_weights_seed7() -> tuple[np.ndarray, np.ndarray]:
    """test/test_simulator.py:157-166: 1000-bin histogram of truncated Exp(3) draws (numpy legacy seed 7)."""
    rs = np.random.RandomState(7)
    values = []
    while len(values) < 1_000_000:
        value = 10.0
        while value > 5.0:
            value = rs.exponential(3.0)
        values.append(value)
    hist, edges = np.histogram(values, bins=1000, density=True)
    return 0.5 * (edges[:-1] + edges[1:]), hist


def parity_chain() -> FlatDag:
    """test/test_parity_suite.py:47-88 (Monte Carlo half)."""
    events = [(0.0, 120.0, 0.0)] * 3
    acts = {(0, 1): (0, 0.0, 1), (1, 2): (1, 0.0, 2)}
    prec = [(1, [(0, 0)]), (2, [(1, 1)])]
    dists = [
        DistSpec(1, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 1, 2, 3]), weights=np.array([0.1, 0.3, 0.4, 0.2])),
        DistSpec(2, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 1, 2, 3]), weights=np.array([0.15, 0.35, 0.3, 0.2])),
    ]
    return FlatDag.from_lists(events, acts, prec, 120.0, dists)


def parity_branching() -> FlatDag:
    """test/test_parity_suite.py:91-135 (Monte Carlo half)."""
    events = [(0.0, 200.0, 0.0)] * 4
    acts = {(0, 1): (0, 0.0, 1), (0, 2): (1, 0.0, 2), (1, 3): (2, 0.0, 2), (2, 3): (3, 0.0, 1)}
    prec = [(1, [(0, 0)]), (2, [(0, 1)]), (3, [(1, 2), (2, 3)])]
    dists = [
        DistSpec(1, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 1, 2, 3, 4]),
                 weights=np.array([0.15, 0.2, 0.25, 0.25, 0.15])),
        DistSpec(2, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 1, 2, 3]), weights=np.array([0.2, 0.35, 0.3, 0.15])),
    ]
    return FlatDag.from_lists(events, acts, prec, 200.0, dists)


def parity_long_chain(n: int = 100) -> FlatDag:
    """test/test_parity_suite.py:138-179 (Monte Carlo half)."""
    events = [(0.0, 1000.0, 0.0)] * (n + 1)
    acts = {(i, i + 1): (i, 0.0, 1) for i in range(n)}
    prec = [(i + 1, [(i, i)]) for i in range(n)]
    dists = [DistSpec(1, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 1, 2]), weights=np.array([0.2, 0.5, 0.3]))]
    return FlatDag.from_lists(events, acts, prec, 1000.0, dists)


def parity_cap() -> FlatDag:
    """test/test_parity_suite.py:182-215 (Monte Carlo half): max_delay caps event 1 at 16."""
    events = [(10.0, 100.0, 10.0), (12.0, 80.0, 12.0)]
    acts = {(0, 1): (0, 0.0, 1)}
    prec = [(1, [(0, 0)])]
    dists = [DistSpec(1, EMPIRICAL_ABSOLUTE, values=np.array([3.0, 4, 5, 6]), weights=np.array([0.25, 0.25, 0.3, 0.2]))]
    return FlatDag.from_lists(events, acts, prec, 4.0, dists)


def random_dag(rng: np.random.Generator, n_events: int, max_preds: int = 4, *, duplicates: bool = False,
               gaps: bool = False, oob_edges: bool = False, stale_free: bool = True, n_types: int = 4,
               max_delay: float | None = None, dists=()) -> FlatDag:
    """A random DAG over a random topological permutation.

    Edges point from earlier to later events of the permutation; some events
    have no predecessors; predecessor lists are shuffled.  Options add the
    reference's undefined or unusual inputs: duplicate precedence entries for
    one target, gaps in activity indices, edge indices outside the link range.
    """
    perm = rng.permutation(n_events)
    earliest = np.round(rng.uniform(0, 50, n_events))  # integer times => many exact ties
    acts = {}
    prec = []
    idx = 0
    for pos in range(1, n_events):
        e = int(perm[pos])
        if rng.random() < 0.15:
            continue
        k = int(rng.integers(1, max_preds + 1))
        cand = perm[:pos]
        srcs = rng.choice(cand, size=min(k, pos), replace=False)
        preds = []
        for s in srcs:
            s = int(s)
            if (s, e) in acts:
                continue
            step = int(rng.integers(2, 4)) if gaps and rng.random() < 0.2 else 1
            idx += step
            a = idx - 1
            acts[(s, e)] = (a, float(np.round(rng.uniform(0, 20))), int(rng.integers(-1, n_types)))
            preds.append((s, a))
        if oob_edges and preds and rng.random() < 0.1:
            preds.append((int(cand[0]), idx + 50 if rng.random() < 0.5 else -3))
        prec.append((e, preds))
    if duplicates and prec:
        for _ in range(max(1, len(prec) // 10)):
            j = int(rng.integers(0, len(prec)))
            tgt, preds = prec[j]
            sub = [p for p in preds if rng.random() < 0.5]
            prec.insert(int(rng.integers(0, len(prec) + 1)), (tgt, sub))
    order = rng.permutation(len(prec))
    prec = [prec[i] for i in order]
    if max_delay is None:
        max_delay = float(rng.choice([0.0, 5.0, 30.0, 1e6]))
    events = [(float(x), float(x) + 100.0, 0.0) for x in earliest]
    return FlatDag.from_lists(events, acts, prec, max_delay, list(dists))


def special_durations(rng: np.random.Generator, n: int, L: int) -> np.ndarray:
    """Injected durations with ties, zeros of both signs, NaN, infinities and negatives."""
    d = np.round(rng.uniform(-2, 20, (n, L)))
    pick = rng.random((n, L))
    d[pick < 0.03] = np.nan
    d[(pick >= 0.03) & (pick < 0.05)] = np.inf
    d[(pick >= 0.05) & (pick < 0.07)] = -np.inf
    d[(pick >= 0.07) & (pick < 0.12)] = -0.0
    d[(pick >= 0.12) & (pick < 0.2)] = 0.0
    return d


STD_DISTS = [
    DistSpec(0, CONSTANT, (1.5, 0.0, INF)),
    DistSpec(1, EXPONENTIAL, (0.3, INF, INF)),
    DistSpec(2, GAMMA, (2.0, 0.2, INF)),
    DistSpec(3, EMPIRICAL_RELATIVE, values=np.array([0.0, 0.5, 1.0, 2.0]), weights=np.array([1.0, 2.0, 0.0, 1.0])),
]


GRAD_DISTS = STD_DISTS + [DistSpec(4, EMPIRICAL_ABSOLUTE, values=np.array([0.0, 2.0, 5.0]), weights=np.array([1.0, 1.0, 2.0]))]


def continuous_dag(seed: int, n_events: int, *, max_delay: float, shared_links: bool) -> FlatDag:
    """random_dag's structure (repeated targets, gaps, out-of-range edges) with continuous times, plus entries
    that repeat a source with another link and entries that read a link another entry reads too."""
    rng = np.random.default_rng(seed)
    d = random_dag(rng, n_events, duplicates=True, gaps=True, oob_edges=True, n_types=5, max_delay=max_delay,
                   dists=GRAD_DISTS)
    d.earliest = rng.uniform(0.0, 40.0, d.earliest.size)
    d.act_dur = rng.uniform(1.0, 20.0, d.act_dur.size)
    if shared_links and d.act_idx.size:
        src, edge = list(d.prec_src), list(d.prec_edge)
        off = list(d.prec_off)
        for j in range(len(d.prec_target)):
            a, b = off[j], off[j + 1]
            if b > a and rng.random() < 0.3:  # repeat this entry's first source with some other link
                ins = b
                src.insert(ins, src[a])
                edge.insert(ins, int(rng.choice(d.act_idx)))
                off = off[: j + 1] + [o + 1 for o in off[j + 1:]]
        d.prec_src = np.asarray(src, dtype=np.int32)
        d.prec_edge = np.asarray(edge, dtype=np.int32)
        d.prec_off = np.asarray(off, dtype=np.int64)
    return d


def long_lived_slots(seed: int, n: int = 260, fan_in: int = 20, max_delay: float = 30.0) -> FlatDag:
    """A reverse-walk stress DAG: n sources s_i, a hub m (predecessor s_0), n events t_i
    (predecessors s_i and m) and a sink z (predecessors t_0 .. t_{fan_in - 1}).

    The t_i's adjoints stay live from their own step to the sources' step, more slots than a
    workgroup's LDS holds at one sample per lane, so some live in global memory; the hub has n
    successors and the sink fan_in predecessors, so their per-event records run past the 64 dwords
    a wave reads with one vector load."""
    rng = np.random.default_rng(seed)
    hub, sink = n, 2 * n + 1
    t = [n + 1 + i for i in range(n)]
    earliest = np.concatenate([rng.uniform(0.0, 10.0, n), [20.0], rng.uniform(30.0, 45.0, n), [80.0]])
    events = [(float(x), float(x) + 200.0, 0.0) for x in earliest]
    acts, prec = {}, []
    dur = lambda: float(rng.uniform(1.0, 20.0))  # noqa: E731
    acts[(0, hub)] = (n, dur(), 1)
    prec.append((hub, [(0, n)]))
    for i in range(n):
        acts[(i, t[i])] = (i, dur(), i % 5)
        acts[(hub, t[i])] = (n + 1 + i, dur(), (i + 2) % 5)
        prec.append((t[i], [(i, i), (hub, n + 1 + i)]))
    zp = []
    for j in range(fan_in):
        acts[(t[j], sink)] = (2 * n + 1 + j, dur(), j % 5)
        zp.append((t[j], 2 * n + 1 + j))
    prec.append((sink, zp))
    return FlatDag.from_lists(events, acts, prec, max_delay, list(GRAD_DISTS))


# ---- fixture (de)serialisation ------------------------------------------------
_ARRAYS = ("earliest", "latest", "actual", "act_src", "act_dst", "act_idx", "act_dur", "act_type", "prec_target",
           "prec_off", "prec_src", "prec_edge")


def dag_to_npz_dict(dag: FlatDag) -> dict:
    out = {f"dag_{k}": getattr(dag, k) for k in _ARRAYS}
    out["dag_max_delay"] = np.float64(dag.max_delay)
    off = [0]
    vals, wts = [], []
    for d in dag.dists:
        if d.values is not None:
            vals.extend(np.asarray(d.values, dtype=np.float64))
            wts.extend(np.asarray(d.weights, dtype=np.float64))
        off.append(len(vals))
    out["dist_type"] = np.asarray([d.activity_type for d in dag.dists], dtype=np.int32)
    out["dist_kind"] = np.asarray([d.kind for d in dag.dists], dtype=np.int32)
    out["dist_params"] = np.asarray([d.params for d in dag.dists], dtype=np.float64).reshape(-1, 3)
    out["dist_off"] = np.asarray(off, dtype=np.int64)
    out["dist_values"] = np.asarray(vals, dtype=np.float64)
    out["dist_weights"] = np.asarray(wts, dtype=np.float64)
    return out


def dag_from_npz(z) -> FlatDag:
    kw = {k: z[f"dag_{k}"] for k in _ARRAYS}
    dists = []
    for i in range(z["dist_type"].shape[0]):
        a, b = int(z["dist_off"][i]), int(z["dist_off"][i + 1])
        kind = int(z["dist_kind"][i])
        has_table = kind in (EMPIRICAL_ABSOLUTE, EMPIRICAL_RELATIVE)
        dists.append(DistSpec(int(z["dist_type"][i]), kind, tuple(float(x) for x in z["dist_params"][i]),
                              z["dist_values"][a:b] if has_table else None,
                              z["dist_weights"][a:b] if has_table else None))
    return FlatDag(max_delay=float(z["dag_max_delay"]), dists=dists, **kw)
