"""The oracle's pathwise gradient, pinned without the GPU.

``OracleCtx.gradient`` differentiates the reference's propagation loop
(_core.cpp:332-350) along the winning path of every event.  These tests pin it
against what the mathematics fixes:

* central finite differences of the oracle's own propagation (``orc_propagate``,
  itself pinned bit for bit against the compiled reference), on random DAGs with
  continuous times, where the objective is piecewise linear and almost surely
  smooth at the evaluation point: caps, stale reads (repeated precedence
  targets), one link read by several entries, repeated sources, edges outside
  the link table, gaps, links with and without distributions;
* closed forms on a chain and on a capped diamond.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from dags import STD_DISTS, continuous_dag
from mc_dagprop_amd.workloads import FlatDag


def durations_of(dmin: np.ndarray, X: np.ndarray, rel: np.ndarray, has_dist: np.ndarray) -> np.ndarray:
    """Per-sample durations as the sampler forms them: d + X*d (relative), d + X (empirical absolute), d (none).
    Multiply, then add: two roundings."""
    extra = np.where(rel == 1, X * dmin, np.where(has_dist, X, 0.0))
    return dmin + extra


def objective(ctx: oracle.OracleCtx, D: np.ndarray, w: np.ndarray, earliest: np.ndarray) -> float:
    r, _ = ctx.propagate(D)
    return float(np.sum(np.maximum(0.0, r - earliest[None, :]) * w[None, :]))


def setup(seed: int, n_events: int, n: int, max_delay: float, shared_links: bool = True):
    dag = continuous_dag(seed, n_events, max_delay=max_delay, shared_links=shared_links)
    ctx = oracle.OracleCtx(dag)
    rng = np.random.default_rng(seed + 1)
    dmin, types = ctx.links()
    rel = ctx.link_relative()
    registered = {d.activity_type for d in dag.dists}
    has_dist = np.asarray([int(t) in registered for t in types])
    X = rng.uniform(0.0, 1.5, (n, ctx.L))
    w = rng.uniform(0.5, 2.0, ctx.E)
    D = durations_of(dmin[None, :], X, rel[None, :], has_dist[None, :])
    return dag, ctx, dmin, rel, has_dist, X, w, D


@pytest.mark.parametrize("seed,max_delay", [(1, 1e6), (2, 25.0), (3, 8.0), (4, 1e6), (5, 15.0)])
def test_gradient_matches_finite_differences(seed: int, max_delay: float) -> None:
    dag, ctx, dmin, rel, has_dist, X, w, D = setup(seed, 40, 24, max_delay)
    r, c = ctx.propagate(D)
    obj, gd, ge = ctx.gradient(r, c, D, w)
    earliest = np.asarray(dag.earliest, dtype=np.float64)
    assert obj == pytest.approx(objective(ctx, D, w, earliest), rel=1e-13)
    h = 1e-6
    for l in range(ctx.L):
        up, dn = dmin.copy(), dmin.copy()
        up[l] += h
        dn[l] -= h
        fd = (objective(ctx, durations_of(up[None, :], X, rel[None, :], has_dist[None, :]), w, earliest)
              - objective(ctx, durations_of(dn[None, :], X, rel[None, :], has_dist[None, :]), w, earliest)) / (2 * h)
        assert gd[l] == pytest.approx(fd, rel=1e-6, abs=2e-4), f"link {l}"
    for e in range(ctx.E):
        vals = []
        for sgn in (1, -1):
            d2 = FlatDag(**{**dag.__dict__})
            d2.earliest = earliest.copy()
            d2.earliest[e] += sgn * h
            c2 = oracle.OracleCtx(d2)
            vals.append(objective(c2, D, w, d2.earliest))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert ge[e] == pytest.approx(fd, rel=1e-6, abs=2e-4), f"event {e}"


def test_case_mix_is_exercised() -> None:
    """The finite-difference DAGs hit every branch: caps, cause -1, stale winners, repeated sources, shared links,
    edges outside the table."""
    caps = none = stale_win = 0
    for seed, md in [(1, 1e6), (2, 25.0), (3, 8.0), (4, 1e6), (5, 15.0)]:
        dag, ctx, dmin, rel, has_dist, X, w, D = setup(seed, 40, 24, md)
        r, c = ctx.propagate(D)
        up = np.asarray(dag.earliest)[None, :] + dag.max_delay
        caps += int(np.sum((r == up) & (c >= 0)))
        none += int(np.sum(c < 0))
        order = ctx.order
        pos = np.empty(ctx.E, dtype=np.int64)
        pos[order] = np.arange(ctx.E)
        ev = np.nonzero(c >= 0)
        stale_win += int(np.sum(pos[c[ev]] > pos[ev[1]]))
    assert caps > 0 and none > 0 and stale_win > 0
    d = continuous_dag(1, 40, max_delay=1e6, shared_links=True)
    assert np.any(d.prec_edge >= d.link_count) or np.any(d.prec_edge < 0)
    assert np.unique(d.prec_edge[(d.prec_edge >= 0) & (d.prec_edge < d.link_count)]).size < np.sum(
        (d.prec_edge >= 0) & (d.prec_edge < d.link_count))


def test_chain_closed_form() -> None:
    """0 -> 1 -> 2, relative delays: J = r1 + r2 = 2 d0' + d1' with d' = d (1 + X)."""
    events = [(0.0, 1e9, 0.0)] * 3
    acts = {(0, 1): (0, 4.0, 1), (1, 2): (1, 6.0, 1)}
    prec = [(1, [(0, 0)]), (2, [(1, 1)])]
    dag = FlatDag.from_lists(events, acts, prec, 1e6, STD_DISTS)  # type 1: exponential (relative)
    ctx = oracle.OracleCtx(dag)
    X = np.array([[0.25, 0.5]])
    D = np.array([[4.0 + 0.25 * 4.0, 6.0 + 0.5 * 6.0]])
    r, c = ctx.propagate(D)
    obj, gd, ge = ctx.gradient(r, c, D, np.ones(3))
    assert obj == 2 * 5.0 + 9.0
    np.testing.assert_allclose(gd, [2 * (1 + X[0, 0]), 1 + X[0, 1]], rtol=1e-15)
    np.testing.assert_allclose(ge, [2.0, -1.0, -1.0], rtol=0, atol=0)


def test_capped_diamond_closed_form() -> None:
    """0 -> {1, 2} -> 3 with the cap binding at 3: event 3's time is earliest + max_delay, so its whole
    adjoint goes to its own earliest time and nothing flows to 1, 2 or their links."""
    events = [(0.0, 1e9, 0.0), (0.0, 1e9, 0.0), (0.0, 1e9, 0.0), (1.0, 1e9, 0.0)]
    acts = {(0, 1): (0, 3.0, -1), (0, 2): (1, 5.0, -1), (1, 3): (2, 4.0, -1), (2, 3): (3, 1.0, -1)}
    prec = [(1, [(0, 0)]), (2, [(0, 1)]), (3, [(1, 2), (2, 3)])]
    dag = FlatDag.from_lists(events, acts, prec, 2.0, [])
    ctx = oracle.OracleCtx(dag)
    D = np.array([[3.0, 5.0, 4.0, 1.0]])
    r, c = ctx.propagate(D)
    # r1 = min(0 + 2, 3) = 2 (capped), r2 = 2 (capped), r3 = 3 (capped: 1 + 2)
    np.testing.assert_array_equal(r[0], [0.0, 2.0, 2.0, 3.0])
    w = np.array([1.0, 2.0, 3.0, 4.0])
    obj, gd, ge = ctx.gradient(r, c, D, w)
    assert obj == 2 * 2.0 + 3 * 2.0 + 4 * 2.0
    np.testing.assert_array_equal(gd, [0.0, 0.0, 0.0, 0.0])
    # each capped event: -w (explicit) + w (through the cap) = 0
    np.testing.assert_array_equal(ge, [0.0, 0.0, 0.0, 0.0])


def test_zero_minimal_duration_needs_the_unit_draw() -> None:
    """A relative link with minimal duration 0: the multiplier 1 + X comes from the duration at d = 1."""
    events = [(0.0, 1e9, 0.0)] * 2
    acts = {(0, 1): (0, 0.0, 1)}
    dag = FlatDag.from_lists(events, acts, [(1, [(0, 0)])], 1e6, STD_DISTS)
    ctx = oracle.OracleCtx(dag)
    D = np.array([[0.0]])
    r, c = ctx.propagate(D)
    with pytest.raises(oracle.OracleError):
        ctx.gradient(r, c, D, np.ones(2))
    obj, gd, ge = ctx.gradient(r, c, D, np.ones(2), mult1=np.array([[1.75]]))
    # r1 = 0 = earliest: inactive, so no adjoint reaches the link
    assert obj == 0.0 and gd[0] == 0.0
    events = [(0.0, 1e9, 0.0), (-1.0, 1e9, 0.0)]
    dag = FlatDag.from_lists(events, acts, [(1, [(0, 0)])], 1e6, STD_DISTS)
    ctx = oracle.OracleCtx(dag)
    r, c = ctx.propagate(D)
    obj, gd, ge = ctx.gradient(r, c, D, np.ones(2), mult1=np.array([[1.75]]))
    assert obj == 1.0 and gd[0] == 1.75


def test_inconsistent_inputs_rejected() -> None:
    dag, ctx, dmin, rel, has_dist, X, w, D = setup(1, 30, 4, 1e6)
    r, c = ctx.propagate(D)
    c2 = c.copy()
    j = np.nonzero(c2[0] >= 0)[0][0]
    c2[0, j] = -1
    with pytest.raises(oracle.OracleError):
        ctx.gradient(r, c2, D, w)
