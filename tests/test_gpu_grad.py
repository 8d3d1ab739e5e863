"""The reverse walk (pathwise gradients) against the FP64 CPU gradient and finite differences.

* ``mcdp_gradient_host`` vs ``oracle.OracleCtx.gradient`` on the GPU's own run_many outputs: the
  same adjoints summed in another order, so within 1e-11 relative to the largest component;
* central finite differences with common random numbers (h = 1e-3 s, the same seeds at both
  points, the per-sample objective difference summed exactly): the gradient of the expectation
  the sampled objective estimates;
* properties: results independent of the chunking (bit for bit), of repeated calls, and
  ``set_times`` equivalent to compiling the new times.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import dags as D
import oracle
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import FlatDag, layered_small

pytestmark = pytest.mark.gpu


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def oracle_means(dag: FlatDag, prop: capi.Propagator, seeds, w, mult1=None):
    r, d, c = prop.run_many(seeds)
    ctx = oracle.OracleCtx(dag)
    obj, gd, ge = ctx.gradient(r, c, d, w, mult1=mult1)
    n = len(seeds)
    return obj / n, gd / n, ge / n


def assert_close(got, want, rtol: float = 1e-11) -> None:
    got, want = np.asarray(got, dtype=np.float64), np.asarray(want, dtype=np.float64)
    scale = max(1.0, float(np.max(np.abs(want))) if want.size else 1.0)
    np.testing.assert_allclose(got, want, rtol=rtol, atol=rtol * scale)


def with_times(dag: FlatDag, dmin_links: np.ndarray | None = None, earliest: np.ndarray | None = None) -> FlatDag:
    """The same DAG with new link minimal durations (through its activities) and earliest times."""
    d2 = FlatDag(**{**dag.__dict__})
    if dmin_links is not None:
        d2.act_dur = np.asarray([dmin_links[i] for i in dag.act_idx], dtype=np.float64)
    if earliest is not None:
        d2.earliest = np.asarray(earliest, dtype=np.float64).copy()
    return d2


CASES = [
    ("small", lambda: layered_small(), 3000),
    ("continuous_caps", lambda: D.continuous_dag(2, 60, max_delay=25.0, shared_links=True), 700),
    ("continuous_open", lambda: D.continuous_dag(4, 80, max_delay=1e6, shared_links=True), 513),
    ("random_stale", lambda: D.random_dag(np.random.default_rng(7), 120, duplicates=True, gaps=True, oob_edges=True,
                                          max_delay=30.0, dists=D.STD_DISTS), 257),
    ("sim6", lambda: D.sim6(D.STD_DISTS), 130),
    # slots in global memory; records past the one-load window (a hub of 260 successors, a sink of 20 inputs)
    ("long_lived_slots", lambda: D.long_lived_slots(3), 200),
]


@pytest.mark.parametrize("name,make,n", CASES, ids=[c[0] for c in CASES])
def test_gradient_matches_oracle(gpu, name, make, n) -> None:
    dag = make()
    prop = capi.Propagator(dag)
    rng = np.random.default_rng(11)
    w = rng.uniform(0.5, 2.0, prop.E)
    seeds = np.arange(1000, 1000 + n, dtype=np.int32)
    obj, gd, ge = prop.gradient(seeds, w)
    o_obj, o_gd, o_ge = oracle_means(dag, prop, seeds, w)
    assert_close([obj], [o_obj])
    assert_close(gd, o_gd)
    assert_close(ge, o_ge)


def test_case_mix_reaches_the_gpu(gpu) -> None:
    """The oracle cases above include stale winners, caps and repeated sources on the GPU's own samples."""
    dag = D.random_dag(np.random.default_rng(7), 120, duplicates=True, gaps=True, oob_edges=True, max_delay=30.0,
                       dists=D.STD_DISTS)
    prop = capi.Propagator(dag)
    assert prop.plan_info()["stale_preds"] > 0
    r, _, c = prop.run_many(np.arange(257, dtype=np.int32))
    up = np.asarray(dag.earliest)[None, :] + dag.max_delay
    assert np.sum((r == up) & (c >= 0)) > 0 and np.sum(c < 0) > 0


def test_chunking_and_repeats_are_bit_exact(gpu) -> None:
    dag = D.continuous_dag(3, 70, max_delay=15.0, shared_links=True)
    prop = capi.Propagator(dag)
    w = np.random.default_rng(3).uniform(0.5, 2.0, prop.E)
    seeds = np.arange(5000, dtype=np.int32)[::-1].copy()
    ref = prop.gradient(seeds, w)
    for chunk in (64, 2048, 4096, 0):
        got = prop.gradient(seeds, w, chunk=chunk)
        assert bits([got[0]]) == bits([ref[0]])
        np.testing.assert_array_equal(bits(got[1]), bits(ref[1]))
        np.testing.assert_array_equal(bits(got[2]), bits(ref[2]))


def test_several_reverse_launches_per_call(gpu) -> None:
    """A reverse call split into launches of 16 tiles (tile offsets, unit rows per launch) gives the sums of one
    launch bit for bit, also with unit-duration links."""
    dag = layered_small()
    prop = capi.Propagator(dag)
    dmin, e = prop.get_times()
    dmin[::4] = 0.0
    prop.set_times(dmin, e - 300.0)
    w = np.random.default_rng(8).uniform(0.5, 2.0, prop.E)
    seeds = np.arange(6000, dtype=np.int32)
    ref = prop.gradient(seeds, w)
    capi.check(capi.lib().mcdp_debug_reverse_tiles(prop.h, 3))  # rounds to one group of 16 tiles
    got = prop.gradient(seeds, w)
    capi.check(capi.lib().mcdp_debug_reverse_tiles(prop.h, 0))
    assert bits([got[0]]) == bits([ref[0]])
    np.testing.assert_array_equal(bits(got[1]), bits(ref[1]))
    np.testing.assert_array_equal(bits(got[2]), bits(ref[2]))


def test_set_times_matches_a_fresh_compile(gpu) -> None:
    dag = D.continuous_dag(5, 60, max_delay=20.0, shared_links=True)
    prop = capi.Propagator(dag)
    rng = np.random.default_rng(5)
    dmin0, e0 = prop.get_times()
    dmin1 = dmin0 * rng.uniform(0.5, 1.5, dmin0.size)
    e1 = e0 + rng.uniform(-3.0, 3.0, e0.size)
    seeds = np.arange(300, dtype=np.int32)
    prop.run_many(seeds)  # the old times are on the device now
    prop.set_times(dmin1, e1)
    d1, got_e = prop.get_times()
    np.testing.assert_array_equal(d1, dmin1)
    np.testing.assert_array_equal(got_e, e1)
    fresh = capi.Propagator(with_times(dag, dmin1, e1))
    for a, b in zip(prop.run_many(seeds), fresh.run_many(seeds)):
        np.testing.assert_array_equal(bits(a) if a.dtype == np.float64 else a, bits(b) if b.dtype == np.float64 else b)
    w = rng.uniform(0.5, 2.0, prop.E)
    g1, g2 = prop.gradient(seeds, w), fresh.gradient(seeds, w)
    np.testing.assert_array_equal(bits(g1[1]), bits(g2[1]))
    np.testing.assert_array_equal(bits(g1[2]), bits(g2[2]))


def test_zero_minimal_duration_draws_again(gpu) -> None:
    """Relative links with minimal duration 0: the multiplier is 1 + X, the duration the same draw gives at
    minimal duration 1 (taken from a run with every minimal duration set to 1)."""
    dag = layered_small()
    prop = capi.Propagator(dag)
    dmin, _ = prop.get_times()
    zero = np.zeros_like(dmin, dtype=bool)
    zero[::3] = True
    dmin[zero] = 0.0
    prop.set_times(dmin)
    seeds = np.arange(777, dtype=np.int32)
    w = np.random.default_rng(9).uniform(0.5, 2.0, prop.E)
    # earliest far in the past, so events are late and adjoints reach the zero links
    _, e = prop.get_times()
    prop.set_times(None, e - 500.0)
    obj, gd, ge = prop.gradient(seeds, w)
    unit = capi.Propagator(dag)
    unit.set_times(np.ones_like(dmin))
    _, mult1, _ = unit.run_many(seeds)
    d2 = with_times(dag, dmin, e - 500.0)
    o_obj, o_gd, o_ge = oracle_means(d2, prop, seeds, w, mult1=mult1)
    assert np.any(o_gd[zero] != 0.0)
    assert_close([obj], [o_obj])
    assert_close(gd, o_gd)
    assert_close(ge, o_ge)


def test_reverse_device_entry_point(gpu) -> None:
    """The device entry point: sums (not means) with a leading dimension above n, accumulate, and n = 0."""
    dag = D.continuous_dag(6, 50, max_delay=20.0, shared_links=True)
    prop = capi.Propagator(dag)
    E, L = prop.E, prop.L
    n, ld = 200, 256
    seeds_h = np.arange(200, dtype=np.int32)
    w_h = np.random.default_rng(6).uniform(0.5, 2.0, E)
    seeds = capi.DeviceBuffer.from_array(np.concatenate([seeds_h, np.zeros(ld - n, np.int32)]))
    w = capi.DeviceBuffer.from_array(w_h)
    real, dur, cause = capi.DeviceBuffer(8 * E * ld), capi.DeviceBuffer(8 * L * ld), capi.DeviceBuffer(4 * E * ld)
    obj, gd, ge = capi.DeviceBuffer(8), capi.DeviceBuffer(8 * L), capi.DeviceBuffer(8 * E)
    prop.run_many_device(seeds, n, ld, real, dur, cause)
    prop.reverse_device(seeds, w, real, dur, cause, n, ld, obj, gd, ge)
    capi.synchronize()
    m_obj, m_gd, m_ge = prop.gradient(seeds_h, w_h)
    assert_close(obj.to_array(np.float64, (1,)) / n, [m_obj])
    assert_close(gd.to_array(np.float64, (L,)) / n, m_gd)
    assert_close(ge.to_array(np.float64, (E,)) / n, m_ge)
    s1 = gd.to_array(np.float64, (L,))
    prop.reverse_device(seeds, w, real, dur, cause, n, ld, obj, gd, ge, accumulate=True)
    capi.synchronize()
    assert_close(gd.to_array(np.float64, (L,)), 2 * s1)
    prop.reverse_device(seeds, w, real, dur, cause, 0, ld, obj, gd, ge)
    capi.synchronize()
    assert not np.any(gd.to_array(np.float64, (L,))) and obj.to_array(np.float64, (1,))[0] == 0.0


def fd_objective_difference(prop_plus, prop_minus, seeds, w, e_plus, e_minus) -> float:
    """sum_s [J_s(+) - J_s(-)], each sample's difference formed first (common random numbers)."""
    rp, _, _ = prop_plus.run_many(seeds)
    rm, _, _ = prop_minus.run_many(seeds)
    lp = np.maximum(0.0, rp - e_plus[None, :]) * w[None, :]
    lm = np.maximum(0.0, rm - e_minus[None, :]) * w[None, :]
    return math.fsum((lp - lm).ravel())


def test_common_random_number_finite_differences(gpu) -> None:
    """Central differences, h = 1e-3 s, on the 200-event layered workload: each derivative of the sample
    mean equals the pathwise gradient up to samples whose winning entry changes inside [-h, h]."""
    dag = layered_small()
    prop = capi.Propagator(dag)
    rng = np.random.default_rng(21)
    w = rng.uniform(0.5, 2.0, prop.E)
    seeds = np.arange(4096, dtype=np.int32)
    n = seeds.size
    _, gd, ge = prop.gradient(seeds, w)
    dmin, e0 = prop.get_times()
    h = 1e-3
    plus, minus = capi.Propagator(dag), capi.Propagator(dag)
    links = rng.choice(np.nonzero(gd)[0], 12, replace=False)
    for l in links:
        dp, dm = dmin.copy(), dmin.copy()
        dp[l] += h
        dm[l] -= h
        plus.set_times(dp)
        minus.set_times(dm)
        fd = fd_objective_difference(plus, minus, seeds, w, e0, e0) / ((dp[l] - dm[l]) * n)
        assert fd == pytest.approx(gd[l], rel=2e-3, abs=1e-6), f"link {l}: fd {fd} vs {gd[l]}"
    plus.set_times(dmin)
    minus.set_times(dmin)
    events = rng.choice(prop.E, 12, replace=False)
    for e in events:
        ep, em = e0.copy(), e0.copy()
        ep[e] += h
        em[e] -= h
        plus.set_times(None, ep)
        minus.set_times(None, em)
        fd = fd_objective_difference(plus, minus, seeds, w, ep, em) / ((ep[e] - em[e]) * n)
        assert fd == pytest.approx(ge[e], rel=2e-3, abs=1e-6), f"event {e}: fd {fd} vs {ge[e]}"


def test_autograd_function_on_the_simulator(gpu) -> None:
    """torch.autograd over the drop-in Simulator: J and both gradients equal the C ABI's, bit for bit."""
    import torch

    import mc_dagprop_amd as api
    from mc_dagprop_amd.grad import expected_lateness

    dag = layered_small()
    ctx, gen = dag.to_api(api)
    sim = api.Simulator(ctx, gen)
    d0, e0 = sim.get_times()
    dmin = torch.tensor(d0, dtype=torch.float64, requires_grad=True)
    earliest = torch.tensor(e0, dtype=torch.float64, requires_grad=True)
    w = np.random.default_rng(1).uniform(0.5, 2.0, len(e0))
    seeds = list(range(2000))
    J = expected_lateness(sim, dmin, earliest, seeds, w)
    (2.0 * J).backward()
    obj, gd, ge = capi.Propagator(dag).gradient(seeds, w)
    assert J.item() == obj
    np.testing.assert_array_equal(bits(dmin.grad.numpy()), bits(2.0 * gd))
    np.testing.assert_array_equal(bits(earliest.grad.numpy()), bits(2.0 * ge))
    # a step of gradient descent on the earliest times lowers the sampled lateness (common random numbers)
    with torch.no_grad():
        e1 = earliest - 0.01 * earliest.grad.sign()
    J1 = expected_lateness(sim, dmin.detach(), e1, seeds, w)
    assert J1.item() < J.item()
