"""Batched run_many: several DAGs in one persistent launch give, for every DAG and sample, the bits
each DAG's own run_many gives (the sampler is keyed by seed and activity, the recurrence is the
reference's), whatever the batch's shape, sample counts, leading dimensions or missing outputs."""

from __future__ import annotations

import numpy as np
import pytest

import dags as D
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import layered_network, layered_small

pytestmark = pytest.mark.gpu


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint64)


def _dags():
    rng = np.random.default_rng(11)
    return [
        layered_small(),
        layered_network(layers=12),
        D.random_dag(rng, 120, max_preds=5, duplicates=True, oob_edges=True, dists=D.STD_DISTS),
        D.random_dag(rng, 7, max_preds=2, dists=D.STD_DISTS),
    ]


@pytest.mark.parametrize("shape", [(0, 0, 0), (8, 16, 1), (5, 4, 2)])
def test_batch_matches_single_runs(gpu, shape) -> None:
    dags = _dags()
    ns = [1000, 300, 129, 1]
    seeds = [np.arange(n, dtype=np.int32) * (k + 3) - 500 for k, n in enumerate(ns)]
    singles = [capi.Propagator(d).run_many(s) for d, s in zip(dags, seeds)]
    B = capi.Batch([capi.Propagator(d) for d in dags], *shape)
    got = B.run_many(seeds)
    for (r, d, c), (r0, d0, c0) in zip(got, singles):
        np.testing.assert_array_equal(bits(r), bits(r0))
        np.testing.assert_array_equal(bits(d), bits(d0))
        np.testing.assert_array_equal(c, c0)


def test_batch_general_path_and_empty_plan(gpu) -> None:
    """A plan that spills predecessors to global memory puts the whole batch on the general
    variant; a plan with no samples contributes no work."""
    from test_gpu_parity import _wide_dag

    dags = [_wide_dag(600, 5), layered_small(), layered_small()]
    props = [capi.Propagator(d) for d in dags]
    B = capi.Batch(props)
    assert props[0].plan_info()["global_preds"] > 0
    seeds = [np.arange(70), np.arange(200, 330), np.arange(0)]
    got = B.run_many(seeds)
    for d, s, (r, dd, c) in zip(dags, seeds, got):
        if s.size == 0:
            assert r.shape[0] == 0
            continue
        r0, d0, c0 = capi.Propagator(d).run_many(s)
        np.testing.assert_array_equal(bits(r), bits(r0))
        np.testing.assert_array_equal(bits(dd), bits(d0))
        np.testing.assert_array_equal(c, c0)


def test_batch_device_leading_dimension_and_missing_outputs(gpu) -> None:
    dags = [layered_small(), layered_network(layers=5)]
    props = [capi.Propagator(d) for d in dags]
    B = capi.Batch(props)
    ns, lds = [777, 200], [800, 203]
    sd = [capi.DeviceBuffer.from_array(np.arange(n, dtype=np.int32)) for n in ns]
    real = [capi.DeviceBuffer.from_array(np.full((P.E, ld), -7.0)) for P, ld in zip(props, lds)]
    cause = [capi.DeviceBuffer.from_array(np.full((P.E, ld), -9, dtype=np.int32)) for P, ld in zip(props, lds)]
    B.run_many_device(sd, ns, lds, real, None, [cause[0], None])
    for k, (P, n, ld) in enumerate(zip(props, ns, lds)):
        R = real[k].to_array(np.float64, (P.E, ld))
        Cc = cause[k].to_array(np.int32, (P.E, ld))
        r0, _, c0 = capi.Propagator(dags[k]).run_many(np.arange(n))
        assert np.all(R[:, n:] == -7.0)
        np.testing.assert_array_equal(bits(R[:, :n].T), bits(r0))
        if k == 0:
            np.testing.assert_array_equal(Cc[:, :n].T, c0)
        else:
            assert np.all(Cc == -9)  # cause output skipped


def test_reconfigured_member_is_rejected(gpu) -> None:
    props = [capi.Propagator(layered_small())]
    B = capi.Batch(props)
    props[0].configure(4, 8, 1)
    with pytest.raises(capi.McdpError, match="reconfigured"):
        B.run_many([np.arange(10)])


@pytest.mark.parametrize("budget_mb", [5, 4096])
def test_batch_stats_match_single_runs(gpu, budget_mb: int) -> None:
    """mcdp_batch_stats_host over DAGs of uneven size, in groups that fit the memory budget (5 MB: the
    2.4 MB DAG alone, then the 4.3 MB one with the two small ones; 4 GB: one group), equals each DAG's own
    run_many followed by the device statistics: means, quantiles and cause counts bit for bit."""
    dags = _dags()
    ns = [1000, 300, 129, 1]
    seeds = [np.arange(n, dtype=np.int32) * (k + 5) + 77 for k, n in enumerate(ns)]
    q = (0.0, 0.5, 0.9, 0.99, 1.0)
    want = []
    for d, s in zip(dags, seeds):
        P = capi.Propagator(d)
        n = s.size
        sd = capi.DeviceBuffer.from_array(s)
        real = capi.DeviceBuffer(8 * max(P.E * n, 1))
        cause = capi.DeviceBuffer(4 * max(P.E * n, 1))
        P.run_many_device(sd, n, n, real, None, cause)
        want.append(P.event_stats_device(real, cause, n, n, q))
    B = capi.Batch([capi.Propagator(d) for d in dags])
    got = B.stats(seeds, q, memory_budget=budget_mb << 20)
    for (m, qq, c), (m0, qq0, c0) in zip(got, want):
        np.testing.assert_array_equal(bits(m), bits(m0))
        np.testing.assert_array_equal(bits(qq), bits(qq0))
        np.testing.assert_array_equal(c, c0)
    means_only = B.stats(seeds, (), memory_budget=budget_mb << 20, cause_counts=False)
    for (m, qq, c), (m0, _, _) in zip(means_only, want):
        np.testing.assert_array_equal(bits(m), bits(m0))
        assert qq.shape[1] == 0 and c is None



def test_batch_launch_replays_from_a_graph(gpu) -> None:
    """The prepared batch launch (queue-head reset kernel + persistent kernel, no memset, copy or sync)
    captured into a HIP graph: three replays, each after the outputs were overwritten and after other eager
    work on the device, reproduce the eager launch bit for bit."""
    dags = _dags()
    ns = [1000, 300, 129, 1]
    seeds = [np.arange(n, dtype=np.int32) * (k + 2) + 9 for k, n in enumerate(ns)]
    props = [capi.Propagator(d) for d in dags]
    B = capi.Batch(props)
    sd = [capi.DeviceBuffer.from_array(s) for s in seeds]
    real = [capi.DeviceBuffer(8 * P.E * n) for P, n in zip(props, ns)]
    dur = [capi.DeviceBuffer(8 * P.L * n) for P, n in zip(props, ns)]
    cause = [capi.DeviceBuffer(4 * P.E * n) for P, n in zip(props, ns)]
    outs = real + dur + cause

    def read():
        return ([r.to_array(np.float64, (P.E, n)) for r, P, n in zip(real, props, ns)],
                [d.to_array(np.float64, (P.L, n)) for d, P, n in zip(dur, props, ns)],
                [c.to_array(np.int32, (P.E, n)) for c, P, n in zip(cause, props, ns)])

    B.prepare_device(sd, ns, ns, real, dur, cause)
    stream = capi.C.c_void_p(0)
    capi.check(capi.lib().mcdp_stream_create(capi.C.byref(stream)))
    B.launch(stream)
    capi.synchronize(stream)
    want = read()
    singles = [capi.Propagator(d).run_many(s) for d, s in zip(dags, seeds)]
    for r, (r0, _, _) in zip(want[0], singles):
        np.testing.assert_array_equal(bits(r.T), bits(r0))
    graph = capi.C.c_void_p(0)
    capi.check(capi.lib().mcdp_graph_capture_begin(stream))
    B.launch(stream)
    capi.check(capi.lib().mcdp_graph_capture_end(stream, capi.C.byref(graph)))
    other = capi.Propagator(layered_small())
    o_seeds = capi.DeviceBuffer.from_array(np.arange(500, dtype=np.int32))
    o_real, o_dur, o_cause = (capi.DeviceBuffer(8 * other.E * 500), capi.DeviceBuffer(8 * other.L * 500),
                              capi.DeviceBuffer(4 * other.E * 500))
    try:
        for rep in range(3):
            for b in outs:
                capi.check(capi.lib().mcdp_memset(b.ptr, 0xFF, b.nbytes))
            other.run_many_device(o_seeds, 500, 500, o_real, o_dur, o_cause)
            capi.synchronize()
            capi.check(capi.lib().mcdp_graph_launch(graph, stream))
            capi.synchronize(stream)
            got = read()
            for kind, g_list, w_list in zip(("realized", "durations", "cause"), got, want):
                for g, w in zip(g_list, w_list):
                    np.testing.assert_array_equal(bits(g) if g.dtype != np.int32 else g,
                                                  bits(w) if w.dtype != np.int32 else w,
                                                  err_msg=f"replay {rep} {kind}")
    finally:
        capi.check(capi.lib().mcdp_graph_destroy(graph))
        capi.check(capi.lib().mcdp_stream_destroy(stream))


def test_batch_matches_oracle(gpu) -> None:
    """Every DAG's batched durations, fed to the CPU oracle (the reference's propagation loop,
    _core.cpp:332-350), reproduce the batch's realized and cause bit for bit; links that need no
    draw carry the oracle's closed-form durations."""
    import oracle

    dags = _dags()
    ns = [700, 300, 129, 1]
    seeds = [np.arange(n, dtype=np.int32) * (k + 7) - 300 for k, n in enumerate(ns)]
    B = capi.Batch([capi.Propagator(d) for d in dags])
    for d, (r, dd, c) in zip(dags, B.run_many(seeds)):
        ctx = oracle.OracleCtx(d)
        r0, c0 = ctx.propagate(dd)
        np.testing.assert_array_equal(bits(r), bits(r0))
        np.testing.assert_array_equal(c, c0)
        base = ctx.base_durations()
        fixed = ~np.isnan(base)
        np.testing.assert_array_equal(bits(dd[:, fixed]), bits(np.broadcast_to(base[fixed], dd[:, fixed].shape)))


def test_batch_stats_match_oracle(gpu) -> None:
    """Batch.stats (grouped launches, device statistics) equals the oracle's per-event statistics
    (math.fsum mean, numpy.quantile, explicit cause counting; demo/monte_carlo.py:45-63 is the
    consumer these replace) over the oracle's propagation of the same seeds' durations."""
    import oracle

    dags = _dags()
    ns = [640, 257, 129, 3]
    seeds = [np.arange(n, dtype=np.int32) * 3 + 11 * k for k, n in enumerate(ns)]
    q = (0.0, 0.5, 0.9, 0.99, 1.0)
    props = [capi.Propagator(d) for d in dags]
    got = capi.Batch(props).stats(seeds, q, memory_budget=5 << 20)  # two groups
    for d, P, s, (m, qq, cc) in zip(dags, props, seeds, got):
        _, dd, _ = capi.Propagator(d).run_many(s)
        r0, c0 = oracle.OracleCtx(d).propagate(dd)
        m0, q0, k0 = oracle.event_stats(r0, c0, oracle.pred_sources(d), q)
        np.testing.assert_array_equal(bits(m), bits(m0))
        np.testing.assert_array_equal(bits(qq), bits(q0))
        np.testing.assert_array_equal(cc, k0)


def test_prepared_launch_refused_after_new_times(gpu) -> None:
    """set_times frees a member's device plan, which a prepared launch (and a graph captured from it)
    points into: the next launch must fail instead of reading freed memory; a new prepare works and
    gives the member's own results with the new times."""
    dags = [layered_small(), layered_network(layers=6)]
    props = [capi.Propagator(d) for d in dags]
    B = capi.Batch(props)
    ns = [300, 200]
    sd = [capi.DeviceBuffer.from_array(np.arange(n, dtype=np.int32)) for n in ns]
    real = [capi.DeviceBuffer(8 * P.E * n) for P, n in zip(props, ns)]
    B.prepare_device(sd, ns, ns, real)
    B.launch()
    capi.synchronize()
    dmin, earliest = props[1].get_times()
    props[1].set_times(dmin * 1.25, earliest)
    with pytest.raises(capi.McdpError, match="prepare the batch again"):
        B.launch()
    with pytest.raises(capi.McdpError, match="no prepared launch"):
        B.launch()  # refused once: the batch stays unprepared
    B.prepare_device(sd, ns, ns, real)
    B.launch()
    capi.synchronize()
    solo = capi.Propagator(dags[1])
    solo.set_times(dmin * 1.25, earliest)
    r0, _, _ = solo.run_many(np.arange(ns[1]))
    np.testing.assert_array_equal(bits(real[1].to_array(np.float64, (props[1].E, ns[1])).T), bits(r0))
