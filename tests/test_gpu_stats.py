"""On-device per-event statistics against the statistics oracle (oracle.event_stats).

* mean: the exactly rounded sum / n (math.fsum), bit for bit;
* quantiles: numpy.quantile's default method, bit for bit (value-equal where the selected order
  statistic is a signed zero, whose sign numpy's partition leaves unspecified);
* cause counts: exact.

Statistics are a pure function of the event-major arrays, so most cases feed synthetic arrays with
the awkward values (ties, whole rows of one value, infinities, NaN, negative numbers, signed zeros,
one sample); the end-to-end cases reduce run_many's own results.
"""

from __future__ import annotations

import numpy as np
import pytest

import dags as D
import oracle
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import layered_network, layered_small

pytestmark = pytest.mark.gpu

Q = [0.0, 0.01, 0.25, 0.5, 0.9, 0.99, 0.999, 1.0]


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint64)


def assert_quantiles_equal(got: np.ndarray, want: np.ndarray) -> None:
    same = bits(got) == bits(want)
    zeros = (got == 0) & (want == 0)
    nans = np.isnan(got) & np.isnan(want)
    bad = ~(same | zeros | nans)
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], want[bad][:5])


def synthetic(rng: np.random.Generator, n: int, E: int, preds) -> tuple[np.ndarray, np.ndarray]:
    """Sample-major realized / cause with per-event character: ties, constant rows, capped rows,
    negative and signed-zero values, infinities, a NaN row."""
    r = np.empty((n, E))
    kinds = rng.integers(0, 8, E)
    for e in range(E):
        k = kinds[e]
        if k == 0:
            r[:, e] = rng.normal(100.0, 30.0, n)
        elif k == 1:
            r[:, e] = 42.0  # one value
        elif k == 2:
            r[:, e] = np.minimum(rng.exponential(50.0, n), 60.0)  # mass at a cap
        elif k == 3:
            r[:, e] = rng.integers(-3, 4, n).astype(float)  # heavy ties, negatives
        elif k == 4:
            r[:, e] = rng.choice([-0.0, 0.0, 1.0], n)  # signed zeros
        elif k == 5:
            r[:, e] = rng.uniform(-1e300, 1e300, n)
        elif k == 6:
            r[:, e] = rng.normal(0, 1, n)
            r[rng.integers(0, n), e] = np.inf
        else:
            r[:, e] = rng.normal(0, 1, n)
            if e % 2:
                r[rng.integers(0, n), e] = np.nan
    c = np.empty((n, E), dtype=np.int32)
    for e in range(E):
        choices = np.array([-1] + list(preds[e]), dtype=np.int32)
        c[:, e] = rng.choice(choices, n)
    return r, c


def run_stats_device(P: capi.Propagator, r: np.ndarray, c: np.ndarray, q, pad: int = 0):
    n, E = r.shape
    ld = n + pad
    R = np.full((E, ld), -5.0)
    R[:, :n] = r.T
    Cc = np.full((E, ld), -7, dtype=np.int32)
    Cc[:, :n] = c.T
    return P.event_stats_device(capi.DeviceBuffer.from_array(R), capi.DeviceBuffer.from_array(Cc), n, ld, q)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 257, 5000])
def test_stats_synthetic_match_oracle(gpu, n: int) -> None:
    rng = np.random.default_rng(n)
    dag = D.random_dag(rng, 40, max_preds=4, duplicates=True)
    P = capi.Propagator(dag)
    preds = oracle.pred_sources(dag)
    r, c = synthetic(rng, n, P.E, preds)
    mean, quant, counts = run_stats_device(P, r, c, Q, pad=3)
    om, oq, oc = oracle.event_stats(r, c, preds, Q)
    np.testing.assert_array_equal(bits(mean[~np.isnan(om)]), bits(om[~np.isnan(om)]))
    assert np.array_equal(np.isnan(mean), np.isnan(om))
    assert_quantiles_equal(quant, oq)
    np.testing.assert_array_equal(counts, oc)


def crafted_rows(rng: np.random.Generator, n: int) -> np.ndarray:
    """Rows whose quantiles lie outside the masses of their extremes, for each route of the selection:
    value bins of ordinary spread, a middle value held by more samples than one bin gathers (radix
    passes), a range of three subnormals (bin scale overflows: radix passes), ranges of infinite width
    (radix passes), signed zeros among positives, heavy ties, and sums near overflow."""
    third = n // 3
    cols = [
        rng.normal(100.0, 30.0, n),
        np.concatenate([rng.normal(-5.0, 1.0, third), np.full(n - 2 * third, 0.5), rng.normal(5.0, 1.0, third)]),
        rng.choice([5e-324, 1e-323, 1.5e-323], n),
        np.where(rng.random(n) < 0.5, rng.normal(0.0, 1.0, n), rng.choice([-1.7e308, 1.7e308], n)),
        np.concatenate([rng.normal(0.0, 1.0, n - 2), [np.inf, -np.inf]]),
        rng.choice([-0.0, 0.0, 1e-300, 2.0, 3.0], n),
        np.round(rng.normal(0.0, 4.0, n)),
        rng.uniform(-1e300, 1e300, n),
        np.exp(rng.normal(0.0, 30.0, n)),
    ]
    r = np.stack([rng.permutation(c) for c in cols], axis=1)
    return r


@pytest.mark.parametrize("n", [5000, 100_003])
def test_quantiles_value_bins_and_radix_fallback(gpu, n: int) -> None:
    rng = np.random.default_rng(n)
    r = crafted_rows(rng, n)
    dag = D.parity_long_chain(r.shape[1] - 1)
    P = capi.Propagator(dag)
    assert P.E == r.shape[1]
    q = [0.001, 0.25, 0.5, 0.75, 0.9, 0.99, 0.999]
    c = np.full(r.shape, -1, dtype=np.int32)
    mean, quant, _ = run_stats_device(P, r, c, q, pad=1)
    om, oq, _ = oracle.event_stats(r, c, oracle.pred_sources(dag), q)
    assert_quantiles_equal(quant, oq)
    finite = np.isfinite(om)
    np.testing.assert_array_equal(bits(mean[finite]), bits(om[finite]))


@pytest.mark.parametrize("n,pad", [(3001, 5), (4096, 0), (77, 1)])
def test_cause_counts_wide_events_and_strays(gpu, n: int, pad: int) -> None:
    """Events with 1 to 14 predecessors (counted in registers up to 8 distinct sources, in LDS
    beyond), causes that are no predecessor of their event (counted nowhere), rows at every
    alignment (odd ld)."""
    rng = np.random.default_rng(n + pad)
    dag = D.random_dag(rng, 60, max_preds=14, duplicates=True)
    P = capi.Propagator(dag)
    preds = oracle.pred_sources(dag)
    assert max(len(set(p)) for p in preds) > 8 and any(0 < len(set(p)) <= 8 for p in preds)
    r, _ = synthetic(rng, n, P.E, preds)
    c = np.empty((n, P.E), dtype=np.int32)
    for e in range(P.E):
        choices = np.array([-1] + list(preds[e]) + [P.E + 3, e, 2**31 - 1, -2], dtype=np.int32)
        c[:, e] = rng.choice(choices, n)
    _, _, counts = run_stats_device(P, r, c, [], pad=pad)
    _, _, oc = oracle.event_stats(r, c, preds, [])
    np.testing.assert_array_equal(counts, oc)


def test_cause_counts_use_first_position_of_repeated_source(gpu) -> None:
    """Event 2 lists source 0 twice (through two activities): its causes count at position 0."""
    from mc_dagprop_amd.workloads import CONSTANT, DistSpec, FlatDag

    events = [(0.0, 100.0, 0.0), (1.0, 100.0, 0.0), (2.0, 100.0, 0.0), (3.0, 100.0, 0.0)]
    acts = {(0, 1): (0, 1.0, 0), (0, 2): (1, 2.0, 0), (1, 2): (2, 1.0, 0), (2, 3): (3, 1.0, 0)}
    prec = [(1, [(0, 0)]), (2, [(0, 1), (1, 2), (0, 0)]), (3, [(2, 3)])]
    dag = FlatDag.from_lists(events, acts, prec, 50.0, [DistSpec(0, CONSTANT, (0.5, 0.0, float("inf")))])
    P = capi.Propagator(dag)
    preds = oracle.pred_sources(dag)
    assert preds[2] == [0, 1, 0]
    rng = np.random.default_rng(7)
    r, c = synthetic(rng, 999, P.E, preds)
    _, _, counts = run_stats_device(P, r, c, [])
    _, _, oc = oracle.event_stats(r, c, preds, [])
    np.testing.assert_array_equal(counts, oc)
    off = P.pred_offsets()
    assert counts.size == off[-1] + P.E
    e2 = counts[off[2] + 2: off[3] + 3]
    assert e2[3] == 0 and e2[1] == np.sum(c[:, 2] == 0)  # [-1, pos 0, pos 1, pos 2 (repeat of pos 0)]


def test_run_stats_matches_run_many(gpu) -> None:
    """Statistics straight from the device equal the oracle's on run_many's own results."""
    dag = layered_small()
    P = capi.Propagator(dag)
    seeds = np.arange(-3000, 7000)
    mean, quant, counts = P.run_stats(seeds, Q)
    r, _, c = P.run_many(seeds)
    om, oq, oc = oracle.event_stats(r, c, oracle.pred_sources(dag), Q)
    np.testing.assert_array_equal(bits(mean), bits(om))
    assert_quantiles_equal(quant, oq)
    np.testing.assert_array_equal(counts, oc)
    assert counts.sum() == seeds.size * P.E  # every sample's cause is -1 or a predecessor


def test_run_stats_network_slice(gpu) -> None:
    """A 30-layer slice of the network workload (8 activity types, most events saturated at the cap)."""
    dag = layered_network(layers=30)
    P = capi.Propagator(dag)
    seeds = np.arange(20000)
    mean, quant, counts = P.run_stats(seeds, [0.5, 0.9, 0.99])
    r, _, c = P.run_many(seeds)
    om, oq, oc = oracle.event_stats(r, c, oracle.pred_sources(dag), [0.5, 0.9, 0.99])
    np.testing.assert_array_equal(bits(mean), bits(om))
    assert_quantiles_equal(quant, oq)
    np.testing.assert_array_equal(counts, oc)


def test_python_api_run_stats(gpu) -> None:
    from mc_dagprop_amd import Activity, DagContext, Event, EventTimestamp, GenericDelayGenerator, Simulator

    events = [Event(str(i), EventTimestamp(t, 100.0, 0.0)) for i, t in enumerate([0.0, 5.0, 10.0])]
    acts = {(0, 1): Activity(idx=0, minimal_duration=3.0, activity_type=1),
            (1, 2): Activity(idx=1, minimal_duration=5.0, activity_type=1)}
    ctx = DagContext(events=events, activities=acts, precedence_list=[(1, [(0, 0)]), (2, [(1, 1)])], max_delay=60.0)
    gen = GenericDelayGenerator()
    gen.add_exponential(1, 2.0, max_scale=10.0)
    sim = Simulator(ctx, gen)
    st = sim.run_stats(range(4096), quantiles=[0.5, 0.99])
    assert st["mean"].shape == (3,) and st["quantiles"].shape == (3, 2)
    res = sim.run_many(range(4096))
    R = np.array([x.realized for x in res])
    np.testing.assert_array_equal(bits(st["quantiles"]), bits(np.quantile(R, [0.5, 0.99], axis=0).T))
    co = st["cause_offsets"]
    cc = st["cause_counts"]
    assert list(co) == [0, 1, 3, 5]
    assert cc[co[0]:co[1]].tolist() == [4096]  # event 0: no predecessors


def test_stats_rejects_bad_quantiles(gpu) -> None:
    P = capi.Propagator(layered_small())
    with pytest.raises(capi.McdpError, match="range"):
        P.run_stats(np.arange(10), [1.5])
    with pytest.raises(capi.McdpError, match="8 quantiles"):
        P.run_stats(np.arange(10), np.linspace(0, 1, 9))


def test_statistics_before_the_first_run_keep_the_plan_consistent(gpu) -> None:
    """Statistics called before any run upload the plan; the later runs must still use the schedule they
    launch with (bit-exact against a propagator that ran first)."""
    dag = layered_small()
    a, b = capi.Propagator(dag), capi.Propagator(dag)
    E = a.E
    n = 300
    realized = capi.DeviceBuffer.from_array(np.zeros((E, n)))
    cause = capi.DeviceBuffer.from_array(np.full((E, n), -1, dtype=np.int32))
    a.event_stats_device(realized, cause, n, n, q=(0.5,))
    seeds = np.arange(n, dtype=np.int32)
    for x, y in zip(a.run_many(seeds), b.run_many(seeds)):
        np.testing.assert_array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
