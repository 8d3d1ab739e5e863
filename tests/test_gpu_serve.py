"""The serving loop: requests over resident contexts, admitted by device-memory footprint.

Each request's statistics must equal those of the same seeds run alone (``run_stats``), bit for bit,
whatever else is queued or running; the pool's peak use stays within its budget; requests wait for
memory instead of failing; requests that could never fit are refused at submission.  Requests above
the request cap run in chunks of samples and still equal ``run_stats`` bit for bit, quantiles
included: rows kept whole for the events whose order statistics may leave their extremes' masses
(chosen from the first chunk and the events' bounds, which must enclose every draw), and further
passes over the chunks that select the other order statistics exactly; with the default limits,
with no kept rows and limits small enough that regions are refined over several passes and targets
wait for refine slots, with too few kept rows, and on rows full of ties.  With exact selection off,
quantiles are within one histogram bin.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest

import dags as D
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import layered_network, layered_small

pytestmark = pytest.mark.gpu


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def same_stats(got, want) -> None:
    mean, quant, counts = got[:3]
    w_mean, w_quant, w_counts = want
    np.testing.assert_array_equal(bits(mean), bits(w_mean))
    ok = (bits(quant) == bits(w_quant)) | ((quant == 0) & (w_quant == 0))
    assert ok.all()
    np.testing.assert_array_equal(counts, w_counts)


BINS = 4096  # histogram bins of statistics over sample chunks


def footprint(E: int, P: int, n: int, nq: int, cb: int = 1) -> int:
    """A request's device bytes run whole (cb: bytes per cause, 1 for winner positions, which every context
    here allows: no event has more than 254 entries)."""
    up = lambda x: (x + 255) // 256 * 256  # noqa: E731
    return up(4 * n) + up(8 * E * n) + up(cb * E * n) + up(8 * E) + up(8 * E * nq) + up(8 * (E + P))


def contexts():
    dags = [layered_small(), layered_network(layers=10), D.continuous_dag(3, 60, max_delay=20.0, shared_links=True)]
    return dags, [capi.Propagator(d) for d in dags]


def test_results_in_the_mix_equal_alone(gpu) -> None:
    dags, props = contexts()
    budget = 48 << 20
    srv = capi.Server(budget, max_running=4, n_streams=3, request_cap=budget)
    ids = [srv.add_context(p) for p in props]
    rng = np.random.default_rng(4)
    reqs = []
    for _ in range(30):
        c = int(rng.integers(len(props)))
        n = int(rng.integers(1, 3000))
        seeds = rng.integers(0, 2**31 - 1, n).astype(np.int32)
        q = sorted(rng.uniform(0, 1, int(rng.integers(0, 4))).tolist())
        reqs.append((srv.submit(ids[c], seeds, q), c, seeds, q))
    for rid, c, seeds, q in reqs:
        got = srv.wait(rid)
        assert got[3] > 0.0
        same_stats(got, props[c].run_stats(seeds, q))
    st = srv.stats()
    assert st["submitted"] == 30 and st["completed"] == 30 and st["failed"] == 0
    assert st["queued"] == 0 and st["running"] == 0 and st["bytes_in_use"] == 0
    assert 0 < st["peak_bytes"] <= budget
    assert st["max_running"] >= 2  # requests overlapped on the device
    srv.close()


def test_requests_wait_for_memory(gpu) -> None:
    """A budget that holds one request at a time: all complete, one at a time, in order."""
    dags, props = contexts()
    p = props[1]
    n = 2000
    need = footprint(p.E, int(p.pred_offsets()[-1]), n, 3)
    srv = capi.Server(need + need // 2, max_running=8, n_streams=4, request_cap=need + need // 2)
    cid = srv.add_context(p)
    seeds = [np.arange(i * n, (i + 1) * n, dtype=np.int32) for i in range(6)]
    rids = [srv.submit(cid, s) for s in seeds]
    for rid, s in zip(rids, seeds):
        same_stats(srv.wait(rid), p.run_stats(s))
    st = srv.stats()
    assert st["completed"] == 6 and st["max_running"] == 1 and st["peak_bytes"] <= need + need // 2
    srv.close()


def test_oversized_request_is_refused(gpu) -> None:
    _, props = contexts()
    p = props[1]
    srv = capi.Server(1 << 20, max_running=2, n_streams=1, request_cap=1 << 20)
    cid = srv.add_context(p)
    with pytest.raises(capi.McdpError, match="cap"):
        srv.submit(cid, np.arange(100_000, dtype=np.int32))
    with pytest.raises(capi.McdpError):
        srv.submit(cid + 5, np.arange(10, dtype=np.int32))
    with pytest.raises(capi.McdpError):
        srv.submit(cid, np.arange(10, dtype=np.int32), q=[1.5])
    small = np.arange(10, dtype=np.int32)
    same_stats(srv.wait(srv.submit(cid, small)), p.run_stats(small))
    srv.close()


def test_threads_submit_and_wait(gpu) -> None:
    dags, props = contexts()
    srv = capi.Server(64 << 20, max_running=6, n_streams=4, request_cap=64 << 20)
    ids = [srv.add_context(p) for p in props]
    errors = []

    def client(k: int) -> None:
        try:
            rng = np.random.default_rng(100 + k)
            for _ in range(5):
                c = int(rng.integers(len(props)))
                seeds = rng.integers(0, 10**6, int(rng.integers(1, 1500))).astype(np.int32)
                same_stats(srv.wait(srv.submit(ids[c], seeds)), props[c].run_stats(seeds))
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert srv.stats()["completed"] == 20
    srv.close()


def test_close_with_requests_in_flight(gpu) -> None:
    """Closing drains running requests and drops queued ones without hanging."""
    _, props = contexts()
    srv = capi.Server(16 << 20, max_running=2, n_streams=2, request_cap=16 << 20)
    cid = srv.add_context(props[1])
    for i in range(10):
        srv.submit(cid, np.arange(i * 1000, (i + 1) * 1000, dtype=np.int32))
    srv.close()


def test_chunked_requests(gpu) -> None:
    """Requests above the request cap run in chunks of samples: mean, cause counts and every quantile equal
    the one-piece statistics bit for bit; chunked and one-piece requests share the device."""
    dags, props = contexts()
    big, small = props[1], props[0]
    srv = capi.Server(96 << 20, max_running=4, n_streams=3, request_cap=24 << 20)
    srv.debug_kept_rows(0)  # every order statistic off the extremes' masses takes further passes
    cb, cs = srv.add_context(big), srv.add_context(small)
    rng = np.random.default_rng(12)
    q = [0.0, 0.25, 0.5, 0.9, 0.99, 1.0]
    reqs = []
    for i in range(6):
        seeds = rng.integers(0, 2**31 - 1, int(rng.integers(3000, 6000))).astype(np.int32)
        reqs.append((srv.submit(cb, seeds, q), big, seeds, dags[1].max_delay))
        s2 = rng.integers(0, 2**31 - 1, 500).astype(np.int32)
        reqs.append((srv.submit(cs, s2, q), small, s2, None))
    assert srv.stats()["chunked"] == 6
    for rid, p, seeds, md in reqs:
        same_stats(srv.wait(rid), p.run_stats(seeds, q))
    st = srv.stats()
    assert st["completed"] == 12 and st["peak_bytes"] <= 96 << 20 and st["max_running"] >= 2
    assert st["exact_rounds"] >= 6
    srv.close()


def jitter_chain(n_events: int = 60, max_delay: float = 500.0):
    """A chain whose links alternately add 0, 2 or 5 (a three-value empirical law) and an exponential of mean
    1e-12: realized times cluster within a few ulps of few values, so histogram bins hold many equal or
    nearly equal values (refined first by value, then by key offset)."""
    from mc_dagprop_amd.workloads import EXPONENTIAL, DistSpec

    dists = list(D.GRAD_DISTS) + [DistSpec(5, EXPONENTIAL, (1e-12, float("inf"), float("inf")))]
    events = [(0.0, 1e6, 0.0)] + [(float(3 * i), 1e6, 0.0) for i in range(1, n_events)]
    acts = {(i, i + 1): (i, 1.0, 4 if i % 2 == 0 else 5) for i in range(n_events - 1)}
    prec = [(i + 1, [(i, i)]) for i in range(n_events - 1)]
    return D.FlatDag.from_lists(events, acts, prec, max_delay, dists)


@pytest.mark.parametrize("kept", [0, -1], ids=["passes", "kept_rows"])
def test_chunked_exact_selection(gpu, kept) -> None:
    """Bit-exact chunked quantiles with the default limits: a layered DAG and a chain of clustered, tied values,
    seven quantiles each; by further passes alone, and with the default kept rows."""
    dags = [layered_network(layers=2), jitter_chain()]
    props = [capi.Propagator(d) for d in dags]
    srv = capi.Server(48 << 20, max_running=2, n_streams=2, request_cap=8 << 20)
    srv.debug_kept_rows(kept)
    ids = [srv.add_context(p) for p in props]
    q = (0.0, 0.1, 0.5, 0.75, 0.9, 0.99, 1.0)
    reqs = []
    for c, p in enumerate(props):
        seeds = (np.arange(20_000, dtype=np.int64) * 7919 + 11 * c).astype(np.int32)
        reqs.append((srv.submit(ids[c], seeds, q), p, seeds))
    assert srv.stats()["chunked"] == 2
    for rid, p, seeds in reqs:
        same_stats(srv.wait(rid), p.run_stats(seeds, q))
    st = srv.stats()
    assert st["completed"] == 2 and st["failed"] == 0
    if kept == 0:
        assert st["exact_rounds"] >= 2 and st["kept_rows"] == 0
    else:
        assert st["kept_rows"] > 0
    srv.close()


@pytest.mark.parametrize("limits", [(8, 1, 16), (2, 2, 16), (1, 1, 16)], ids=["refine", "tiny", "single"])
def test_chunked_exact_selection_small_limits(gpu, limits) -> None:
    """Regions larger than one pass may gather: at most 8, 2 or 1 values gathered per region, 1-2 candidates
    per target and 16 refine slots force refining passes (by value, then by key offset) and targets that wait
    for candidate space or a slot; still bit for bit."""
    dags = [D.continuous_dag(3, 40, max_delay=20.0, shared_links=True), jitter_chain(12)]
    props = [capi.Propagator(d) for d in dags]
    srv = capi.Server(16 << 20, max_running=2, n_streams=2, request_cap=2 << 20)
    srv.debug_exact_limits(*limits)
    srv.debug_kept_rows(0)
    ids = [srv.add_context(p) for p in props]
    q = (0.1, 0.5, 0.9)
    reqs = []
    for c, p in enumerate(props):
        seeds = (np.arange(20_000, dtype=np.int64) * 104729 + 5 * c).astype(np.int32)
        reqs.append((srv.submit(ids[c], seeds, q), p, seeds))
    assert srv.stats()["chunked"] == 2
    for rid, p, seeds in reqs:
        same_stats(srv.wait(rid), p.run_stats(seeds, q))
    st = srv.stats()
    assert st["completed"] == 2 and st["failed"] == 0 and st["exact_rounds"] >= 4
    srv.close()


def test_ranges_bound_every_draw(gpu) -> None:
    """Every duration the device draws lies in its link's range, and every realized time in its event's range
    (mcdp_propagator_ranges), over every sampling method the compiler picks: serving trusts a chunk's extreme
    that equals its event's bound never to be passed by a later chunk."""
    from test_ranges import mixed_dag

    for dag in (mixed_dag(0), mixed_dag(1, 500, 1e3), layered_network(layers=12)):
        p = capi.Propagator(dag)
        llo, lhi, elo, ehi = p.ranges()
        seeds = np.arange(8192, dtype=np.int32) * 31 + 7
        real, dur, _ = p.run_many(seeds)
        assert np.all((dur >= llo[None, :]) & (dur <= lhi[None, :]))
        assert np.all((real >= elo[None, :]) & (real <= ehi[None, :]))
        # the bounds are attained where they are point masses (constants, caps, earliest times)
        assert np.any(real == elo[None, :]) and np.any(real == ehi[None, :])


def kept_case():
    """The network workload's literal distributions saturate most events at their cap: a few events per
    hundred have order statistics off their extremes' masses, the rest on them."""
    d = layered_network(layers=20)
    p = capi.Propagator(d)
    seeds = (np.arange(24_000, dtype=np.int64) * 7919 + 3).astype(np.int32)
    return d, p, seeds


def test_kept_rows_need_no_further_pass(gpu) -> None:
    """With the default kept rows a chunked request on the network workload needs no further pass, and its
    statistics equal run_stats bit for bit."""
    d, p, seeds = kept_case()
    q = (0.5, 0.9, 0.99)
    srv = capi.Server(256 << 20, max_running=2, n_streams=2, request_cap=64 << 20)
    cid = srv.add_context(p)
    got = srv.wait(srv.submit(cid, seeds, q))
    st = srv.stats()
    srv.close()
    same_stats(got, p.run_stats(seeds, q))
    assert st["chunked"] == 1 and st["failed"] == 0
    assert 0 < st["kept_rows"] <= max(16, p.E // 128)
    assert st["exact_rounds"] == 0


@pytest.mark.parametrize("rows", [1, 2])
def test_too_few_kept_rows(gpu, rows) -> None:
    """Kept rows past the limit fall back to further passes: still bit for bit."""
    d, p, seeds = kept_case()
    q = (0.5, 0.9, 0.99)
    srv = capi.Server(256 << 20, max_running=2, n_streams=2, request_cap=64 << 20)
    srv.debug_kept_rows(rows)
    cid = srv.add_context(p)
    got = srv.wait(srv.submit(cid, seeds, q))
    st = srv.stats()
    srv.close()
    same_stats(got, p.run_stats(seeds, q))
    assert st["chunked"] == 1 and st["failed"] == 0 and st["kept_rows"] == rows and st["exact_rounds"] >= 1


def test_kept_rows_on_continuous_rows(gpu) -> None:
    """Continuous rows put every order statistic off the extremes: the kept rows fill up (16 of them) and the
    other events take further passes; bit for bit."""
    d = D.continuous_dag(5, 120, max_delay=30.0, shared_links=True)
    p = capi.Propagator(d)
    seeds = (np.arange(20_000, dtype=np.int64) * 104729 + 9).astype(np.int32)
    q = (0.1, 0.5, 0.9)
    srv = capi.Server(64 << 20, max_running=2, n_streams=2, request_cap=16 << 20)
    cid = srv.add_context(p)
    got = srv.wait(srv.submit(cid, seeds, q))
    st = srv.stats()
    srv.close()
    same_stats(got, p.run_stats(seeds, q))
    assert st["chunked"] == 1 and st["kept_rows"] == 16 and st["exact_rounds"] >= 1


def wide_fan_in(n_src: int = 300, max_delay: float = 40.0):
    """One event with n_src entries (more than a byte can number) and a successor."""
    events = [(float(i % 7), 1e6, 0.0) for i in range(n_src + 2)]
    acts = {(i, n_src): (i, 1.0 + (i % 5), i % 4) for i in range(n_src)}
    acts[(n_src, n_src + 1)] = (n_src, 2.0, 1)
    prec = [(n_src, [(i, i) for i in range(n_src)]), (n_src + 1, [(n_src, n_src)])]
    return D.FlatDag.from_lists(events, acts, prec, max_delay, list(D.STD_DISTS))


@pytest.mark.parametrize("fan_in", [254, 300], ids=["byte_positions", "cause_events"])
def test_chunked_causes_at_the_position_limit(gpu, fan_in) -> None:
    """Chunks store the winning entry's position in a byte while no event has more than 254 entries, the cause
    event otherwise; the cause counts (and every other statistic) equal run_stats either way."""
    d = wide_fan_in(fan_in)
    p = capi.Propagator(d)
    seeds = (np.arange(6_000, dtype=np.int64) * 31 + 17).astype(np.int32)
    q = (0.25, 0.5, 0.9)
    srv = capi.Server(32 << 20, max_running=2, n_streams=2, request_cap=8 << 20)  # chunks of 512-1,152 samples
    cid = srv.add_context(p)
    got = srv.wait(srv.submit(cid, seeds, q))
    st = srv.stats()
    srv.close()
    same_stats(got, p.run_stats(seeds, q))
    assert st["chunked"] == 1 and st["failed"] == 0


def test_chunked_extremes_and_no_quantiles(gpu) -> None:
    """A chunked request whose order statistics all lie on an extreme's mass (quantile 1: rank n - 1 only; quantile
    0 would also read rank 1, with weight 0) resolves without a further pass; a chunked request with no quantiles
    needs no selection at all.  Both equal run_stats."""
    d = layered_network(layers=2)
    p = capi.Propagator(d)
    srv = capi.Server(48 << 20, max_running=2, n_streams=2, request_cap=8 << 20)
    cid = srv.add_context(p)
    seeds = np.arange(20_000, dtype=np.int32) * 13 + 1
    for q in ((1.0,), ()):
        got = srv.wait(srv.submit(cid, seeds, q))
        same_stats(got, p.run_stats(seeds, q))
    st = srv.stats()
    assert st["chunked"] == 2 and st["exact_rounds"] == 0 and st["failed"] == 0
    assert st["kept_rows"] == 0  # ranks 0 and n - 1 always lie on the extremes' masses: no row is kept for them
    srv.close()


def test_chunked_approximate_quantiles(gpu) -> None:
    """With exact selection off, a chunked request makes no further passes: mean, counts and the extremes
    are exact, every other quantile within one histogram bin (max_delay / 4096)."""
    d = layered_network(layers=10)
    p = capi.Propagator(d)
    srv = capi.Server(96 << 20, max_running=2, n_streams=2, request_cap=24 << 20, exact_chunked_quantiles=False)
    cid = srv.add_context(p)
    seeds = np.arange(5000, dtype=np.int32) * 3 + 5
    q = [0.0, 0.25, 0.5, 0.9, 0.99, 1.0]
    mean, quant, counts, _ = srv.wait(srv.submit(cid, seeds, q))
    w_mean, w_quant, w_counts = p.run_stats(seeds, q)
    np.testing.assert_array_equal(bits(mean), bits(w_mean))
    np.testing.assert_array_equal(counts, w_counts)
    assert np.max(np.abs(quant - w_quant)) <= d.max_delay / BINS * (1 + 1e-9) + 1e-9
    np.testing.assert_array_equal(bits(quant[:, [0, -1]]), bits(w_quant[:, [0, -1]]))
    st = srv.stats()
    assert st["chunked"] == 1 and st["exact_rounds"] == 0
    srv.close()


def test_chunked_request_needs_finite_max_delay(gpu) -> None:
    dag = D.continuous_dag(3, 60, max_delay=float("inf"), shared_links=True)
    p = capi.Propagator(dag)
    srv = capi.Server(8 << 20, max_running=2, n_streams=1, request_cap=1 << 20)
    cid = srv.add_context(p)
    with pytest.raises(capi.McdpError, match="finite max_delay"):
        srv.submit(cid, np.arange(20_000, dtype=np.int32))
    srv.close()


def test_served_results_match_oracle(gpu) -> None:
    """One request per resident context (in a mix, several on the device at once) against the oracle:
    the same seeds' device durations propagated by the CPU oracle (_core.cpp:332-350), then the
    oracle's per-event statistics (math.fsum mean, numpy.quantile, explicit cause counting; the
    consumer of demo/monte_carlo.py:45-63).  Bit for bit, including quantiles."""
    import oracle

    dags, props = contexts()
    srv = capi.Server(64 << 20, max_running=4, n_streams=3, request_cap=64 << 20)
    ids = [srv.add_context(p) for p in props]
    q = (0.0, 0.25, 0.5, 0.9, 0.99, 1.0)
    reqs = []
    for k in range(2):
        for c in range(len(props)):
            n = 700 + 333 * c + 101 * k
            seeds = np.arange(n, dtype=np.int32) * 5 + 1_000 * c - 77 * k
            reqs.append((srv.submit(ids[c], seeds, q), c, seeds))
    for rid, c, seeds in reqs:
        mean, quant, counts, _ = srv.wait(rid)
        _, dd, _ = capi.Propagator(dags[c]).run_many(seeds)
        r0, c0 = oracle.OracleCtx(dags[c]).propagate(dd)
        m0, q0, k0 = oracle.event_stats(r0, c0, oracle.pred_sources(dags[c]), q)
        np.testing.assert_array_equal(bits(mean), bits(m0))
        np.testing.assert_array_equal(bits(quant), bits(q0))
        np.testing.assert_array_equal(counts, k0)
    srv.close()


def test_chunked_request_matches_oracle(gpu) -> None:
    """A request above the request cap runs in chunks of samples: its mean, cause counts and quantiles equal
    the oracle's bit for bit."""
    import oracle

    dags, props = contexts()
    d, p = dags[1], props[1]
    n = 5000
    P = int(p.pred_offsets()[-1])
    cap = footprint(p.E, P, n, 4) // 2  # forces chunks (about a sixth of the samples, beside the 16 MiB histograms)
    srv = capi.Server(4 * cap, max_running=2, n_streams=2, request_cap=cap)
    cid = srv.add_context(p)
    seeds = np.arange(n, dtype=np.int32) + 12345
    q = (0.0, 0.5, 0.9, 1.0)
    mean, quant, counts, _ = srv.wait(srv.submit(cid, seeds, q))
    assert srv.stats()["chunked"] == 1
    srv.close()
    _, dd, _ = capi.Propagator(d).run_many(seeds)
    r0, c0 = oracle.OracleCtx(d).propagate(dd)
    m0, q0, k0 = oracle.event_stats(r0, c0, oracle.pred_sources(d), q)
    np.testing.assert_array_equal(bits(mean), bits(m0))
    np.testing.assert_array_equal(counts, k0)
    np.testing.assert_array_equal(bits(quant), bits(q0))
