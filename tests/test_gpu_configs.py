"""The benchmark configurations at their full sizes, on the GPU, against the oracle.

* Configuration 1 (the headline): the network DAG (20,000 events, 50,005 activities, 8 activity
  types), 100,000 samples, full realized / durations / cause outputs resident on the device.  Every
  one of the 100,000 samples -- the full rounds of tiles and the wide steps of the last partial
  round -- is propagated again by the CPU oracle from the device's own durations
  (orc_propagate_columns, pinned against orc_propagate, which is pinned against the compiled
  reference) and must match bit for bit; links that need no draw must carry the closed-form
  durations; realized must lie in [earliest, earliest + max_delay], each cause must be -1 or a
  predecessor, and the winner's capped candidate must equal realized.  Anchor: _core.cpp:312-361.
* Configuration 4: the pathwise gradient on the same DAG at 65,536 samples: sums split over four
  calls equal one call bit for bit; 256 samples spread over the range match the FP64 oracle
  gradient; central finite differences with common random numbers (h = 1e-3 s) over all 65,536
  samples match the pathwise derivative.
* Configuration 2: the 64-DAG batch (819,700 events, 4,096-65,536 samples per DAG) in one
  persistent launch: every DAG's statistics equal its own run's bit for bit; sample columns of
  every DAG equal its own run and the oracle's propagation; the smallest DAGs' statistics equal
  the oracle's statistics over the oracle's propagation of all their samples.
* Configuration 3: serving over contexts of 1,000-200,000 events with requests of 16,384-262,144
  samples, many run in chunks: results equal the same seeds run alone bit for bit (chunked
  quantiles too), cause counts of requests too large to run in one piece add up from pieces, and
  one request per context matches the oracle's statistics bit for bit.
"""

from __future__ import annotations

import ctypes as C
import gc
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import SplitMix64, batch_workload, layered_network, serve_contexts, serve_stream

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint64)


def log(msg: str) -> None:
    print(msg, flush=True)


def pull(buf: capi.DeviceBuffer, ld: int, c0: int, out: np.ndarray) -> np.ndarray:
    """Columns [c0, c0 + out.shape[1]) of an event-major device array with leading dimension ld into
    the rows of `out` (any row stride)."""
    item = out.dtype.itemsize
    rows, w = out.shape
    if rows and w:
        capi.check(capi.lib().mcdp_memcpy2d_dtoh(out.ctypes.data_as(C.c_void_p), out.strides[0],
                                                 C.c_void_p(buf.ptr.value + c0 * item), ld * item, w * item, rows))
    return out


def oracle_columns(ctx: oracle.OracleCtx, dur: np.ndarray, r_out: np.ndarray, c_out: np.ndarray) -> None:
    """The oracle's propagation of event-major durations (L, w), on host threads over column ranges."""
    w = dur.shape[1]
    step = max(64, -(-w // THREADS))
    with ThreadPoolExecutor(THREADS) as ex:
        list(ex.map(lambda a: ctx.propagate_columns(dur[:, a:a + step], r_out[:, a:a + step], c_out[:, a:a + step]),
                    range(0, w, step)))


def pred_table(dag) -> tuple[np.ndarray, np.ndarray]:
    """(K, E) tables of each event's predecessor sources and edges in list order (-1 / -1 padding)."""
    E = dag.n_events
    src = [[] for _ in range(E)]
    edge = [[] for _ in range(E)]
    for i, t in enumerate(dag.prec_target):
        a, b = int(dag.prec_off[i]), int(dag.prec_off[i + 1])
        src[int(t)] = [int(x) for x in dag.prec_src[a:b]]
        edge[int(t)] = [int(x) for x in dag.prec_edge[a:b]]
    K = max((len(s) for s in src), default=0)
    S = np.full((K, E), -1, dtype=np.int64)
    G = np.full((K, E), -1, dtype=np.int64)
    for e in range(E):
        S[:len(src[e]), e] = src[e]
        G[:len(edge[e]), e] = edge[e]
    return S, G


# ---- configuration 1 ------------------------------------------------------------------------
def test_config1_every_sample_matches_oracle(gpu) -> None:
    gc.collect()
    dag = layered_network()
    P = capi.Propagator(dag)
    E, L, S = P.E, P.L, 100_000
    assert (E, L) == (20_000, 50_005)
    sd = capi.DeviceBuffer.from_array(np.arange(S, dtype=np.int32))
    R, D, Cc = capi.DeviceBuffer(8 * E * S), capi.DeviceBuffer(8 * L * S), capi.DeviceBuffer(4 * E * S)
    t0 = time.perf_counter()
    P.run_many_device(sd, S, S, R, D, Cc)
    capi.synchronize()
    log(f"config1: run_many of {S} samples in {1000 * (time.perf_counter() - t0):.1f} ms, plan {P.plan_info()}")
    ctx = oracle.OracleCtx(dag)
    base = ctx.base_durations()
    fixed = np.nonzero(~np.isnan(base))[0]
    earliest = np.asarray(dag.earliest, dtype=np.float64)[:, None]
    upper = earliest + dag.max_delay
    src_t, edge_t = pred_table(dag)
    W = 12_500
    hr, hc, hd = np.empty((E, W)), np.empty((E, W), np.int32), np.empty((L, W))
    orr, oc = np.empty((E, W)), np.empty((E, W), np.int32)
    for c0 in range(0, S, W):
        w = min(W, S - c0)
        r, c, d, o_r, o_c = hr[:, :w], hc[:, :w], hd[:, :w], orr[:, :w], oc[:, :w]
        pull(R, S, c0, r)
        pull(Cc, S, c0, c)
        pull(D, S, c0, d)
        oracle_columns(ctx, d, o_r, o_c)
        np.testing.assert_array_equal(bits(r), bits(o_r), err_msg=f"realized, samples {c0}..{c0 + w - 1}")
        np.testing.assert_array_equal(c, o_c, err_msg=f"cause, samples {c0}..{c0 + w - 1}")
        np.testing.assert_array_equal(bits(d[fixed]), bits(np.broadcast_to(base[fixed][:, None], (fixed.size, w))))
        assert np.all(r >= earliest) and np.all(r <= upper)
        member = c == -1
        for k in range(src_t.shape[0]):
            member |= (c == src_t[k][:, None]) & (src_t[k][:, None] >= 0)
        assert member.all(), f"a cause outside its event's predecessors, samples {c0}..{c0 + w - 1}"
        if c0 == 0 or c0 + w == S:  # the first round of tiles and the wide steps' samples
            win = (c == -1) & (r == earliest)
            for k in range(src_t.shape[0]):
                has = src_t[k] >= 0
                s_k, g_k = np.where(has, src_t[k], 0), np.where(has, edge_t[k], 0)
                cand = r[s_k] + d[g_k]
                cand = np.where(upper < cand, upper, cand)
                win |= has[:, None] & (c == s_k[:, None]) & (cand == r)
            assert win.all(), "a winner whose capped candidate is not the realized time"
        log(f"config1: samples {c0}..{c0 + w - 1} bit-exact against the oracle")


# ---- configuration 4 ------------------------------------------------------------------------
def test_config4_gradient_full_size(gpu) -> None:
    gc.collect()
    dag = layered_network()
    P = capi.Propagator(dag)
    E, L, S = P.E, P.L, 65_536
    rng = SplitMix64(20260004)
    w_h = np.asarray([rng.uniform(0.5, 2.0) for _ in range(E)])
    seeds_h = np.arange(S, dtype=np.int32)
    sd, wt = capi.DeviceBuffer.from_array(seeds_h), capi.DeviceBuffer.from_array(w_h)
    R, D, Cc = capi.DeviceBuffer(8 * E * S), capi.DeviceBuffer(8 * L * S), capi.DeviceBuffer(4 * E * S)
    outs = [capi.DeviceBuffer(8), capi.DeviceBuffer(8 * L), capi.DeviceBuffer(8 * E)]
    lib = capi.lib()

    def reverse(c0: int, n: int, accumulate: bool, dst) -> None:
        capi.check(lib.mcdp_reverse_device(P.h, C.c_void_p(sd.ptr.value + 4 * c0), wt.ptr,
                                           C.c_void_p(R.ptr.value + 8 * c0), C.c_void_p(D.ptr.value + 8 * c0),
                                           C.c_void_p(Cc.ptr.value + 4 * c0), n, S, int(accumulate), dst[0].ptr,
                                           dst[1].ptr, dst[2].ptr, None))

    def read(dst):
        return (dst[0].to_array(np.float64, (1,)), dst[1].to_array(np.float64, (L,)),
                dst[2].to_array(np.float64, (E,)))

    P.run_many_device(sd, S, S, R, D, Cc)
    reverse(0, S, False, outs)
    capi.synchronize()
    full = read(outs)
    log(f"config4: J = {full[0][0] / S:.6f}, gradient info {P.gradient_info()}")
    # (1) four calls of 16,384 samples (whole reduction groups) accumulate to the same bits
    for k in range(4):
        reverse(k * 16_384, 16_384, k > 0, outs)
    capi.synchronize()
    for a, b in zip(read(outs), full):
        np.testing.assert_array_equal(bits(a), bits(b))
    # (2) 256 samples spread over the range against the FP64 oracle gradient
    ctx = oracle.OracleCtx(dag)
    blocks = [0, 21_000, 43_000, S - 64]
    for i, c0 in enumerate(blocks):
        reverse(c0, 64, i > 0, outs)
    capi.synchronize()
    dev = read(outs)
    r = np.concatenate([pull(R, S, c0, np.empty((E, 64))) for c0 in blocks], axis=1).T
    c = np.concatenate([pull(Cc, S, c0, np.empty((E, 64), np.int32)) for c0 in blocks], axis=1).T
    d = np.concatenate([pull(D, S, c0, np.empty((L, 64))) for c0 in blocks], axis=1).T
    o_obj, o_gd, o_ge = ctx.gradient(r, c, d, w_h)
    for got, want in zip(dev, ([o_obj], o_gd, o_ge)):
        want = np.asarray(want, dtype=np.float64)
        scale = max(1.0, float(np.max(np.abs(want))))
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * scale)
    # (3) central finite differences with common random numbers over all 65,536 samples
    gd, ge = full[1], full[2]
    dmin0, e0 = P.get_times()
    pm = capi.Propagator(dag)
    tmp = [capi.DeviceBuffer(8), capi.DeviceBuffer(8 * L), capi.DeviceBuffer(8 * E)]
    h = 1e-3

    def objective(dmin, earliest) -> float:
        pm.set_times(dmin, earliest)
        pm.run_many_device(sd, S, S, R, D, Cc)
        capi.check(lib.mcdp_reverse_device(pm.h, sd.ptr, wt.ptr, R.ptr, D.ptr, Cc.ptr, S, S, 0, tmp[0].ptr, tmp[1].ptr,
                                           tmp[2].ptr, None))
        capi.synchronize()
        return float(tmp[0].to_array(np.float64, (1,))[0])

    links = list(np.argsort(-np.abs(gd))[:3]) + [int(np.nonzero(gd)[0][len(np.nonzero(gd)[0]) // 2])]
    for l in links:
        dp, dm = dmin0.copy(), dmin0.copy()
        dp[l] += h
        dm[l] -= h
        fd = (objective(dp, e0) - objective(dm, e0)) / (dp[l] - dm[l])
        log(f"config4: link {l}: finite difference {fd:.6f}, pathwise {gd[l]:.6f}")
        assert fd == pytest.approx(gd[l], rel=2e-3, abs=1e-3 * S)
    events = list(np.argsort(-np.abs(ge))[:3]) + [int(np.nonzero(ge)[0][len(np.nonzero(ge)[0]) // 2])]
    for e in events:
        ep, em = e0.copy(), e0.copy()
        ep[e] += h
        em[e] -= h
        fd = (objective(dmin0, ep) - objective(dmin0, em)) / (ep[e] - em[e])
        log(f"config4: event {e}: finite difference {fd:.6f}, pathwise {ge[e]:.6f}")
        assert fd == pytest.approx(ge[e], rel=2e-3, abs=1e-3 * S)


# ---- configuration 2 ------------------------------------------------------------------------
def test_config2_batch_full_size(gpu) -> None:
    gc.collect()
    dags, ns = batch_workload()
    seeds = [np.arange(n, dtype=np.int32) for n in ns]
    q = (0.5, 0.9, 0.99)
    # (1) one persistent launch of all 64 DAGs (200 GiB budget), statistics on the device, against
    #     each DAG's own run and statistics
    props = [capi.Propagator(d) for d in dags]
    B = capi.Batch(props)
    t0 = time.perf_counter()
    got = B.stats(seeds, q, memory_budget=200 << 30)
    log(f"config2: batched statistics of {len(dags)} DAGs in {time.perf_counter() - t0:.2f} s (first call)")
    del B
    gc.collect()
    for k, (d, s) in enumerate(zip(dags, seeds)):
        solo = capi.Propagator(d)
        m0, q0, c0 = solo.run_stats(s, q)
        m, qq, cc = got[k]
        np.testing.assert_array_equal(bits(m), bits(m0), err_msg=f"DAG {k} mean")
        np.testing.assert_array_equal(bits(qq), bits(q0), err_msg=f"DAG {k} quantiles")
        np.testing.assert_array_equal(cc, c0, err_msg=f"DAG {k} cause counts")
        del solo
    log("config2: statistics of all 64 DAGs equal their own runs'")
    # (2) the batch's per-sample rows (realized + cause of every DAG, ~198 GB) against sample columns
    #     of each DAG's own run and the oracle
    B = capi.Batch(props)
    sd = [capi.DeviceBuffer.from_array(s) for s in seeds]
    R = [capi.DeviceBuffer(8 * p.E * n) for p, n in zip(props, ns)]
    Cc = [capi.DeviceBuffer(4 * p.E * n) for p, n in zip(props, ns)]
    B.run_many_device(sd, ns, ns, R, None, Cc)
    capi.synchronize()
    for k, (d, n) in enumerate(zip(dags, ns)):
        E = props[k].E
        cols = sorted({0, n // 3, (2 * n) // 3, n - 64})
        r = np.concatenate([pull(R[k], n, c0, np.empty((E, 64))) for c0 in cols], axis=1).T
        c = np.concatenate([pull(Cc[k], n, c0, np.empty((E, 64), np.int32)) for c0 in cols], axis=1).T
        s = np.concatenate([seeds[k][c0:c0 + 64] for c0 in cols])
        solo = capi.Propagator(d)
        r0, d0, c_0 = solo.run_many(s)
        del solo
        np.testing.assert_array_equal(bits(r), bits(r0), err_msg=f"DAG {k}")
        np.testing.assert_array_equal(c, c_0, err_msg=f"DAG {k}")
        ro, co = oracle.OracleCtx(d).propagate(d0)
        np.testing.assert_array_equal(bits(r), bits(ro), err_msg=f"DAG {k} vs oracle")
        np.testing.assert_array_equal(c, co, err_msg=f"DAG {k} vs oracle")
    log("config2: sample columns of all 64 DAGs equal their own runs and the oracle")
    # (3) the smallest DAGs: the oracle's statistics over its propagation of every sample
    small = sorted(range(len(dags)), key=lambda k: props[k].E * ns[k])[:3]
    for k in small:
        d, n = dags[k], ns[k]
        E = props[k].E
        solo = capi.Propagator(d)
        _, d0, _ = solo.run_many(seeds[k])
        del solo
        ro, co = oracle.OracleCtx(d).propagate(d0)
        r = pull(R[k], n, 0, np.empty((E, n)))
        np.testing.assert_array_equal(bits(r), bits(ro.T))
        m0, q0, k0 = oracle.event_stats(ro, co, oracle.pred_sources(d), q)
        m, qq, cc = got[k]
        np.testing.assert_array_equal(bits(m), bits(m0))
        np.testing.assert_array_equal(bits(qq), bits(q0))
        np.testing.assert_array_equal(cc, k0)
        log(f"config2: DAG {k} ({E} events x {n} samples) statistics equal the oracle's")


# ---- configuration 3 ------------------------------------------------------------------------
def test_config3_serving_full_size(gpu) -> None:
    gc.collect()
    all_ctx = serve_contexts()
    order = np.argsort([d.n_events for d in all_ctx])
    pick = sorted(set(order[np.linspace(0, len(order) - 1, 24).round().astype(int)].tolist()))
    dags = [all_ctx[i] for i in pick]
    props = [capi.Propagator(d) for d in dags]
    sizes = [p.E for p in props]
    assert min(sizes) < 1500 and max(sizes) > 150_000
    budget, cap = 160 << 30, 16 << 30
    srv = capi.Server(budget, max_running=8, n_streams=4, request_cap=cap)
    ids = [srv.add_context(p) for p in props]
    q = (0.5, 0.9, 0.99)
    stream = serve_stream(len(props), 40, 40.0)
    reqs = []
    t0 = time.perf_counter()
    for i in range(len(stream.samples)):
        delay = t0 + stream.arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        c = int(stream.context[i])
        seeds = stream.seeds(i)
        reqs.append((srv.submit(ids[c], seeds, q), c, seeds))
        if i < len(props):  # one small request per context for the oracle leg, in the same mix
            s_small = np.arange(5_000_000 + 1000 * i, 5_000_000 + 1000 * i + 64, dtype=np.int32)
            reqs.append((srv.submit(ids[i], s_small, (0.0, 0.5, 1.0)), i, s_small))
    results = [(srv.wait(rid), c, s) for rid, c, s in reqs]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.asarray([r[0][3] for r in results])
    log(f"config3: {len(results)} requests over {len(props)} contexts ({sum(sizes)} events) in {wall:.2f} s, "
        f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms; server {st}")
    assert st["failed"] == 0 and st["chunked"] >= 5 and st["max_running"] >= 2 and st["peak_bytes"] <= budget
    gc.collect()
    n_alone = n_pieces = n_oracle = n_chunked_alone = 0
    for (mean, quant, counts, _), c, seeds in results:
        P, d, n = props[c], dags[c], seeds.size
        E = P.E
        if n == 64:  # the oracle leg
            _, d0, _ = P.run_many(seeds)
            ro, co = oracle.OracleCtx(d).propagate(d0)
            m0, q0, k0 = oracle.event_stats(ro, co, oracle.pred_sources(d), (0.0, 0.5, 1.0))
            np.testing.assert_array_equal(bits(mean), bits(m0))
            np.testing.assert_array_equal(bits(quant), bits(q0))
            np.testing.assert_array_equal(counts, k0)
            n_oracle += 1
            continue
        off = P.pred_offsets()
        earliest = np.asarray(d.earliest)
        # every sample lies in [earliest, earliest + max_delay]; the mean (exactly rounded sum / n) within
        # a few ulps of that range (the sum of n copies of a time rounds once, then divides)
        tol = 4 * np.spacing(np.abs(earliest) + d.max_delay)
        assert np.all(mean >= earliest - tol) and np.all(mean <= earliest + d.max_delay + tol)
        assert np.all(np.diff(quant, axis=1) >= 0)
        per_event = np.add.reduceat(counts, off[:-1] + np.arange(E))
        assert np.all(per_event == n)
        chunked = 12 * E * n > cap
        if 12 * E * n <= (40 << 30):  # one piece fits beside the test's other buffers: run it alone
            m0, q0, c0 = P.run_stats(seeds, q)
            np.testing.assert_array_equal(bits(mean), bits(m0))
            np.testing.assert_array_equal(counts, c0)
            np.testing.assert_array_equal(bits(quant), bits(q0))  # chunked requests too: exact selection
            n_chunked_alone += int(chunked)
            n_alone += 1
        elif n_pieces < 2:  # too large to run in one piece: cause counts add up over pieces
            k = math.ceil(12 * E * n / (32 << 30))
            total = np.zeros_like(counts)
            for part in np.array_split(seeds, k):
                total += P.run_stats(part, ())[2]
            np.testing.assert_array_equal(counts, total)
            n_pieces += 1
    log(f"config3: {n_alone} requests ({n_chunked_alone} chunked) equal their seeds run alone, {n_pieces} chunked "
        f"requests' counts add up over pieces, {n_oracle} requests equal the oracle")
    assert n_oracle == len(props) and n_alone >= 10 and n_chunked_alone >= 1
