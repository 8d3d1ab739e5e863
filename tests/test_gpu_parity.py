"""GPU parity: the gfx950 path against the oracle, bit for bit.

Every test calls libmcdp through its C ABI (directly via ctypes, or through the
pybind11 module, which is a thin layer over the same ABI).

* Philox words: device == oracle, bit-exact.
* Propagation with injected durations: device == the reference's own results
  (golden fixtures) and == the oracle on edge-case DAGs and durations
  (NaN, +-inf, -0.0, ties, caps, duplicate targets, index gaps, edges outside
  the link range, wide levels that spill past LDS).
* Fused sampling + propagation (run_many): the device's own durations, fed to the
  oracle, reproduce the device's realized/cause bit-exactly; durations that need
  no draw equal the oracle's closed form bit-exactly.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import dags as D
import oracle
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import CONSTANT, EMPIRICAL_ABSOLUTE, DistSpec, FlatDag, layered_network, layered_small

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
GOLDEN_FILES = sorted(GOLDEN.glob("*.npz"))
INF = float("inf")


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint64)


def test_device_philox_matches_oracle(gpu) -> None:
    rng = np.random.default_rng(1)
    n = 4096
    ctr4 = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    key4 = rng.integers(0, 2**32, (n, 2), dtype=np.uint64).astype(np.uint32)
    ctr4[:3] = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    key4[:3] = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    np.testing.assert_array_equal(capi.debug_philox(4, ctr4, key4), oracle.philox4x32(ctr4, key4))
    ctr2 = ctr4[:, :2].copy()
    key2 = key4[:, 0].copy()
    np.testing.assert_array_equal(capi.debug_philox(2, ctr2, key2), oracle.philox2x32(ctr2, key2))
    assert tuple(capi.debug_philox(4, ctr4[:1], key4[:1])[0]) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)


@pytest.mark.parametrize("path", GOLDEN_FILES, ids=[p.stem for p in GOLDEN_FILES])
def test_injected_matches_reference_fixture(gpu, path: Path) -> None:
    """Given the reference's own durations, device realized/cause equal the reference's, bit for bit."""
    z = np.load(path)
    dag = D.dag_from_npz(z)
    P = capi.Propagator(dag)
    r, c = P.propagate(z["durations"])
    np.testing.assert_array_equal(bits(r), bits(z["realized"]))
    np.testing.assert_array_equal(c, z["cause"])


@pytest.mark.parametrize("path", GOLDEN_FILES, ids=[p.stem for p in GOLDEN_FILES])
def test_run_many_consistent_on_fixture_dags(gpu, path: Path) -> None:
    """Device durations -> oracle propagation reproduces device realized/cause; fixed durations match."""
    z = np.load(path)
    dag = D.dag_from_npz(z)
    P = capi.Propagator(dag)
    seeds = np.concatenate([z["seeds"], np.arange(100, 200)]).astype(np.int32)
    r, d, c = P.run_many(seeds)
    ctx = oracle.OracleCtx(dag)
    rr, cc = ctx.propagate(d)
    np.testing.assert_array_equal(bits(r), bits(rr))
    np.testing.assert_array_equal(c, cc)
    base = ctx.base_durations()
    fixed = ~np.isnan(base)
    np.testing.assert_array_equal(bits(d[:, fixed]), bits(np.broadcast_to(base[fixed], d[:, fixed].shape)))
    # Realized times stay inside [earliest, earliest + max_delay] (_core.cpp:336, 348) unless inputs are NaN.
    e = dag.earliest[None, :]
    assert np.all((r >= e) & (r <= e + dag.max_delay))


@pytest.mark.parametrize("case", range(12))
def test_injected_special_values(gpu, case: int) -> None:
    rng = np.random.default_rng(100 + case)
    kw = [{}, {"duplicates": True}, {"gaps": True}, {"oob_edges": True}][case % 4]
    dag = D.random_dag(rng, int(rng.integers(3, 150)), max_preds=int(rng.integers(1, 7)), **kw)
    P = capi.Propagator(dag)
    ctx = oracle.OracleCtx(dag)
    n = int(rng.choice([1, 63, 64, 65, 200]))
    dur = D.special_durations(rng, n, ctx.L)
    r, c = P.propagate(dur)
    rr, cc = ctx.propagate(dur)
    np.testing.assert_array_equal(bits(r), bits(rr))
    np.testing.assert_array_equal(c, cc)


def _wide_dag(n_src: int, depth: int) -> FlatDag:
    """n_src sources all read again `depth` levels later: more live values than LDS slots."""
    events = [(float(i % 7), 1e9, 0.0) for i in range(n_src)]
    acts, prec = {}, []
    chain_prev = None
    idx = 0
    for k in range(depth):  # a chain that creates the levels
        e = len(events)
        events.append((0.0, 1e9, 0.0))
        if chain_prev is not None:
            acts[(chain_prev, e)] = (idx, 1.0, 0)
            prec.append((e, [(chain_prev, idx)]))
            idx += 1
        chain_prev = e
    for i in range(n_src):
        e = len(events)
        events.append((0.0, 1e9, 0.0))
        acts[(i, e)] = (idx, 2.0, 0)
        acts[(chain_prev, e)] = (idx + 1, 0.5, 0)
        prec.append((e, [(i, idx), (chain_prev, idx + 1)]))
        idx += 2
    return FlatDag.from_lists(events, acts, prec, 1e6, [DistSpec(0, CONSTANT, (0.25, 0, INF))])


def test_global_memory_fallback(gpu) -> None:
    dag = _wide_dag(600, 5)
    P = capi.Propagator(dag)
    info = P.plan_info()
    assert info["global_preds"] > 0, info
    ctx = oracle.OracleCtx(dag)
    rng = np.random.default_rng(3)
    dur = np.round(rng.uniform(0, 9, (130, ctx.L)))
    r, c = P.propagate(dur)
    rr, cc = ctx.propagate(dur)
    np.testing.assert_array_equal(bits(r), bits(rr))
    np.testing.assert_array_equal(c, cc)
    r, d, c = P.run_many(np.arange(70))
    rr, cc = ctx.propagate(d)
    np.testing.assert_array_equal(bits(r), bits(rr))
    np.testing.assert_array_equal(c, cc)
    # run_stats walks the same plan storing one-byte winner positions (the general path's variant)
    seeds = np.arange(300)
    q = (0.1, 0.5, 0.9)
    mean, quant, counts = P.run_stats(seeds, q)
    r, _, c = P.run_many(seeds)
    om, oq, oc = oracle.event_stats(r, c, oracle.pred_sources(dag), q)
    np.testing.assert_array_equal(bits(mean), bits(om))
    np.testing.assert_array_equal(quant, oq)
    np.testing.assert_array_equal(counts, oc)


def test_device_api_leading_dimension(gpu) -> None:
    """Device entry points with ld > n write only the first n columns of each row."""
    dag = layered_small()
    P = capi.Propagator(dag)
    n, ld = 1000, 1088
    seeds = capi.DeviceBuffer.from_array(np.arange(n, dtype=np.int32))
    sentinel = np.full((P.E, ld), -7.0)
    real = capi.DeviceBuffer.from_array(sentinel)
    dur = capi.DeviceBuffer.from_array(np.full((P.L, ld), -7.0))
    cause = capi.DeviceBuffer.from_array(np.full((P.E, ld), -9, dtype=np.int32))
    P.run_many_device(seeds, n, ld, real, dur, cause)
    capi.synchronize()
    R = real.to_array(np.float64, (P.E, ld))
    Dd = dur.to_array(np.float64, (P.L, ld))
    Cc = cause.to_array(np.int32, (P.E, ld))
    assert np.all(R[:, n:] == -7.0) and np.all(Dd[:, n:] == -7.0) and np.all(Cc[:, n:] == -9)
    hr, hd, hc = P.run_many(np.arange(n))
    np.testing.assert_array_equal(bits(R[:, :n].T), bits(hr))
    np.testing.assert_array_equal(bits(Dd[:, :n].T), bits(hd))
    np.testing.assert_array_equal(Cc[:, :n].T, hc)
    # injected path on the same buffers
    r2 = capi.DeviceBuffer(P.E * ld * 8)
    c2 = capi.DeviceBuffer(P.E * ld * 4)
    P.propagate_device(dur, n, ld, r2, c2)
    capi.synchronize()
    np.testing.assert_array_equal(bits(r2.to_array(np.float64, (P.E, ld))[:, :n]), bits(R[:, :n]))
    np.testing.assert_array_equal(c2.to_array(np.int32, (P.E, ld))[:, :n], Cc[:, :n])


def test_seed_is_the_key(gpu) -> None:
    """A sample's result depends on its seed only: not on its position, batch size or neighbours."""
    dag = layered_small()
    P = capi.Propagator(dag)
    seeds = np.array([5, -1, 2**31 - 1, -(2**31), 0, 5, 77] * 20, dtype=np.int32)
    r, d, c = P.run_many(seeds)
    for i, s in enumerate(seeds[:7]):
        r1, d1, c1 = P.run_many([s])
        np.testing.assert_array_equal(bits(r1[0]), bits(r[i]))
        np.testing.assert_array_equal(bits(d1[0]), bits(d[i]))
        np.testing.assert_array_equal(c1[0], c[i])
    np.testing.assert_array_equal(bits(r[0]), bits(r[5]))
    assert not np.array_equal(d[0], d[1])


def test_network_slice_consistent(gpu) -> None:
    """A 40-layer slice of the network workload (8 activity types), 300 samples."""
    dag = layered_network(layers=40)
    P = capi.Propagator(dag)
    r, d, c = P.run_many(np.arange(300))
    ctx = oracle.OracleCtx(dag)
    rr, cc = ctx.propagate(d)
    np.testing.assert_array_equal(bits(r), bits(rr))
    np.testing.assert_array_equal(c, cc)
    base = ctx.base_durations()
    fixed = ~np.isnan(base)
    np.testing.assert_array_equal(bits(d[:, fixed]), bits(np.broadcast_to(base[fixed], d[:, fixed].shape)))
    assert np.all(np.isfinite(d)) and np.all(d >= dag.act_dur[None, :])


def _host_chunk(dag: FlatDag) -> int:
    """Samples per device chunk of the host calls: 2 GiB of results and seeds, rounded down to whole waves."""
    per_sample = 12 * dag.n_events + 8 * len(dag.act_idx) + 8
    return (2**31 // per_sample) // 64 * 64


@pytest.mark.parametrize("name", ["network", "long_chain"])
def test_host_run_many_over_several_chunks(gpu, name: str) -> None:
    """A host run_many over more samples than one device chunk equals the same seeds in calls of at most one chunk,
    bit for bit, and so does propagating its durations.  Results and injected durations move between the caller's
    arrays and the device in pieces through pinned staging buffers: row groups of the network DAG (26 KB device
    rows), and rows longer than a staging buffer, split, for the 100-event chain (8.6 MB device rows of its
    1,073,728-sample chunks)."""
    dag = layered_network() if name == "network" else D.parity_long_chain(99)
    P = capi.Propagator(dag)
    chunk = _host_chunk(dag)
    n = chunk + (700 if name == "network" else 5000)
    seeds = np.arange(n, dtype=np.int32)
    r, d, c = P.run_many(seeds)
    for lo, hi in ((0, chunk), (chunk, n)):
        r1, d1, c1 = P.run_many(seeds[lo:hi])
        np.testing.assert_array_equal(bits(r1), bits(r[lo:hi]))
        np.testing.assert_array_equal(bits(d1), bits(d[lo:hi]))
        np.testing.assert_array_equal(c1, c[lo:hi])
        del r1, d1, c1
    # injected durations over several chunks, uploaded through the same staging: the same results again
    rr, cc = P.propagate(d)
    np.testing.assert_array_equal(bits(rr), bits(r))
    np.testing.assert_array_equal(cc, c)


@pytest.mark.parametrize("name", ["network_slice", "long_chain"])
def test_short_host_runs_replay_a_graph(gpu, name: str) -> None:
    """Host runs of a few samples on plans of 32 steps or more: the first call with a given sample count launches
    directly, the second captures its launches into a graph, later calls replay it (the network slice's 41 wide
    launches; the 100-step chain's one tile launch).  Every call, with other work and synchronizations between
    them, equals the same seeds in one direct run_many, bit for bit; injected runs too, and a change of times is
    not hidden by a replay."""
    if name == "network_slice":
        dag = layered_network(layers=40)
    else:
        dag = D.parity_long_chain(99)
    P = capi.Propagator(dag)
    other = capi.Propagator(layered_small())
    seeds = np.arange(48, dtype=np.int32) * 13 + 5
    r0, d0, c0 = P.run_many(seeds)  # 48 samples: a direct launch
    for k in range(8):
        sl = slice(6 * k, 6 * k + 6)
        r, d, c = P.run_many(seeds[sl])
        np.testing.assert_array_equal(bits(r), bits(r0[sl]), err_msg=f"call {k}")
        np.testing.assert_array_equal(bits(d), bits(d0[sl]), err_msg=f"call {k}")
        np.testing.assert_array_equal(c, c0[sl], err_msg=f"call {k}")
        rr, cc = P.propagate(d0[sl])
        np.testing.assert_array_equal(bits(rr), bits(r0[sl]), err_msg=f"injected call {k}")
        np.testing.assert_array_equal(cc, c0[sl], err_msg=f"injected call {k}")
        other.run_many(np.arange(50 + k))  # other work between the replays
        capi.synchronize()
    # new times: the next calls use the new plan, not the graph of the old one
    _, e0 = P.get_times()
    earliest = e0 + 7.0
    P.set_times(earliest=earliest)
    moved = FlatDag(**{**dag.__dict__})
    moved.earliest = earliest.copy()
    rf, df, cf = capi.Propagator(moved).run_many(seeds[:6])
    for k in range(3):
        r, d, c = P.run_many(seeds[:6])
        np.testing.assert_array_equal(bits(r), bits(rf), err_msg=f"call {k} after new times")
        np.testing.assert_array_equal(bits(d), bits(df), err_msg=f"call {k} after new times")
        np.testing.assert_array_equal(c, cf, err_msg=f"call {k} after new times")


SHAPES = [(1, 4, 1), (2, 16, 1), (5, 8, 1), (16, 12, 1), (16, 16, 1), (1, 4, 2), (3, 8, 2), (7, 4, 2), (16, 8, 2)]


@pytest.mark.parametrize("shape", SHAPES, ids=[f"w{w}k{k}s{s}" for w, k, s in SHAPES])
def test_results_independent_of_schedule_shape(gpu, shape) -> None:
    """Every schedule shape (waves per workgroup, events per wave and step, samples per lane) gives
    the same bits; 200 samples leave a partial last tile in both tile sizes (64 and 128)."""
    dag = layered_network(layers=12)
    seeds = np.arange(-50, 150)
    ref = capi.Propagator(dag)
    r0, d0, c0 = ref.run_many(seeds)
    P = capi.Propagator(dag)
    P.configure(*shape)
    info = P.plan_info()
    assert (info["waves"], info["events_per_wave"], info["samples_per_lane"]) == shape
    r, d, c = P.run_many(seeds)
    np.testing.assert_array_equal(bits(r), bits(r0))
    np.testing.assert_array_equal(bits(d), bits(d0))
    np.testing.assert_array_equal(c, c0)
    rr, cc = P.propagate(d0)
    np.testing.assert_array_equal(bits(rr), bits(r0))
    np.testing.assert_array_equal(cc, c0)


def test_autotune_keeps_results(gpu) -> None:
    dag = layered_network(layers=12)
    seeds = np.arange(300)
    P = capi.Propagator(dag)
    r0, d0, c0 = P.run_many(seeds)
    w, k, spl, ms = P.autotune(4096)
    assert 1 <= w <= 16 and k in (4, 8, 12, 16) and spl in (1, 2) and ms > 0
    info = P.plan_info()
    assert (info["waves"], info["events_per_wave"], info["samples_per_lane"]) == (w, k, spl)
    r, d, c = P.run_many(seeds)
    np.testing.assert_array_equal(bits(r), bits(r0))
    np.testing.assert_array_equal(bits(d), bits(d0))
    np.testing.assert_array_equal(c, c0)


def test_autotune_choice_persists_in_a_file(gpu, tmp_path) -> None:
    """The first tuning times the shapes and appends one line; a new propagator of the same structure (other
    times) reads it back without timing; unparsable lines and other keys are skipped; a recorded choice for
    another sample-count bucket does not match."""
    dag = layered_network(layers=12)
    path = tmp_path / "shapes.txt"
    path.write_text("garbage line\nmcdp-shape 1 0000000000000000 gfx000 1 4096 4 4 1 1.0\n")
    P = capi.Propagator(dag)
    w, k, spl, ms = P.autotune(4096, path=str(path))
    assert not P.last_tune_from_file and ms > 0
    lines = path.read_text().splitlines()
    assert len(lines) == 3
    f = lines[-1].split()
    assert f[:3] == ["mcdp-shape", "1", f"{P.fingerprint():016x}"] and f[2] != "0000000000000000"
    assert f[5] == "4096" and (int(f[6]), int(f[7]), int(f[8])) == (w, k, spl) and float(f[9]) > 0
    # same structure, other times: the recorded shape, without timing
    d2 = FlatDag(**{**dag.__dict__})
    d2.act_dur = dag.act_dur * 1.5
    Q = capi.Propagator(d2)
    assert Q.fingerprint() == P.fingerprint()
    assert Q.autotune(3000, path=str(path))[:3] == (w, k, spl) and Q.last_tune_from_file
    info = Q.plan_info()
    assert (info["waves"], info["events_per_wave"], info["samples_per_lane"]) == (w, k, spl)
    # a later line for the key wins
    other = (8, 4, 1) if (w, k, spl) != (8, 4, 1) else (16, 4, 1)
    with open(path, "a") as fh:
        fh.write(" ".join(f[:6]) + f" {other[0]} {other[1]} {other[2]} 9.5\n")
    R = capi.Propagator(dag)
    assert R.autotune(4096, path=str(path)) == (*other, 9.5) and R.last_tune_from_file
    # another bucket of sample counts tunes again and appends
    R.autotune(5000, path=str(path))
    assert not R.last_tune_from_file and len(path.read_text().splitlines()) == 5
    seeds = np.arange(200)
    np.testing.assert_array_equal(bits(R.run_many(seeds)[0]), bits(P.run_many(seeds)[0]))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 127, 128, 129, 300])
def test_two_samples_per_lane_edges(gpu, n: int) -> None:
    """Tiles of 128 samples: sizes around the tile and half-tile boundaries, device API with ld > n,
    injected durations, all against the one-sample-per-lane kernel."""
    dag = layered_small()
    A = capi.Propagator(dag)
    A.configure(8, 16, 1)
    B = capi.Propagator(dag)
    B.configure(8, 8, 2)
    seeds = np.arange(1000, 1000 + n)
    ra, da, ca = A.run_many(seeds)
    rb, db, cb = B.run_many(seeds)
    np.testing.assert_array_equal(bits(ra), bits(rb))
    np.testing.assert_array_equal(bits(da), bits(db))
    np.testing.assert_array_equal(ca, cb)
    rr, cc = B.propagate(da)
    np.testing.assert_array_equal(bits(rr), bits(ra))
    np.testing.assert_array_equal(cc, ca)
    ld = n + 3  # odd leading dimension
    sd = capi.DeviceBuffer.from_array(seeds.astype(np.int32))
    real = capi.DeviceBuffer.from_array(np.full((B.E, ld), -7.0))
    dur = capi.DeviceBuffer.from_array(np.full((B.L, ld), -7.0))
    cause = capi.DeviceBuffer.from_array(np.full((B.E, ld), -9, dtype=np.int32))
    B.run_many_device(sd, n, ld, real, dur, cause)
    capi.synchronize()
    R = real.to_array(np.float64, (B.E, ld))
    Dd = dur.to_array(np.float64, (B.L, ld))
    Cc = cause.to_array(np.int32, (B.E, ld))
    assert np.all(R[:, n:] == -7.0) and np.all(Dd[:, n:] == -7.0) and np.all(Cc[:, n:] == -9)
    np.testing.assert_array_equal(bits(R[:, :n].T), bits(ra))
    np.testing.assert_array_equal(bits(Dd[:, :n].T), bits(da))
    np.testing.assert_array_equal(Cc[:, :n].T, ca)


def test_tail_split_matches(gpu) -> None:
    """Two samples per lane with a short last round: the samples past the last full round run in the
    one-sample-per-lane kernel (a second launch on the same stream); every sample matches the
    one-sample-per-lane run and the oracle."""
    dag = layered_small()
    n = 3 * 256 * 128 + 77  # three full rounds of 256 resident tiles of 128, plus 77 samples
    seeds = np.arange(n, dtype=np.int32) - 1000
    A = capi.Propagator(dag)
    A.configure(16, 8, 1)
    B = capi.Propagator(dag)
    B.configure(16, 8, 2)
    ra, da, ca = A.run_many(seeds)
    rb, db, cb = B.run_many(seeds)
    np.testing.assert_array_equal(bits(rb), bits(ra))
    np.testing.assert_array_equal(bits(db), bits(da))
    np.testing.assert_array_equal(cb, ca)
    rr, cc = B.propagate(da)
    np.testing.assert_array_equal(bits(rr), bits(ra))
    np.testing.assert_array_equal(cc, ca)
    pick = np.r_[0:64, n - 200:n]
    r0, c0 = oracle.OracleCtx(dag).propagate(db[pick])
    np.testing.assert_array_equal(bits(r0), bits(rb[pick]))
    np.testing.assert_array_equal(c0, cb[pick])


PARTIAL_ROUND_DAGS = {
    "layered_small": lambda: layered_small(),
    "edge_cases": lambda: D.continuous_dag(7, 60, max_delay=12.0, shared_links=True),
}


@pytest.mark.parametrize("name", sorted(PARTIAL_ROUND_DAGS))
def test_partial_round_routes_match(gpu, name: str) -> None:
    """The samples past the last full round of tiles run as one-sample-per-lane tiles (route 1) or as wide
    steps, one launch per step with every event a wave of its own (route 2); both, and the automatic choice,
    give the bits of a run with no split, and the oracle agrees on the last samples.  The edge-case DAG has
    stale reads, repeated targets and sources, shared links, gaps and edges outside the link table."""
    dag = PARTIAL_ROUND_DAGS[name]()
    n = 256 * 128 + 77  # one full round of 256 resident tiles of 128 on a 256-CU device, plus 77 samples
    seeds = np.arange(n, dtype=np.int32) * 3 - 5000
    A = capi.Propagator(dag)
    A.configure(16, 8, 1)
    ra, da, ca = A.run_many(seeds)
    for mode in (1, 2, 0):
        B = capi.Propagator(dag)
        B.configure(16, 8, 2)
        capi.check(capi.lib().mcdp_debug_tail_mode(B.h, mode))
        rb, db, cb = B.run_many(seeds)
        np.testing.assert_array_equal(bits(rb), bits(ra), err_msg=f"route {mode}")
        np.testing.assert_array_equal(bits(db), bits(da), err_msg=f"route {mode}")
        np.testing.assert_array_equal(cb, ca, err_msg=f"route {mode}")
        rr, cc = B.propagate(da)
        np.testing.assert_array_equal(bits(rr), bits(ra), err_msg=f"route {mode} injected")
        np.testing.assert_array_equal(cc, ca, err_msg=f"route {mode} injected")
    pick = np.r_[0:64, n - 200:n]
    r0, c0 = oracle.OracleCtx(dag).propagate(da[pick])
    np.testing.assert_array_equal(bits(r0), bits(ra[pick]))
    np.testing.assert_array_equal(c0, ca[pick])


def test_tail_mode_rejects_unknown_routes(gpu) -> None:
    P = capi.Propagator(layered_small())
    assert capi.lib().mcdp_debug_tail_mode(P.h, 3) != 0
    assert capi.lib().mcdp_debug_tail_mode(P.h, -1) != 0


def test_run_many_with_wide_steps_replays_from_a_graph(gpu) -> None:
    """run_many_device whose last partial round runs as wide steps on the side stream (forked from and
    joined to the caller's stream) captured into a HIP graph: three replays, each after the outputs were
    overwritten and after other eager work, reproduce the eager run bit for bit."""
    dag = layered_small()
    n = 256 * 128 + 77
    P = capi.Propagator(dag)
    P.configure(16, 8, 2)
    capi.check(capi.lib().mcdp_debug_tail_mode(P.h, 2))
    sd = capi.DeviceBuffer.from_array(np.arange(n, dtype=np.int32) * 7 - 3)
    real, dur, cause = capi.DeviceBuffer(8 * P.E * n), capi.DeviceBuffer(8 * P.L * n), capi.DeviceBuffer(4 * P.E * n)
    stream = capi.C.c_void_p(0)
    capi.check(capi.lib().mcdp_stream_create(capi.C.byref(stream)))

    def read():
        return (real.to_array(np.float64, (P.E, n)), dur.to_array(np.float64, (P.L, n)),
                cause.to_array(np.int32, (P.E, n)))

    P.run_many_device(sd, n, n, real, dur, cause, stream)
    capi.synchronize(stream)
    want = read()
    graph = capi.C.c_void_p(0)
    capi.check(capi.lib().mcdp_graph_capture_begin(stream))
    P.run_many_device(sd, n, n, real, dur, cause, stream)
    capi.check(capi.lib().mcdp_graph_capture_end(stream, capi.C.byref(graph)))
    other = capi.Propagator(layered_network(layers=12))
    try:
        for rep in range(3):
            for b in (real, dur, cause):
                capi.check(capi.lib().mcdp_memset(b.ptr, 0xFF, b.nbytes))
            other.run_many(np.arange(300))
            capi.check(capi.lib().mcdp_graph_launch(graph, stream))
            capi.synchronize(stream)
            got = read()
            np.testing.assert_array_equal(bits(got[0]), bits(want[0]), err_msg=f"replay {rep}")
            np.testing.assert_array_equal(bits(got[1]), bits(want[1]), err_msg=f"replay {rep}")
            np.testing.assert_array_equal(got[2], want[2], err_msg=f"replay {rep}")
    finally:
        capi.check(capi.lib().mcdp_graph_destroy(graph))
        capi.check(capi.lib().mcdp_stream_destroy(stream))


@pytest.mark.parametrize("case", range(4))
def test_wide_steps_special_values(gpu, case: int) -> None:
    """Injected durations with NaN, +-inf, -0.0, ties and caps through a run whose last partial round is
    forced onto wide steps: every sample equals the oracle, the wide samples included."""
    rng = np.random.default_rng(300 + case)
    kw = [{}, {"duplicates": True}, {"gaps": True}, {"oob_edges": True}][case]
    dag = D.random_dag(rng, int(rng.integers(40, 150)), max_preds=int(rng.integers(2, 7)), **kw)
    P = capi.Propagator(dag)
    P.configure(16, 8, 2)
    capi.check(capi.lib().mcdp_debug_tail_mode(P.h, 2))
    ctx = oracle.OracleCtx(dag)
    n = 256 * 128 + 130  # one full round of 256 resident tiles of 128 on a 256-CU device, then 130 samples
    dur = D.special_durations(rng, n, ctx.L)
    r, c = P.propagate(dur)
    pick = np.r_[0:64, n - 300:n]
    rr, cc = ctx.propagate(dur[pick])
    np.testing.assert_array_equal(bits(r[pick]), bits(rr))
    np.testing.assert_array_equal(c[pick], cc)
    Q = capi.Propagator(dag)
    Q.configure(16, 8, 1)
    r1, c1 = Q.propagate(dur)
    np.testing.assert_array_equal(bits(r), bits(r1))
    np.testing.assert_array_equal(c, c1)


@pytest.mark.parametrize("name", sorted(PARTIAL_ROUND_DAGS))
@pytest.mark.parametrize("n", [1, 77, 1000])
def test_short_run_routes_match(gpu, name: str, n: int) -> None:
    """Runs shorter than one round: one partial round of tiles (route 1) or wide steps alone (route 2), and
    the automatic choice, give the same bits, sampled and injected; the oracle agrees."""
    dag = PARTIAL_ROUND_DAGS[name]()
    seeds = np.arange(n, dtype=np.int32) * 11 + 4
    out = []
    for mode in (1, 2, 0):
        P = capi.Propagator(dag)
        P.configure(16, 8, 2)
        capi.check(capi.lib().mcdp_debug_tail_mode(P.h, mode))
        out.append(P.run_many(seeds))
    for r, d, c in out[1:]:
        np.testing.assert_array_equal(bits(r), bits(out[0][0]))
        np.testing.assert_array_equal(bits(d), bits(out[0][1]))
        np.testing.assert_array_equal(c, out[0][2])
    r0, d0, c0 = out[0]
    P = capi.Propagator(dag)
    P.configure(16, 8, 2)
    capi.check(capi.lib().mcdp_debug_tail_mode(P.h, 2))
    rr, cc = P.propagate(d0)
    np.testing.assert_array_equal(bits(rr), bits(r0))
    np.testing.assert_array_equal(cc, c0)
    ro, co = oracle.OracleCtx(dag).propagate(d0[: min(n, 200)])
    np.testing.assert_array_equal(bits(ro), bits(r0[: min(n, 200)]))
    np.testing.assert_array_equal(co, c0[: min(n, 200)])
