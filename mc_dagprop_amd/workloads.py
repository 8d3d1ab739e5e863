"""Flat DAG descriptions and the synthetic benchmark workloads.

A :class:`FlatDag` holds a DAG context and a delay generator as plain numpy
arrays: the form the C ABI (``include/mcdp.h``) takes.  It can also build the
Python API objects (``DagContext``, ``GenericDelayGenerator``).

The layered generators reproduce the benchmark workloads of the path:

* ``layered_small``   -- 20 layers x 10 events, 1-3 predecessors from the previous
  layer (~380 activities), 3 activity types (constant 1.5, exponential mean 60,
  gamma(2, 30));
* ``layered_network`` -- 200 layers x 100 events, 1-4 predecessors (~50,000
  activities), 8 activity types (constant 1.5, exponential means 30 and 120,
  gamma (2, 30) and (0.5, 200), 64-bin empirical absolute and relative tables on
  [0, 900] s with Dirichlet(1) weights, one type without a distribution).

Timestamps: minimal_duration ~ U[30, 600) s; earliest = critical-path time +
U[0, 120) s slack; latest = earliest + 3600; actual = 0; max_delay = 1800.
All randomness comes from one SplitMix64 stream (seed 20260000 + variant) so
the workloads are identical on every machine and numpy version.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Distribution kinds, in the order of the reference's DistVar alternatives
# (_core.cpp:143).
CONSTANT, EXPONENTIAL, GAMMA, EMPIRICAL_ABSOLUTE, EMPIRICAL_RELATIVE = range(5)


@dataclass
class DistSpec:
    activity_type: int
    kind: int
    params: tuple[float, float, float] = (0.0, 0.0, float("inf"))
    values: np.ndarray | None = None
    weights: np.ndarray | None = None


@dataclass
class FlatDag:
    earliest: np.ndarray
    latest: np.ndarray
    actual: np.ndarray
    act_src: np.ndarray
    act_dst: np.ndarray
    act_idx: np.ndarray
    act_dur: np.ndarray
    act_type: np.ndarray
    prec_target: np.ndarray
    prec_off: np.ndarray
    prec_src: np.ndarray
    prec_edge: np.ndarray
    max_delay: float
    dists: list[DistSpec] = field(default_factory=list)

    @property
    def n_events(self) -> int:
        return int(self.earliest.shape[0])

    @property
    def link_count(self) -> int:
        return int(self.act_idx.max()) + 1 if self.act_idx.size else 0

    # -- builders ---------------------------------------------------------
    @staticmethod
    def from_lists(events, activities, precedence_list, max_delay, dists=()) -> "FlatDag":
        """events: [(earliest, latest, actual)]; activities: {(src, dst): (idx, dur, type)};
        precedence_list: [(target, [(src, edge), ...])]."""
        ev = np.asarray(events, dtype=np.float64).reshape(-1, 3)
        keys = list(activities.keys())
        vals = [activities[k] for k in keys]
        off = [0]
        src, edge = [], []
        for _, preds in precedence_list:
            for s, e in preds:
                src.append(s)
                edge.append(e)
            off.append(len(src))
        return FlatDag(
            earliest=ev[:, 0].copy(),
            latest=ev[:, 1].copy(),
            actual=ev[:, 2].copy(),
            act_src=np.asarray([k[0] for k in keys], dtype=np.int32),
            act_dst=np.asarray([k[1] for k in keys], dtype=np.int32),
            act_idx=np.asarray([v[0] for v in vals], dtype=np.int32),
            act_dur=np.asarray([v[1] for v in vals], dtype=np.float64),
            act_type=np.asarray([v[2] for v in vals], dtype=np.int32),
            prec_target=np.asarray([t for t, _ in precedence_list], dtype=np.int32),
            prec_off=np.asarray(off, dtype=np.int64),
            prec_src=np.asarray(src, dtype=np.int32),
            prec_edge=np.asarray(edge, dtype=np.int32),
            max_delay=float(max_delay),
            dists=list(dists),
        )

    def to_api(self, api):
        """(DagContext, GenericDelayGenerator) built through module ``api``."""
        events = [
            api.Event(str(i), api.EventTimestamp(float(a), float(b), float(c)))
            for i, (a, b, c) in enumerate(zip(self.earliest, self.latest, self.actual))
        ]
        acts = {
            (int(s), int(d)): api.Activity(idx=int(i), minimal_duration=float(u), activity_type=int(t))
            for s, d, i, u, t in zip(self.act_src, self.act_dst, self.act_idx, self.act_dur, self.act_type)
        }
        prec = []
        for i, tgt in enumerate(self.prec_target):
            a, b = int(self.prec_off[i]), int(self.prec_off[i + 1])
            prec.append((int(tgt), [(int(s), int(e)) for s, e in zip(self.prec_src[a:b], self.prec_edge[a:b])]))
        ctx = api.DagContext(events=events, activities=acts, precedence_list=prec, max_delay=self.max_delay)
        return ctx, self.generator(api)

    def generator(self, api):
        gen = api.GenericDelayGenerator()
        for d in self.dists:
            a, b, c = d.params
            if d.kind == CONSTANT:
                gen.add_constant(d.activity_type, a)
            elif d.kind == EXPONENTIAL:
                gen.add_exponential(d.activity_type, a, b)
            elif d.kind == GAMMA:
                gen.add_gamma(d.activity_type, a, b, c)
            elif d.kind == EMPIRICAL_ABSOLUTE:
                gen.add_empirical_absolute(d.activity_type, list(map(float, d.values)), list(map(float, d.weights)))
            else:
                gen.add_empirical_relative(d.activity_type, list(map(float, d.values)), list(map(float, d.weights)))
        return gen


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the workload generators' only source of randomness."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int) -> None:
        self.s = seed & self.MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0**-53)

    def below(self, n: int) -> int:
        return min(int(((self.next() >> 11) * 2.0**-53) * n), n - 1)


def _histogram_tables(rng: SplitMix64) -> tuple[np.ndarray, np.ndarray]:
    centres = 7.03125 + 14.0625 * np.arange(64)
    w = np.array([-np.log1p(-rng.uniform()) for _ in range(64)])  # Gamma(1) = Exp(1) => Dirichlet(1)
    return centres, w / w.sum()


def layered(layers: int, width: int, max_preds: int, n_types: int, variant: int, dists: list[DistSpec]) -> FlatDag:
    rng = SplitMix64(20260000 + variant)
    E = layers * width
    earliest = np.zeros(E)
    act_src, act_dst, act_dur, act_type = [], [], [], []
    prec_target, prec_off, prec_src, prec_edge = [], [0], [], []
    for j in range(width):
        earliest[j] = rng.uniform(0.0, 120.0)
    for layer in range(1, layers):
        for j in range(width):
            e = layer * width + j
            k = 1 + rng.below(max_preds)
            pool = list(range((layer - 1) * width, layer * width))
            preds = []
            for i in range(k):  # partial Fisher-Yates, draw order kept
                r = i + rng.below(width - i)
                pool[i], pool[r] = pool[r], pool[i]
                preds.append(pool[i])
            arrive = -np.inf
            for p in preds:
                idx = len(act_src)
                dmin = rng.uniform(30.0, 600.0)
                t = rng.below(n_types)
                act_src.append(p)
                act_dst.append(e)
                act_dur.append(dmin)
                act_type.append(t)
                prec_src.append(p)
                prec_edge.append(idx)
                arrive = max(arrive, earliest[p] + dmin)
            earliest[e] = arrive + rng.uniform(0.0, 120.0)
            prec_target.append(e)
            prec_off.append(len(prec_src))
    A = len(act_src)
    return FlatDag(
        earliest=earliest,
        latest=earliest + 3600.0,
        actual=np.zeros(E),
        act_src=np.asarray(act_src, dtype=np.int32),
        act_dst=np.asarray(act_dst, dtype=np.int32),
        act_idx=np.arange(A, dtype=np.int32),
        act_dur=np.asarray(act_dur),
        act_type=np.asarray(act_type, dtype=np.int32),
        prec_target=np.asarray(prec_target, dtype=np.int32),
        prec_off=np.asarray(prec_off, dtype=np.int64),
        prec_src=np.asarray(prec_src, dtype=np.int32),
        prec_edge=np.asarray(prec_edge, dtype=np.int32),
        max_delay=1800.0,
        dists=dists,
    )


INF = float("inf")


class SplitMix64Stream:
    """The SplitMix64 sequence of :class:`SplitMix64` drawn in blocks: draw i (0-based) of the stream
    seeded with ``seed`` is mix(seed + (i + 1) * 0x9E3779B97F4A7C15 mod 2^64), so a block is one
    vectorized evaluation (numpy uint64 arithmetic wraps modulo 2^64 on every platform).  Uniforms
    and bounded integers are derived exactly as :class:`SplitMix64` derives them."""

    GAMMA = np.uint64(0x9E3779B97F4A7C15)

    def __init__(self, seed: int) -> None:
        self.seed = np.uint64(seed & SplitMix64.MASK)
        self.pos = 0

    def raw(self, n: int) -> np.ndarray:
        i = np.arange(self.pos + 1, self.pos + 1 + n, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            z = self.seed + i * self.GAMMA
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.raw(n) >> np.uint64(11)).astype(np.float64) * 2.0**-53
        return lo + (hi - lo) * u

    def below(self, n: int, m) -> np.ndarray:
        """n integers, the i-th uniform on [0, m_i) (m an int or an array of n ints)."""
        m = np.asarray(m, dtype=np.int64)
        u = (self.raw(n) >> np.uint64(11)).astype(np.float64) * 2.0**-53
        return np.minimum((u * m).astype(np.int64), m - 1)


def layered_vectorized(layers: int, width: int, max_preds: int, dists: list[DistSpec], n_types: int,
                       seed: int) -> FlatDag:
    """The layered construction of :func:`layered` (1..max_preds distinct predecessors from the previous
    layer, kept in draw order; minimal durations U[30, 600) s; earliest = latest predecessor arrival +
    U[0, 120) s slack; max_delay 1800), drawn from one :class:`SplitMix64Stream` in blocks over the
    whole DAG, fast enough for hundreds of contexts of up to 200,000 events.  Draw order: the
    predecessor count of every event of layers 1.. (event order); max_preds rounds of a partial
    Fisher-Yates shuffle of each such event's copy of the previous layer (one draw per event and
    round, consumed whatever the event's count); the minimal duration of every activity, then its
    type (activity order: by destination event, predecessors in draw order); the slack of every
    event (layer 0 first).  Identical on every machine and numpy version."""
    rng = SplitMix64Stream(seed)
    E = layers * width
    N = (layers - 1) * width  # events with predecessors
    k = 1 + rng.below(N, max_preds)
    pool = np.tile(np.arange(width, dtype=np.int16 if width < 32768 else np.int64), (N, 1))
    rows = np.arange(N)
    for i in range(min(max_preds, width)):
        r = i + rng.below(N, width - i)
        a, b = pool[rows, i].copy(), pool[rows, r].copy()
        pool[rows, i], pool[rows, r] = b, a
    take = np.arange(width)[None, :] < k[:, None]
    dst = width + np.repeat(rows, k)  # destination event of each activity, in activity order
    src = (dst // width - 1) * width + pool[take].astype(np.int64)
    A = src.size
    dmin = rng.uniform(A, 30.0, 600.0)
    typ = rng.below(A, n_types)
    slack = rng.uniform(E, 0.0, 120.0)
    earliest = np.zeros(E)
    earliest[:width] = slack[:width]
    seg = np.concatenate([[0], np.cumsum(k)])  # activities of event width + j: seg[j] .. seg[j + 1]
    for layer in range(1, layers):
        j0, j1 = (layer - 1) * width, layer * width
        a0, a1 = seg[j0], seg[j1]
        cand = earliest[src[a0:a1]] + dmin[a0:a1]
        arrive = np.maximum.reduceat(cand, seg[j0:j1] - a0)
        earliest[layer * width:(layer + 1) * width] = arrive + slack[layer * width:(layer + 1) * width]
    return FlatDag(
        earliest=earliest,
        latest=earliest + 3600.0,
        actual=np.zeros(E),
        act_src=src.astype(np.int32),
        act_dst=dst.astype(np.int32),
        act_idx=np.arange(A, dtype=np.int32),
        act_dur=dmin,
        act_type=typ.astype(np.int32),
        prec_target=np.arange(width, E, dtype=np.int32),
        prec_off=seg.astype(np.int64),
        prec_src=src.astype(np.int32),
        prec_edge=np.arange(A, dtype=np.int32),
        max_delay=1800.0,
        dists=dists,
    )


def dag_digest(dag: FlatDag) -> str:
    """sha256 of a DAG's arrays and distribution parameters (fixed dtypes, little-endian)."""
    import hashlib

    h = hashlib.sha256()
    for a, t in ((dag.earliest, "<f8"), (dag.latest, "<f8"), (dag.actual, "<f8"), (dag.act_src, "<i4"),
                 (dag.act_dst, "<i4"), (dag.act_idx, "<i4"), (dag.act_dur, "<f8"), (dag.act_type, "<i4"),
                 (dag.prec_target, "<i4"), (dag.prec_off, "<i8"), (dag.prec_src, "<i4"), (dag.prec_edge, "<i4")):
        h.update(np.ascontiguousarray(a, dtype=t).tobytes())
    h.update(np.float64(dag.max_delay).astype("<f8").tobytes())
    for d in dag.dists:
        h.update(np.asarray([d.activity_type, d.kind], dtype="<i8").tobytes())
        h.update(np.asarray(d.params, dtype="<f8").tobytes())
        for a in (d.values, d.weights):
            if a is not None:
                h.update(np.ascontiguousarray(a, dtype="<f8").tobytes())
    return h.hexdigest()


def batch_workload(n_dags: int = 64) -> tuple[list[FlatDag], list[int]]:
    """The batched workload: ``n_dags`` layered DAGs (width 100, 1-4 predecessors, 8 activity types
    with per-DAG parameters from :func:`seeded_network_dists`), event counts log-uniform in
    [1,000, 50,000] (rounded to whole layers) and sample counts log-uniform in [4,096, 65,536], all
    drawn from SplitMix64 (seed 20260002)."""
    rng = SplitMix64(20260002)
    dags, ns = [], []
    for i in range(n_dags):
        events = int(round(1000.0 * 50.0 ** rng.uniform(0.0, 1.0), -2))
        dags.append(layered_vectorized(events // 100, 100, 4, seeded_network_dists(i), 8, 20260002 + i))
        ns.append(int(4096.0 * 16.0 ** rng.uniform(0.0, 1.0)))
    return dags, ns


@dataclass
class ServeStream:
    """The serving workload's request stream: arrival times (s), context of each request, and its
    sample count; request i's seeds are 1,000,000 i + 0 .. n_i - 1."""

    arrivals: np.ndarray
    context: np.ndarray
    samples: np.ndarray

    def seeds(self, i: int) -> np.ndarray:
        return np.arange(1_000_000 * i, 1_000_000 * i + int(self.samples[i]), dtype=np.int32)


def serve_contexts(n_contexts: int = 200) -> list[FlatDag]:
    """The serving workload's resident contexts: event counts log-uniform in [1,000, 200,000] (whole
    layers of width 100, at least 10 layers), 1-4 predecessors, the network workload's eight
    activity types; SplitMix64 (seed 20260003 for the sizes, 7000 + i for context i)."""
    rng = SplitMix64Stream(20260003)
    events = np.exp(rng.uniform(n_contexts, np.log(1000.0), np.log(200000.0)))
    layers = np.maximum(10, np.round(events / 100.0)).astype(int)
    dists = network_dists()
    return [layered_vectorized(int(m), 100, 4, dists, 8, 7000 + i) for i, m in enumerate(layers)]


def serve_stream(n_contexts: int, n_requests: int, rate: float, min_samples: int = 16384,
                 max_samples: int = 262144) -> ServeStream:
    """A Poisson stream of ``n_requests`` requests at ``rate`` per second, each on a context drawn
    uniformly and asking for a log-uniform number of samples in [min_samples, max_samples]; SplitMix64
    (seed 20260013)."""
    rng = SplitMix64Stream(20260013)
    gaps = -np.log1p(-rng.uniform(n_requests)) / rate
    ctx = rng.below(n_requests, n_contexts)
    n = np.exp(rng.uniform(n_requests, np.log(float(min_samples)), np.log(float(max_samples)))).astype(np.int64)
    return ServeStream(np.cumsum(gaps), ctx, n)


def network_dists() -> list[DistSpec]:
    """The eight activity types of the network workload (type 7 has no distribution)."""
    rng = SplitMix64(20260001 + 1000)
    v_abs, w_abs = _histogram_tables(rng)
    v_rel, w_rel = _histogram_tables(rng)
    return [
        DistSpec(0, CONSTANT, (1.5, 0.0, INF)),
        DistSpec(1, EXPONENTIAL, (30.0, INF, INF)),
        DistSpec(2, EXPONENTIAL, (120.0, INF, INF)),
        DistSpec(3, GAMMA, (2.0, 30.0, INF)),
        DistSpec(4, GAMMA, (0.5, 200.0, INF)),
        DistSpec(5, EMPIRICAL_ABSOLUTE, values=v_abs, weights=w_abs),
        DistSpec(6, EMPIRICAL_RELATIVE, values=v_rel, weights=w_rel),
    ]


def seeded_network_dists(seed: int) -> list[DistSpec]:
    """The network workload's eight activity types with parameters drawn from `seed` (one DAG of a batch):
    constant factor U[1, 2]; exponential scales U[15, 60] and U[60, 240]; gamma (2, U[15, 60]) and
    (0.5, U[100, 400]); empirical absolute and relative 64-bin histograms on [0, 900] with Dirichlet(1)
    weights; type 7 has no distribution."""
    rng = SplitMix64(20260002 + 7919 * seed)
    c = 1.0 + rng.uniform()
    e1, e2 = 15.0 + 45.0 * rng.uniform(), 60.0 + 180.0 * rng.uniform()
    g1, g2 = 15.0 + 45.0 * rng.uniform(), 100.0 + 300.0 * rng.uniform()
    v_abs, w_abs = _histogram_tables(rng)
    v_rel, w_rel = _histogram_tables(rng)
    return [
        DistSpec(0, CONSTANT, (c, 0.0, INF)),
        DistSpec(1, EXPONENTIAL, (e1, INF, INF)),
        DistSpec(2, EXPONENTIAL, (e2, INF, INF)),
        DistSpec(3, GAMMA, (2.0, g1, INF)),
        DistSpec(4, GAMMA, (0.5, g2, INF)),
        DistSpec(5, EMPIRICAL_ABSOLUTE, values=v_abs, weights=w_abs),
        DistSpec(6, EMPIRICAL_RELATIVE, values=v_rel, weights=w_rel),
    ]


def layered_small() -> FlatDag:
    """200 events / ~380 activities, 3 activity types (the CPU-sized workload)."""
    dists = [
        DistSpec(0, CONSTANT, (1.5, 0.0, INF)),
        DistSpec(1, EXPONENTIAL, (60.0, INF, INF)),
        DistSpec(2, GAMMA, (2.0, 30.0, INF)),
    ]
    return layered(20, 10, 3, 3, 0, dists)


def layered_network(layers: int = 200, width: int = 100) -> FlatDag:
    """20,000 events / ~50,000 activities, 8 activity types (the network-scale workload)."""
    rng = SplitMix64(20260001 + 1000)
    v_abs, w_abs = _histogram_tables(rng)
    v_rel, w_rel = _histogram_tables(rng)
    dists = [
        DistSpec(0, CONSTANT, (1.5, 0.0, INF)),
        DistSpec(1, EXPONENTIAL, (30.0, INF, INF)),
        DistSpec(2, EXPONENTIAL, (120.0, INF, INF)),
        DistSpec(3, GAMMA, (2.0, 30.0, INF)),
        DistSpec(4, GAMMA, (0.5, 200.0, INF)),
        DistSpec(5, EMPIRICAL_ABSOLUTE, values=v_abs, weights=w_abs),
        DistSpec(6, EMPIRICAL_RELATIVE, values=v_rel, weights=w_rel),
    ]  # type 7 has no distribution: duration = minimal_duration
    return layered(layers, width, 4, 8, 1, dists)
