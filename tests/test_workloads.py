"""The synthetic workloads are pinned: every generator draws from SplitMix64 (no numpy Generator
stream, which may change between numpy versions), and the sha256 of the generated arrays of each
benchmark workload is committed here, so a bench line and a GPU test always run the same DAGs.

* configuration 0: layered_small, 200 events;
* configuration 1: layered_network, 20,000 events / 50,005 activities;
* configuration 2: batch_workload, 64 DAGs (819,700 events), with their sample counts;
* configuration 3: serve_contexts, 200 contexts (7,785,800 events), and the 300-request Poisson
  stream at 12 requests/s the serving bench submits.
"""

from __future__ import annotations

import hashlib

import numpy as np

from mc_dagprop_amd.workloads import (SplitMix64, SplitMix64Stream, batch_workload, dag_digest, layered_network,
                                      layered_small, layered_vectorized, network_dists, serve_contexts, serve_stream)

DIGESTS = {
    "config0": "86dd127a5196035732f19f77a9308b9516feb0ae3ea4b91efdb4def0e6705d46",
    "config1": "30d0d2b811cf26a8c216631f6c29e2c107544cf98824a46f2bc576fbf6994851",
    "config2": "a11deea87fb542acaf4c6aa29eba8f13f6acc56783034f6c89a16878fc7dd4bc",
    "config3": "b234cfadb05fee839c931543df1d91e62200c0fae61491cd8c732e9e2216af46",
}


def test_splitmix64_known_answers() -> None:
    """The published first outputs of SplitMix64 seeded with 1234567 (Steele, Lea, Flood 2014; the
    seeding generator of the xoshiro family)."""
    r = SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                             4593380528125082431, 16408922859458223821]


def test_block_stream_is_the_sequential_stream() -> None:
    for seed in (0, 1, 20260003, 2**64 - 1):
        a, b = SplitMix64(seed), SplitMix64Stream(seed)
        want = [a.next() for _ in range(1000)]
        got = np.concatenate([b.raw(1), b.raw(998), b.raw(1)])
        assert [int(v) for v in got] == want
        a, b = SplitMix64(seed), SplitMix64Stream(seed)
        assert np.array_equal(b.uniform(50, 30.0, 600.0), [a.uniform(30.0, 600.0) for _ in range(50)])
        assert np.array_equal(b.below(50, 7), [a.below(7) for _ in range(50)])


def test_vectorized_layered_dag_rules() -> None:
    d = layered_vectorized(40, 100, 4, network_dists(), 8, 7001)
    counts = np.diff(d.prec_off)
    assert counts.min() >= 1 and counts.max() <= 4 and d.prec_target.tolist() == list(range(100, 4000))
    for j, e in enumerate(d.prec_target):
        a, b = int(d.prec_off[j]), int(d.prec_off[j + 1])
        src = d.prec_src[a:b]
        assert len(set(src.tolist())) == b - a and np.all(src // 100 == e // 100 - 1)
        assert np.all(d.act_dst[d.prec_edge[a:b]] == e) and np.all(d.act_src[d.prec_edge[a:b]] == src)
        slack = d.earliest[e] - np.max(d.earliest[src] + d.act_dur[a:b])
        assert 0.0 <= slack < 120.0
    assert np.all((d.act_dur >= 30.0) & (d.act_dur < 600.0)) and set(np.unique(d.act_type)) == set(range(8))
    assert np.all(d.earliest[:100] < 120.0)


def test_benchmark_workload_digests() -> None:
    assert dag_digest(layered_small()) == DIGESTS["config0"]
    net = layered_network()
    assert (net.n_events, net.link_count) == (20_000, 50_005)
    assert dag_digest(net) == DIGESTS["config1"]
    dags, ns = batch_workload()
    h = hashlib.sha256()
    for d in dags:
        h.update(dag_digest(d).encode())
    h.update(np.asarray(ns, dtype="<i8").tobytes())
    assert sum(d.n_events for d in dags) == 819_700 and min(ns) >= 4096 and max(ns) <= 65536
    assert h.hexdigest() == DIGESTS["config2"]


def test_serving_workload_digest() -> None:
    ctx = serve_contexts()
    h = hashlib.sha256()
    for d in ctx:
        h.update(dag_digest(d).encode())
    st = serve_stream(200, 300, 12.0)
    for a in (st.arrivals.astype("<f8"), st.context.astype("<i8"), st.samples.astype("<i8")):
        h.update(a.tobytes())
    sizes = [d.n_events for d in ctx]
    assert sum(sizes) == 7_785_800 and min(sizes) >= 1000 and max(sizes) <= 200_000
    assert st.samples.min() >= 16384 and st.samples.max() <= 262144
    assert h.hexdigest() == DIGESTS["config3"]
