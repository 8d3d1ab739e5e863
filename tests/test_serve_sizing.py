"""The serving footprint rule (mcdp_server_submit's chunk sizing), host only.

A request whose whole footprint exceeds the request cap runs in chunks of samples.  The chunk
is the largest multiple of 128 samples whose footprint fits the cap; a request that cannot fit
even 128 samples is refused.  Near the boundary (the fixed part of the footprint within 1,024
bytes of the cap) an unsigned subtraction used to wrap, giving a chunk of every sample and a
footprint above the cap, which the strictly ordered admission could never place.

The footprint below restates the device layout (256-byte aligned blocks: seeds, realized and
cause rows (one byte per cause, the winning entry's position, when no event has more than 254
entries; four otherwise), mean, quantiles, cause counts, and for chunked requests the running
statistics of
56 bytes per event and a 4,096-bin histogram per event, and with quantiles the exact selection's
targets, candidate values and refine slots, and the rows kept whole: one per 128 events, at least
16 and at most E, within a quarter of the cap, each with its n samples, its quantiles and its event
id; they give way to none when the chunks would not fit beside them); the test checks the
library's answers against it: every accepted request fits the cap, and no larger chunk would.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from mc_dagprop_amd import capi

CHUNK_STAT_BYTES = 56
CHUNK_BINS = 4096
MAX_RANKS = 16           # two order statistics per quantile, at most 8 quantiles
EXACT_TARGET_BYTES = 80  # one (event, rank) target of the exact selection
# candidate values per target: n / 512 (16..256); refine slots of 256 sub-bins: one per 64 targets (16..1024)


def _a(x: int) -> int:
    return (x + 255) // 256 * 256


def footprint(n: int, rows: int, E: int, P: int, nq: int, chunked: bool, kept: int = 0, cb: int = 4) -> int:
    b = _a(n * 4) + _a(E * rows * 8) + _a(E * rows * cb) + _a(E * 8) + _a(E * nq * 8) + _a((E + P) * 8)
    if chunked:
        b += _a(E * CHUNK_STAT_BYTES) + _a(E * CHUNK_BINS * 4)
    if chunked and nq > 0:  # exact quantiles: targets, candidates, refine slots' counts and key bounds, counters
        T = E * min(2 * nq, MAX_RANKS)
        per_target = min(256, max(16, (n + 511) // 512))
        slots = min(1024, max(16, (T + 63) // 64))
        b += _a(T * EXACT_TARGET_BYTES) + _a(T * per_target * 8) + _a(slots * 256 * 4)
        b += 2 * _a(slots * 256 * 8) + 256
        if kept:  # rows kept whole: event ids, samples, quantiles
            b += _a(kept * 4) + _a(kept * n * 8) + _a(kept * nq * 8)
    return b


def kept_rows(n: int, E: int, nq: int, cap: int) -> int:
    if nq <= 0:
        return 0
    return min(E, max(16, (E + 127) // 128), cap // 4 // (n * 8))


def chunk_of(n: int, E: int, P: int, nq: int, cap: int, kept: int, cb: int) -> int:
    """The largest multiple of 128 samples (at most n) whose footprint, with 1,024 bytes of slack, fits."""
    fixed = footprint(n, 0, E, P, nq, True, kept, cb)
    if fixed + 1024 >= cap:
        return 0
    m = (cap - fixed - 1024) // ((8 + cb) * max(E, 1)) // 128 * 128
    m = min(m, n)
    return m if m >= 128 and footprint(n, m, E, P, nq, True, kept, cb) <= cap else 0


def plan(n: int, E: int, P: int, nq: int, cap: int, cb: int) -> tuple[int, int]:
    kept = kept_rows(n, E, nq, cap)
    m = chunk_of(n, E, P, nq, cap, kept, cb)
    if m < 1 and kept > 0:
        kept, m = 0, chunk_of(n, E, P, nq, cap, 0, cb)
    return m, kept


def sizing(n: int, E: int, P: int, nq: int, cap: int, max_entries: int = 300) -> tuple[int, int, int]:
    """The library's answer; max_entries above 254 keeps four-byte causes in chunks."""
    chunk, nbytes = C.c_int64(-1), C.c_int64(-1)
    rc = capi.lib().mcdp_debug_serve_chunk(n, E, P, max_entries, nq, cap, C.byref(chunk), C.byref(nbytes))
    return rc, chunk.value, nbytes.value


def test_whole_request_fits() -> None:
    rc, chunk, nbytes = sizing(1000, 50, 80, 3, 1 << 30)
    assert (rc, chunk, nbytes) == (0, 0, footprint(1000, 1000, 50, 80, 3, False))
    rc, chunk, nbytes = sizing(1000, 50, 80, 3, 1 << 30, max_entries=4)  # one-byte winner positions
    assert (rc, chunk, nbytes) == (0, 0, footprint(1000, 1000, 50, 80, 3, False, 0, 1))


@pytest.mark.parametrize("slack", [0, 1, 255, 256, 1000, 1023, 1024, 1025])
def test_cap_just_above_the_fixed_part_is_refused(slack: int) -> None:
    """fixed + slack with slack <= 1024 + a few bytes: not even 128 samples fit, so NOMEM."""
    n, E, P, nq = 262_144, 200_000, 450_000, 3
    fixed = footprint(n, 0, E, P, nq, True)
    rc, chunk, nbytes = sizing(n, E, P, nq, fixed + slack)
    assert rc == 6, (rc, chunk, nbytes)
    assert b"128 samples" in capi.lib().mcdp_last_error()


def test_smallest_chunk_at_the_boundary() -> None:
    n, E, P, nq = 262_144, 200_000, 450_000, 3
    fixed = footprint(n, 0, E, P, nq, True)
    need128 = footprint(n, 128, E, P, nq, True)
    rc, chunk, nbytes = sizing(n, E, P, nq, need128 + 1024)
    assert rc == 0 and chunk == 128 and nbytes == need128 and nbytes <= need128 + 1024
    # one byte below what the rule's slack asks: refused, never an oversized chunk
    rc, chunk, _ = sizing(n, E, P, nq, fixed + 1024 + 128 * 12 * E - 1)
    assert rc == 6


def test_random_requests_fit_their_cap_and_chunks_are_maximal() -> None:
    rng = np.random.default_rng(7)
    with_kept = with_bytes = 0
    for _ in range(3000):
        E = int(rng.integers(1, 250_000))
        P = int(rng.integers(0, 3 * E + 1))
        nq = int(rng.integers(0, 9))
        n = int(rng.integers(1, 300_000))
        entries = int(rng.choice([1, 4, 254, 255, 1000]))
        cb = 1 if entries <= 254 else 4
        whole = footprint(n, n, E, P, nq, False, 0, cb)
        cap = int(rng.integers(1, max(2, 2 * whole)))
        rc, chunk, nbytes = sizing(n, E, P, nq, cap, entries)
        if whole <= cap:
            assert (rc, chunk, nbytes) == (0, 0, whole)
            continue
        if rc != 0:
            assert rc == 6 and plan(n, E, P, nq, cap, cb)[0] == 0
            assert footprint(n, min(128, n), E, P, nq, True, 0, cb) + 1024 > cap
            continue
        assert 1 <= chunk <= n and (chunk % 128 == 0 or chunk == n), chunk
        m, kept = plan(n, E, P, nq, cap, cb)
        assert m == chunk
        with_kept += kept > 0
        with_bytes += cb == 1
        assert nbytes == footprint(n, chunk, E, P, nq, True, kept, cb) and nbytes <= cap
        if chunk < n and chunk + 128 <= n:
            assert footprint(n, chunk + 128, E, P, nq, True, kept, cb) + 1024 > cap
    assert with_kept > 100  # the kept rows' share of the footprint is exercised, not only its fallback
    assert with_bytes > 100  # and so are chunks of one-byte causes


def test_byte_causes_hold_more_samples_per_chunk() -> None:
    """The network-scale serving request: 9 bytes per event and sample instead of 12."""
    n, E, P, nq, cap = 262_144, 200_000, 500_000, 3, 50 << 30
    _, m4, _ = sizing(n, E, P, nq, cap, 300)
    _, m1, b1 = sizing(n, E, P, nq, cap, 4)
    assert m1 > m4 * 1.25 and b1 <= cap


def test_invalid_arguments() -> None:
    assert sizing(0, 10, 10, 0, 1 << 20)[0] == 1
    assert sizing(10, 10, 10, 9, 1 << 20)[0] == 1
