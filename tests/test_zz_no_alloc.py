"""mcdp_reverse_device_noalloc: the reverse walk on the library's fixed scratch array, called
through the C ABI, gives the bits mcdp_reverse_device gives, on two calls in a row (the second
with fewer samples, so the scratch holds the first call's contents), with more samples than one
launch of the scratch holds (several launches), with and without accumulation, and it refuses a
propagator whose plans are not on the device yet."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import layered_network, layered_small

pytestmark = pytest.mark.gpu


def _bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint64)


class _Run:
    """run_many outputs of n samples on the device, and gradient output buffers."""

    def __init__(self, prop: capi.Propagator, n: int, seed0: int) -> None:
        self.n = n
        s = np.arange(seed0, seed0 + n, dtype=np.int32)
        self.seeds = capi.DeviceBuffer.from_array(s)
        self.real = capi.DeviceBuffer(8 * prop.E * n)
        self.dur = capi.DeviceBuffer(8 * prop.L * n)
        self.cause = capi.DeviceBuffer(4 * prop.E * n)
        prop.run_many_device(self.seeds, n, n, self.real, self.dur, self.cause)
        capi.synchronize()


def _reverse(fn, prop, run: _Run, w, accumulate: int, init=None):
    outs = [capi.DeviceBuffer(8), capi.DeviceBuffer(8 * prop.L), capi.DeviceBuffer(8 * prop.E)]
    if init is not None:
        for b, a in zip(outs, init):
            capi.check(capi.lib().mcdp_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
    rc = fn(prop.h, run.seeds.ptr, w.ptr, run.real.ptr, run.dur.ptr, run.cause.ptr, run.n, run.n, accumulate,
            outs[0].ptr, outs[1].ptr, outs[2].ptr, None)
    capi.check(rc)
    capi.synchronize()
    return (outs[0].to_array(np.float64, (1,)), outs[1].to_array(np.float64, (prop.L,)),
            outs[2].to_array(np.float64, (prop.E,)))


def _same(a, b) -> None:
    for x, y in zip(a, b):
        np.testing.assert_array_equal(_bits(x), _bits(y))


@pytest.mark.parametrize("which", ["network", "small"])
def test_noalloc_matches_reverse_device_twice(gpu, which: str) -> None:
    lib = capi.lib()
    dag = layered_network() if which == "network" else layered_small()
    prop = capi.Propagator(dag)
    rng = np.random.default_rng(5)
    w = capi.DeviceBuffer.from_array(rng.uniform(0.5, 2.0, prop.E))
    # network: 65,536 samples need 512 tiles of 70,021 partial-row columns; the 256 MiB scratch holds
    # 448 per launch, so the call takes two launches
    sizes = (65_536, 20_000) if which == "network" else (9_000, 700)
    print(f"noalloc {which}: E={prop.E} L={prop.L} sizes={sizes} {prop.gradient_info()}", flush=True)
    first = None
    for k, n in enumerate(sizes):
        run = _Run(prop, n, 1000 * k)
        want = _reverse(lib.mcdp_reverse_device, prop, run, w, 0)
        got = _reverse(lib.mcdp_reverse_device_noalloc, prop, run, w, 0)
        _same(got, want)
        # accumulating onto earlier sums
        init = [np.full(1, 3.0), np.linspace(-1.0, 1.0, prop.L), np.linspace(2.0, -2.0, prop.E)]
        want_acc = _reverse(lib.mcdp_reverse_device, prop, run, w, 1, init)
        got_acc = _reverse(lib.mcdp_reverse_device_noalloc, prop, run, w, 1, init)
        _same(got_acc, want_acc)
        if first is None:
            first = (run, want)
    # the first call's inputs again, after the smaller call left the scratch behind
    run, want = first
    _same(_reverse(lib.mcdp_reverse_device_noalloc, prop, run, w, 0), want)


def test_noalloc_refuses_plans_not_on_the_device(gpu) -> None:
    lib = capi.lib()
    prop = capi.Propagator(layered_small())
    n = 128
    seeds = capi.DeviceBuffer.from_array(np.arange(n, dtype=np.int32))
    w = capi.DeviceBuffer.from_array(np.ones(prop.E))
    bufs = [capi.DeviceBuffer(8 * prop.E * n), capi.DeviceBuffer(8 * prop.L * n), capi.DeviceBuffer(4 * prop.E * n),
            capi.DeviceBuffer(8), capi.DeviceBuffer(8 * prop.L), capi.DeviceBuffer(8 * prop.E)]
    r, d, c, o, gd, ge = (b.ptr for b in bufs)
    rc = lib.mcdp_reverse_device_noalloc(prop.h, seeds.ptr, w.ptr, r, d, c, n, n, 0, o, gd, ge, None)
    assert rc == 1 and b"not resident" in lib.mcdp_last_error()
    # one call of the allocating entry point with n = 0 makes them resident
    capi.check(lib.mcdp_reverse_device(prop.h, None, None, None, None, None, 0, 0, 0, o, gd, ge, None))
    prop.run_many_device(seeds, n, n, bufs[0], bufs[1], bufs[2])
    capi.check(lib.mcdp_reverse_device_noalloc(prop.h, seeds.ptr, w.ptr, r, d, c, n, n, 0, o, gd, ge, None))
    capi.synchronize()
