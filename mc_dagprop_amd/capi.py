"""ctypes view of libmcdp's C ABI (``include/mcdp.h``).

The pybind11 module ``mc_dagprop_amd.monte_carlo._core`` is the drop-in API;
this module is the thin, torch-free route to the same library for callers that
keep results on the device (the benchmark, statistics consumers) and for tests
that check the C ABI itself.
"""

from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

import numpy as np

from .workloads import CONSTANT, EMPIRICAL_ABSOLUTE, EXPONENTIAL, GAMMA, FlatDag

LIB_PATH = Path(__file__).resolve().parent / "libmcdp.so"
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "mcdp.h"

_P = C.c_void_p
_i32, _i64, _f64, _u32 = C.c_int32, C.c_int64, C.c_double, C.c_uint32
_pi32, _pi64, _pf64, _pu32 = C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_f64), C.POINTER(_u32)

_SIGNATURES = {
    "mcdp_last_error": (C.c_char_p, []),
    "mcdp_abi_version": (C.c_int, []),
    "mcdp_generator_create": (C.c_int, [C.POINTER(_P)]),
    "mcdp_generator_destroy": (None, [_P]),
    "mcdp_generator_set_seed": (C.c_int, [_P, _i32]),
    "mcdp_generator_add_constant": (C.c_int, [_P, _i32, _f64]),
    "mcdp_generator_add_exponential": (C.c_int, [_P, _i32, _f64, _f64]),
    "mcdp_generator_add_gamma": (C.c_int, [_P, _i32, _f64, _f64, _f64]),
    "mcdp_generator_add_empirical_absolute": (C.c_int, [_P, _i32, _pf64, _pf64, _i64]),
    "mcdp_generator_add_empirical_relative": (C.c_int, [_P, _i32, _pf64, _pf64, _i64]),
    "mcdp_propagator_create": (
        C.c_int,
        [_i32, _pf64, _i32, _pi32, _pf64, _pi32, _i32, _pi32, _pi64, _pi32, _pi32, _f64, _P, C.POINTER(_P)],
    ),
    "mcdp_propagator_destroy": (None, [_P]),
    "mcdp_propagator_node_count": (_i32, [_P]),
    "mcdp_propagator_activity_count": (_i32, [_P]),
    "mcdp_run_many_device": (C.c_int, [_P, _P, _i64, _i64, _P, _P, _P, _P]),
    "mcdp_run_many_host": (C.c_int, [_P, _pi32, _i64, _pf64, _pf64, _pi32]),
    "mcdp_propagate_device": (C.c_int, [_P, _P, _i64, _i64, _P, _P, _P]),
    "mcdp_propagate_host": (C.c_int, [_P, _pf64, _i64, _pf64, _pi32]),
    "mcdp_propagator_plan_info": (C.c_int, [_P, _pi64, C.c_int]),
    "mcdp_propagator_ranges": (C.c_int, [_P, _pf64, _pf64, _pf64, _pf64]),
    "mcdp_propagator_configure": (C.c_int, [_P, _i32, _i32, _i32]),
    "mcdp_propagator_autotune": (
        C.c_int,
        [_P, _i64, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32), C.POINTER(C.c_float)],
    ),
    "mcdp_propagator_fingerprint": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "mcdp_propagator_autotune_cached": (
        C.c_int,
        [_P, _i64, C.c_char_p, C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i32), C.POINTER(C.c_float),
         C.POINTER(_i32)],
    ),
    "mcdp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mcdp_malloc": (C.c_int, [C.POINTER(_P), C.c_size_t]),
    "mcdp_free": (C.c_int, [_P]),
    "mcdp_memcpy_htod": (C.c_int, [_P, _P, C.c_size_t]),
    "mcdp_memcpy_dtoh": (C.c_int, [_P, _P, C.c_size_t]),
    "mcdp_memset": (C.c_int, [_P, C.c_int, C.c_size_t]),
    "mcdp_memcpy2d_dtoh": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, C.c_size_t, C.c_size_t]),
    "mcdp_stream_create": (C.c_int, [C.POINTER(_P)]),
    "mcdp_stream_destroy": (C.c_int, [_P]),
    "mcdp_stream_synchronize": (C.c_int, [_P]),
    "mcdp_event_create": (C.c_int, [C.POINTER(_P)]),
    "mcdp_event_destroy": (C.c_int, [_P]),
    "mcdp_event_record": (C.c_int, [_P, _P]),
    "mcdp_event_elapsed_ms": (C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    "mcdp_debug_philox": (C.c_int, [C.c_int, _i64, _pu32, _pu32, _pu32]),
    "mcdp_debug_copy": (C.c_int, [_P, _P, _i64, _P]),
    "mcdp_debug_math": (C.c_int, [C.c_int, _i64, _pf64, _pf64]),
    "mcdp_propagator_pred_offsets": (C.c_int, [_P, _pi64]),
    "mcdp_event_stats_device": (C.c_int, [_P, _P, _P, _i64, _i64, _i32, _pf64, _P, _P, _P, _P]),
    "mcdp_run_stats_host": (C.c_int, [_P, _pi32, _i64, _i32, _pf64, _pf64, _pf64, _pi64]),
    "mcdp_batch_create": (C.c_int, [C.POINTER(_P), _i32, _i32, _i32, _i32, C.POINTER(_P)]),
    "mcdp_batch_destroy": (None, [_P]),
    "mcdp_batch_run_device": (C.c_int, [_P, C.POINTER(_P), _pi64, _pi64, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P), _P]),
    "mcdp_batch_prepare_device": (C.c_int, [_P, C.POINTER(_P), _pi64, _pi64, C.POINTER(_P), C.POINTER(_P),
                                            C.POINTER(_P)]),
    "mcdp_batch_launch": (C.c_int, [_P, _P]),
    "mcdp_graph_capture_begin": (C.c_int, [_P]),
    "mcdp_graph_capture_end": (C.c_int, [_P, C.POINTER(_P)]),
    "mcdp_graph_launch": (C.c_int, [_P, _P]),
    "mcdp_graph_destroy": (C.c_int, [_P]),
    "mcdp_batch_stats_host": (C.c_int, [_P, C.POINTER(_P), _pi64, C.c_int32, _pf64, _i64, C.POINTER(_P), C.POINTER(_P),
                                        C.POINTER(_P)]),
    "mcdp_propagator_set_times": (C.c_int, [_P, _pf64, _pf64]),
    "mcdp_propagator_get_times": (C.c_int, [_P, _pf64, _pf64]),
    "mcdp_reverse_device": (C.c_int, [_P, _P, _P, _P, _P, _P, _i64, _i64, C.c_int, _P, _P, _P, _P]),
    "mcdp_reverse_device_noalloc": (C.c_int, [_P, _P, _P, _P, _P, _P, _i64, _i64, C.c_int, _P, _P, _P, _P]),
    "mcdp_gradient_host": (C.c_int, [_P, _pi32, _i64, _pf64, _i64, _pf64, _pf64, _pf64]),
    "mcdp_debug_reverse_tiles": (C.c_int, [_P, _i64]),
    "mcdp_debug_tail_mode": (C.c_int, [_P, C.c_int32]),
    "mcdp_debug_serve_chunk": (C.c_int, [_i64, _i64, _i64, _i32, _i32, _i64, _pi64, _pi64]),
    "mcdp_propagator_gradient_info": (C.c_int, [_P, _pi64, C.c_int]),
    "mcdp_server_create": (C.c_int, [_i64, _i64, _i32, _i32, C.POINTER(_P)]),
    "mcdp_server_destroy": (None, [_P]),
    "mcdp_server_add_context": (C.c_int, [_P, _P, C.POINTER(_i32)]),
    "mcdp_server_submit": (C.c_int, [_P, _i32, _pi32, _i64, _i32, _pf64, C.POINTER(_i64)]),
    "mcdp_server_wait": (C.c_int, [_P, _i64, _pf64, _pf64, _pi64, C.POINTER(_f64)]),
    "mcdp_server_stats": (C.c_int, [_P, _pi64, C.c_int]),
    "mcdp_server_set_exact_quantiles": (C.c_int, [_P, _i32]),
    "mcdp_debug_server_exact_limits": (C.c_int, [_P, _i32, _i64, _i32]),
    "mcdp_debug_server_kept_rows": (C.c_int, [_P, _i32]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded libmcdp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python mc_dagprop_amd/build.py`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def header_symbols() -> list[str]:
    """Every function name declared in include/mcdp.h."""
    text = HEADER_PATH.read_text()
    return sorted(set(re.findall(r"\b(mcdp_[a-z0-9_]+)\s*\(", text)))


class McdpError(RuntimeError):
    """A nonzero mcdp_status; the message is mcdp_last_error()."""

    def __init__(self, code: int, message: str) -> None:
        super().__init__(message)
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        raise McdpError(rc, lib().mcdp_last_error().decode())


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().mcdp_device_count(C.byref(n)))
    return n.value


class DeviceBuffer:
    """Device allocation owned by this object (freed on close / garbage collection)."""

    def __init__(self, nbytes: int) -> None:
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p(0)
        if self.nbytes:
            check(lib().mcdp_malloc(C.byref(self.ptr), self.nbytes))

    @staticmethod
    def from_array(a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = DeviceBuffer(a.nbytes)
        if a.nbytes:
            check(lib().mcdp_memcpy_htod(buf.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return buf

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            check(lib().mcdp_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def columns(self, dtype, rows: int, ld: int, c0: int, width: int) -> np.ndarray:
        """Columns [c0, c0 + width) of a [rows][ld] device array, as a (width, rows) host array."""
        item = np.dtype(dtype).itemsize
        out = np.empty((rows, width), dtype=dtype)
        if out.nbytes:
            check(lib().mcdp_memcpy2d_dtoh(out.ctypes.data_as(C.c_void_p), width * item,
                                           C.c_void_p(self.ptr.value + c0 * item), ld * item, width * item, rows))
        return out.T

    def close(self) -> None:
        if self.ptr and self.ptr.value:
            lib().mcdp_free(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


class Generator:
    """GenericDelayGenerator over the C ABI."""

    def __init__(self, dists=()) -> None:
        self.h = C.c_void_p(0)
        check(lib().mcdp_generator_create(C.byref(self.h)))
        for d in dists:
            self.add(d)

    def add(self, d) -> None:
        L = lib()
        a, b, c = d.params
        if d.kind == CONSTANT:
            check(L.mcdp_generator_add_constant(self.h, d.activity_type, a))
        elif d.kind == EXPONENTIAL:
            check(L.mcdp_generator_add_exponential(self.h, d.activity_type, a, b))
        elif d.kind == GAMMA:
            check(L.mcdp_generator_add_gamma(self.h, d.activity_type, a, b, c))
        else:
            v = np.ascontiguousarray(d.values, dtype=np.float64)
            w = np.ascontiguousarray(d.weights, dtype=np.float64)
            fn = (
                L.mcdp_generator_add_empirical_absolute
                if d.kind == EMPIRICAL_ABSOLUTE
                else L.mcdp_generator_add_empirical_relative
            )
            check(fn(self.h, d.activity_type, _ptr(v, _pf64), _ptr(w, _pf64), v.size))

    def __del__(self) -> None:  # pragma: no cover
        try:
            lib().mcdp_generator_destroy(self.h)
        except Exception:
            pass


class Propagator:
    """MonteCarloPropagator over the C ABI, built from a :class:`FlatDag`."""

    def __init__(self, dag: FlatDag) -> None:
        gen = Generator(dag.dists)
        self._keep = [
            np.ascontiguousarray(x, dtype=t)
            for x, t in (
                (dag.earliest, np.float64),
                (dag.act_idx, np.int32),
                (dag.act_dur, np.float64),
                (dag.act_type, np.int32),
                (dag.prec_target, np.int32),
                (dag.prec_off, np.int64),
                (dag.prec_src, np.int32),
                (dag.prec_edge, np.int32),
            )
        ]
        e, ai, ad, at, pt, po, ps, pe = self._keep
        self.h = C.c_void_p(0)
        self.last_tune_from_file = False
        check(
            lib().mcdp_propagator_create(
                e.size,
                _ptr(e, _pf64),
                ai.size,
                _ptr(ai, _pi32),
                _ptr(ad, _pf64),
                _ptr(at, _pi32),
                pt.size,
                _ptr(pt, _pi32),
                _ptr(po, _pi64),
                _ptr(ps, _pi32),
                _ptr(pe, _pi32),
                dag.max_delay,
                gen.h,
                C.byref(self.h),
            )
        )
        self.E = lib().mcdp_propagator_node_count(self.h)
        self.L = lib().mcdp_propagator_activity_count(self.h)

    PLAN_FIELDS = ["steps", "waves", "lds_slots", "lds_bytes", "global_preds", "stale_preds", "orphans",
                   "max_events_per_wave", "events_per_wave", "samples_per_lane", "tail_tile_cost", "tail_wide_cost"]

    def plan_info(self) -> dict:
        v = np.zeros(len(self.PLAN_FIELDS), dtype=np.int64)
        check(lib().mcdp_propagator_plan_info(self.h, _ptr(v, _pi64), v.size))
        return dict(zip(self.PLAN_FIELDS, map(int, v)))

    def ranges(self, events: bool = True) -> tuple:
        """Bounds that hold for every seed: (link_lo [L], link_hi [L], event_lo [E], event_hi [E]); the
        event bounds are None when an earliest time or max_delay is not finite (or events=False)."""
        llo, lhi = np.empty(self.L), np.empty(self.L)
        check(lib().mcdp_propagator_ranges(self.h, _ptr(llo, _pf64), _ptr(lhi, _pf64), None, None))
        if not events:
            return llo, lhi, None, None
        elo, ehi = np.empty(self.E), np.empty(self.E)
        rc = lib().mcdp_propagator_ranges(self.h, None, None, _ptr(elo, _pf64), _ptr(ehi, _pf64))
        if rc:
            return llo, lhi, None, None
        return llo, lhi, elo, ehi

    def configure(self, waves: int, events_per_wave: int, samples_per_lane: int = 1) -> None:
        """Fixes the schedule shape (results do not depend on it)."""
        check(lib().mcdp_propagator_configure(self.h, waves, events_per_wave, samples_per_lane))

    def pred_offsets(self) -> np.ndarray:
        """off[0..E]: start of each event's predecessor list; cause counters of event e are
        counts[off[e] + e : off[e + 1] + e + 1] (cause -1 first, then each predecessor position)."""
        off = np.zeros(self.E + 1, dtype=np.int64)
        check(lib().mcdp_propagator_pred_offsets(self.h, _ptr(off, _pi64)))
        return off

    def run_stats(self, seeds, q=(0.5, 0.9, 0.99)):
        """run_many on the device, then per-event statistics; returns (mean [E], quantiles [E, len(q)],
        cause_counts [off[E] + E])."""
        seeds = np.ascontiguousarray(seeds, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.float64)
        off = self.pred_offsets()
        mean = np.empty(self.E)
        quant = np.empty((self.E, q.size))
        counts = np.empty(int(off[-1]) + self.E, dtype=np.int64)
        check(lib().mcdp_run_stats_host(self.h, _ptr(seeds, _pi32), seeds.size, q.size, _ptr(q, _pf64),
                                        _ptr(mean, _pf64), _ptr(quant, _pf64), _ptr(counts, _pi64)))
        return mean, quant, counts

    def event_stats_device(self, realized: "DeviceBuffer", cause: "DeviceBuffer", n: int, ld: int, q=(0.5,),
                           stream=None):
        """Statistics of event-major device arrays; results copied to the host."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        off = self.pred_offsets()
        d_mean = DeviceBuffer(8 * max(self.E, 1))
        d_quant = DeviceBuffer(8 * max(self.E * q.size, 1))
        d_cnt = DeviceBuffer(8 * max(int(off[-1]) + self.E, 1))
        check(lib().mcdp_event_stats_device(self.h, realized.ptr, cause.ptr if cause is not None else None, n, ld,
                                            q.size, _ptr(q, _pf64), d_mean.ptr, d_quant.ptr if q.size else None,
                                            d_cnt.ptr if cause is not None else None, stream))
        synchronize(stream)
        mean = d_mean.to_array(np.float64, (self.E,))
        quant = d_quant.to_array(np.float64, (self.E, q.size)) if q.size else np.empty((self.E, 0))
        counts = d_cnt.to_array(np.int64, (int(off[-1]) + self.E,)) if cause is not None else None
        return mean, quant, counts

    def set_times(self, dmin=None, earliest=None) -> None:
        """Replaces the minimal duration of every link ([L], by Activity.idx) and/or the earliest time of
        every event ([E]); the schedule is kept."""
        d = None if dmin is None else np.ascontiguousarray(dmin, dtype=np.float64)
        e = None if earliest is None else np.ascontiguousarray(earliest, dtype=np.float64)
        if d is not None and d.shape != (self.L,):
            raise ValueError(f"dmin must have shape ({self.L},)")
        if e is not None and e.shape != (self.E,):
            raise ValueError(f"earliest must have shape ({self.E},)")
        check(lib().mcdp_propagator_set_times(self.h, None if d is None else _ptr(d, _pf64),
                                              None if e is None else _ptr(e, _pf64)))

    def get_times(self) -> tuple[np.ndarray, np.ndarray]:
        """(minimal duration per link [L], earliest per event [E])."""
        d = np.empty(self.L)
        e = np.empty(self.E)
        check(lib().mcdp_propagator_get_times(self.h, _ptr(d, _pf64), _ptr(e, _pf64)))
        return d, e

    GRADIENT_FIELDS = ["steps", "waves", "slots", "lds_slots", "columns", "samples_per_lane"]

    def gradient_info(self) -> dict:
        """Facts about the reverse walk's plan (built on the host)."""
        v = np.zeros(len(self.GRADIENT_FIELDS), dtype=np.int64)
        check(lib().mcdp_propagator_gradient_info(self.h, _ptr(v, _pi64), v.size))
        return dict(zip(self.GRADIENT_FIELDS, map(int, v)))

    def gradient(self, seeds, weights, chunk: int = 0) -> tuple[float, np.ndarray, np.ndarray]:
        """Means over the seeds of J = sum_e w_e max(0, realized_e - earliest_e) and of its pathwise
        derivatives: (J, dJ/d minimal_duration [L], dJ/d earliest [E])."""
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (self.E,):
            raise ValueError(f"weights must have shape ({self.E},)")
        obj = np.zeros(1)
        gd = np.empty(self.L)
        ge = np.empty(self.E)
        check(lib().mcdp_gradient_host(self.h, _ptr(s, _pi32), s.size, _ptr(w, _pf64), chunk, _ptr(obj, _pf64),
                                       _ptr(gd, _pf64), _ptr(ge, _pf64)))
        return float(obj[0]), gd, ge

    def reverse_device(self, seeds, weights, realized, durations, cause, n: int, ld: int, out_obj, out_dmin,
                       out_earliest, accumulate: bool = False, stream=None) -> None:
        """The reverse walk over device results of run_many_device; outputs are device buffers of sums."""
        def p(b):
            return b.ptr if b is not None else None

        check(lib().mcdp_reverse_device(self.h, p(seeds), p(weights), p(realized), p(durations), p(cause), n, ld,
                                        int(accumulate), p(out_obj), p(out_dmin), p(out_earliest), stream))

    def autotune(self, n_samples: int, path: str | None = None) -> tuple[int, int, int, float]:
        """Times every candidate shape with n_samples samples; keeps and returns
        (waves, events_per_wave, samples_per_lane, ms).  With a tuning file, a choice recorded there for
        this DAG structure and device is configured without timing anything, and a new choice is added to it
        (``last_tune_from_file`` says which happened)."""
        w, k, spl, ms = _i32(0), _i32(0), _i32(0), C.c_float(0)
        if path is None:
            check(lib().mcdp_propagator_autotune(self.h, n_samples, C.byref(w), C.byref(k), C.byref(spl),
                                                 C.byref(ms)))
            self.last_tune_from_file = False
        else:
            hit = _i32(0)
            check(lib().mcdp_propagator_autotune_cached(self.h, n_samples, os.fsencode(path), C.byref(w), C.byref(k),
                                                        C.byref(spl), C.byref(ms), C.byref(hit)))
            self.last_tune_from_file = bool(hit.value)
        return w.value, k.value, spl.value, float(ms.value)

    def fingerprint(self) -> int:
        """64-bit fingerprint of what the schedule shape's speed depends on (structure, distribution kinds)."""
        out = C.c_uint64(0)
        check(lib().mcdp_propagator_fingerprint(self.h, C.byref(out)))
        return int(out.value)

    def run_many(self, seeds) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Host run_many: (realized (n,E), durations (n,L), cause (n,E)) as transposed views."""
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        n = s.size
        r = np.empty((self.E, n))
        d = np.empty((self.L, n))
        c = np.empty((self.E, n), dtype=np.int32)
        check(lib().mcdp_run_many_host(self.h, _ptr(s, _pi32), n, _ptr(r, _pf64), _ptr(d, _pf64), _ptr(c, _pi32)))
        return r.T, d.T, c.T

    def propagate(self, durations: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Host propagation with injected durations (n, L): (realized (n,E), cause (n,E))."""
        dT = np.ascontiguousarray(np.asarray(durations, dtype=np.float64).T)
        n = dT.shape[1] if dT.ndim == 2 else 0
        r = np.empty((self.E, n))
        c = np.empty((self.E, n), dtype=np.int32)
        check(lib().mcdp_propagate_host(self.h, _ptr(dT, _pf64), n, _ptr(r, _pf64), _ptr(c, _pi32)))
        return r.T, c.T

    def run_many_device(self, seeds: DeviceBuffer, n: int, ld: int, realized, durations, cause, stream=None) -> None:
        def p(b):
            return b.ptr if b is not None else None

        check(lib().mcdp_run_many_device(self.h, seeds.ptr, n, ld, p(realized), p(durations), p(cause), stream))

    def propagate_device(self, durations: DeviceBuffer, n: int, ld: int, realized, cause, stream=None) -> None:
        def p(b):
            return b.ptr if b is not None else None

        check(lib().mcdp_propagate_device(self.h, durations.ptr, n, ld, p(realized), p(cause), stream))

    def __del__(self) -> None:  # pragma: no cover
        try:
            lib().mcdp_propagator_destroy(self.h)
        except Exception:
            pass


def debug_philox(variant: int, ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Device Philox-4x32-10 (variant 4) or -2x32-10 (variant 2) on n counters."""
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(key, dtype=np.uint32)
    n = ctr.shape[0]
    out = np.empty_like(ctr)
    check(lib().mcdp_debug_philox(variant, n, _ptr(ctr, _pu32), _ptr(key, _pu32), _ptr(out, _pu32)))
    return out


def debug_math(fn: str, x: np.ndarray) -> np.ndarray:
    """The device samplers' log / cos2pi / exp on host inputs."""
    code = {"log": 0, "cos2pi": 1, "exp": 2}[fn]
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    check(lib().mcdp_debug_math(code, x.size, _ptr(x, _pf64), _ptr(out, _pf64)))
    return out


class Batch:
    """Several propagators run in one persistent launch (mcdp_batch_*)."""

    def __init__(self, props: list, waves: int = 0, events_per_wave: int = 0, samples_per_lane: int = 0) -> None:
        self.props = list(props)  # keeps the propagators alive
        arr = (_P * len(self.props))(*[p.h for p in self.props])
        self.h = C.c_void_p(0)
        check(lib().mcdp_batch_create(arr, len(self.props), waves, events_per_wave, samples_per_lane,
                                      C.byref(self.h)))

    def run_many_device(self, seeds, n, ld, realized, durations=None, cause=None, stream=None) -> None:
        """Device buffers per propagator (lists of DeviceBuffer or None)."""
        k = len(self.props)

        def ptrs(bufs):
            if bufs is None:
                return None
            return (_P * k)(*[(b.ptr if b is not None else None) for b in bufs])

        nn = np.ascontiguousarray(n, dtype=np.int64)
        ll = np.ascontiguousarray(ld, dtype=np.int64)
        check(lib().mcdp_batch_run_device(self.h, ptrs(seeds), _ptr(nn, _pi64), _ptr(ll, _pi64), ptrs(realized),
                                          ptrs(durations), ptrs(cause), stream))

    def prepare_device(self, seeds, n, ld, realized, durations=None, cause=None) -> None:
        """Uploads the launch descriptors for these device buffers (mcdp_batch_prepare_device)."""
        k = len(self.props)

        def ptrs(bufs):
            if bufs is None:
                return None
            return (_P * k)(*[(b.ptr if b is not None else None) for b in bufs])

        nn = np.ascontiguousarray(n, dtype=np.int64)
        ll = np.ascontiguousarray(ld, dtype=np.int64)
        check(lib().mcdp_batch_prepare_device(self.h, ptrs(seeds), _ptr(nn, _pi64), _ptr(ll, _pi64), ptrs(realized),
                                              ptrs(durations), ptrs(cause)))

    def launch(self, stream=None) -> None:
        """Enqueues the prepared launch (capturable into a graph)."""
        check(lib().mcdp_batch_launch(self.h, stream))

    def run_many(self, seed_lists) -> list:
        """Host convenience: [(realized (n, E), durations (n, L), cause (n, E))] per propagator."""
        bufs = []
        for P, sd in zip(self.props, seed_lists):
            sd = np.ascontiguousarray(sd, dtype=np.int32)
            n = sd.size
            bufs.append((n, DeviceBuffer.from_array(sd), DeviceBuffer(8 * P.E * max(n, 1)),
                         DeviceBuffer(8 * P.L * max(n, 1)), DeviceBuffer(4 * P.E * max(n, 1))))
        self.run_many_device([b[1] for b in bufs], [b[0] for b in bufs], [max(b[0], 1) for b in bufs],
                             [b[2] for b in bufs], [b[3] for b in bufs], [b[4] for b in bufs])
        out = []
        for P, (n, _, r, d, c) in zip(self.props, bufs):
            out.append((r.to_array(np.float64, (P.E, n)).T, d.to_array(np.float64, (P.L, n)).T,
                        c.to_array(np.int32, (P.E, n)).T))
        return out

    def stats(self, seed_lists, q=(0.5, 0.9, 0.99), memory_budget: int = 64 << 30, cause_counts: bool = True) -> list:
        """Per-event statistics of every DAG over its seeds (mcdp_batch_stats_host): DAGs in groups whose
        device buffers fit memory_budget bytes, one batched launch per group.  [(mean (E,), quantiles
        (E, n_q), counts or None)] per propagator."""
        k = len(self.props)
        seeds = [np.ascontiguousarray(sd, dtype=np.int32) for sd in seed_lists]
        if len(seeds) != k:
            raise ValueError(f"need {k} seed lists")
        q = np.ascontiguousarray(q, dtype=np.float64)
        nn = np.array([sd.size for sd in seeds], dtype=np.int64)
        means = [np.empty(P.E) for P in self.props]
        quants = [np.empty((P.E, q.size)) for P in self.props]
        counts = [np.empty(int(P.pred_offsets()[-1]) + P.E, dtype=np.int64) for P in self.props] if cause_counts \
            else None

        def ptrs(arrs):
            return None if arrs is None else (_P * k)(*[a.ctypes.data_as(_P) for a in arrs])

        check(lib().mcdp_batch_stats_host(self.h, ptrs(seeds), _ptr(nn, _pi64), q.size, _ptr(q, _pf64),
                                          int(memory_budget), ptrs(means), ptrs(quants) if q.size else None,
                                          ptrs(counts)))
        return [(means[i], quants[i], counts[i] if counts is not None else None) for i in range(k)]

    def __del__(self) -> None:  # pragma: no cover
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.mcdp_batch_destroy(self.h)
            self.h = C.c_void_p(0)


class Server:
    """Statistics requests over resident contexts, admitted by device-memory footprint (mcdp_server_*)."""

    STAT_FIELDS = ["submitted", "completed", "failed", "queued", "running", "max_running", "bytes_in_use",
                   "peak_bytes", "chunked", "exact_rounds", "kept_rows"]

    def __init__(self, memory_budget: int, max_running: int = 8, n_streams: int = 4, request_cap: int = 0,
                 exact_chunked_quantiles: bool = True) -> None:
        """request_cap: a request needing more device memory runs in chunks of samples (0: half the pool).
        exact_chunked_quantiles: such requests' quantiles are exact (further passes over their chunks); False
        places them inside a histogram bin instead (within max_delay / 4096), for lower latency."""
        self.h = C.c_void_p(0)
        check(lib().mcdp_server_create(int(memory_budget), int(request_cap), max_running, n_streams, C.byref(self.h)))
        self._ctx: list[Propagator] = []
        self._req: dict[int, tuple[int, int, int]] = {}
        if not exact_chunked_quantiles:
            check(lib().mcdp_server_set_exact_quantiles(self.h, 0))

    def debug_exact_limits(self, c_max: int, cand_per_target: int, refine_slots: int) -> None:
        """Lower limits of the exact selection for later requests (test hook)."""
        check(lib().mcdp_debug_server_exact_limits(self.h, c_max, cand_per_target, refine_slots))

    def debug_kept_rows(self, rows: int) -> None:
        """Rows kept whole per chunked request, for later requests (test hook): -1 the default, 0 none."""
        check(lib().mcdp_debug_server_kept_rows(self.h, rows))

    def add_context(self, prop: Propagator) -> int:
        cid = _i32(0)
        check(lib().mcdp_server_add_context(self.h, prop.h, C.byref(cid)))
        self._ctx.append(prop)
        return cid.value

    def submit(self, ctx: int, seeds, q=(0.5, 0.9, 0.99)) -> int:
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        rid = _i64(0)
        check(lib().mcdp_server_submit(self.h, ctx, _ptr(s, _pi32), s.size, qq.size, _ptr(qq, _pf64), C.byref(rid)))
        p = self._ctx[ctx]
        self._req[rid.value] = (p.E, int(p.pred_offsets()[-1]), qq.size)
        return rid.value

    def wait(self, rid: int):
        """(mean [E], quantiles [E, n_q], cause_counts [off[E] + E], latency_ms)."""
        E, P, nq = self._req.pop(rid)
        mean = np.empty(E)
        quant = np.empty((E, nq))
        counts = np.empty(E + P, dtype=np.int64)
        lat = _f64(0)
        check(lib().mcdp_server_wait(self.h, rid, _ptr(mean, _pf64), _ptr(quant, _pf64), _ptr(counts, _pi64),
                                     C.byref(lat)))
        return mean, quant, counts, lat.value

    def stats(self) -> dict:
        v = np.zeros(len(self.STAT_FIELDS), dtype=np.int64)
        check(lib().mcdp_server_stats(self.h, _ptr(v, _pi64), v.size))
        return dict(zip(self.STAT_FIELDS, map(int, v)))

    def close(self) -> None:
        if self.h:
            lib().mcdp_server_destroy(self.h)
            self.h = C.c_void_p(0)

    def __del__(self) -> None:  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """HIP events on a stream: elapsed milliseconds between two records."""

    def __init__(self) -> None:
        self.a, self.b = C.c_void_p(0), C.c_void_p(0)
        check(lib().mcdp_event_create(C.byref(self.a)))
        check(lib().mcdp_event_create(C.byref(self.b)))

    def start(self, stream=None) -> None:
        check(lib().mcdp_event_record(self.a, stream))

    def stop(self, stream=None) -> None:
        check(lib().mcdp_event_record(self.b, stream))

    def elapsed_ms(self) -> float:
        ms = C.c_float(0)
        check(lib().mcdp_event_elapsed_ms(self.a, self.b, C.byref(ms)))
        return float(ms.value)


def synchronize(stream=None) -> None:
    check(lib().mcdp_stream_synchronize(stream))
