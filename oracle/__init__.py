"""CPU oracle of the Monte Carlo DAG propagation path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The product
(``mc_dagprop_amd``) never imports it.

Parts (the C sources say, function by function, which reference lines they
follow and what pins them):

* :class:`OracleCtx` -- ``mcdp_oracle.c``: the reference's context compile
  (_core.cpp:193-307) and propagation loop (_core.cpp:316-350) restated in C.
  Pinned against the compiled reference (:class:`RefSim`) and the reference's
  own test KATs.
* :func:`philox4x32` / :func:`philox2x32` -- Philox-4x32-10 / -2x32-10, pinned by
  the Random123 known-answer vectors.
* :meth:`OracleCtx.gradient` -- ``orc_gradient``: the pathwise gradient of the
  expected weighted lateness, pinned by finite differences of the oracle's own
  propagation and by closed forms.
* :class:`RefSim` -- the reference engine itself, compiled from its sources by
  ``oracle/Makefile`` into ``oracle/_ref/`` (the stream-exact durations of
  Xoshiro256++ + libstdc++ come only from here).
* Closed forms of the distributions the reference samples (_core.cpp:72-141):
  truncated exponential and gamma CDFs, the empirical pmf, and the
  no-randomness durations (constant / no distribution).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "liboracle_mcdp.so"
REF_LIB = HERE / "_ref" / "libref_mcdp.so"

_P = C.c_void_p
_pi32, _pi64, _pf64, _pu32 = (C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_uint32))


def _ptr(a, t):
    return a.ctypes.data_as(t)


_olib = None


def oracle_lib() -> C.CDLL:
    global _olib
    if _olib is None:
        if not ORACLE_LIB.exists():
            raise FileNotFoundError(f"{ORACLE_LIB} missing: run `make -C oracle oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        L.orc_compile.restype = C.c_int
        L.orc_compile.argtypes = [
            C.c_int, _pf64, C.c_int, _pi32, _pf64, _pi32, C.c_int, _pi32, _pi64, _pi32, _pi32,
            C.c_double, C.c_int, _pi32, C.POINTER(_P), C.c_char_p, C.c_int,
        ]
        L.orc_free.argtypes = [_P]
        L.orc_event_count.argtypes = [_P]
        L.orc_link_count.argtypes = [_P]
        L.orc_get_order.argtypes = [_P, _pi32]
        L.orc_base_durations.argtypes = [_P, C.c_int, _pi32, _pi32, _pf64, _pf64]
        L.orc_propagate.argtypes = [_P, _pf64, C.c_int64, C.c_int64, _pf64, _pi32]
        L.orc_propagate_columns.argtypes = [_P, _pf64, C.c_int64, C.c_int64, _pf64, _pi32, C.c_int64]
        L.orc_philox4x32_many.argtypes = [C.c_int, C.c_int64, _pu32, _pu32, _pu32]
        L.orc_philox2x32_many.argtypes = [C.c_int, C.c_int64, _pu32, _pu32, _pu32]
        L.orc_get_links.argtypes = [_P, _pf64, _pi32]
        L.orc_gradient.restype = C.c_int
        L.orc_gradient.argtypes = [_P, _pf64, _pi32, _pf64, C.c_int64, _pf64, C.c_int64, _pf64, _pi32, _pf64, _pf64,
                                   _pf64, C.c_char_p, C.c_int]
        _olib = L
    return _olib


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str) -> None:
        super().__init__(msg)
        self.code = code


class OracleCtx:
    """The reference's compiled context (Kahn order + CSR) and its propagation loop, in C."""

    def __init__(self, dag) -> None:
        L = oracle_lib()
        self.dag = dag
        reg = np.asarray(sorted({d.activity_type for d in dag.dists}), dtype=np.int32)
        self._keep = [
            np.ascontiguousarray(dag.earliest, dtype=np.float64),
            np.ascontiguousarray(dag.act_idx, dtype=np.int32),
            np.ascontiguousarray(dag.act_dur, dtype=np.float64),
            np.ascontiguousarray(dag.act_type, dtype=np.int32),
            np.ascontiguousarray(dag.prec_target, dtype=np.int32),
            np.ascontiguousarray(dag.prec_off, dtype=np.int64),
            np.ascontiguousarray(dag.prec_src, dtype=np.int32),
            np.ascontiguousarray(dag.prec_edge, dtype=np.int32),
            reg,
        ]
        e, ai, ad, at, pt, po, ps, pe, rg = self._keep
        err = C.create_string_buffer(256)
        self.h = _P(0)
        rc = L.orc_compile(
            e.size, _ptr(e, _pf64), ai.size, _ptr(ai, _pi32), _ptr(ad, _pf64), _ptr(at, _pi32), pt.size,
            _ptr(pt, _pi32), _ptr(po, _pi64), _ptr(ps, _pi32), _ptr(pe, _pi32), float(dag.max_delay), rg.size,
            _ptr(rg, _pi32), C.byref(self.h), err, 256,
        )
        if rc:
            raise OracleError(rc, err.value.decode())
        self.E = L.orc_event_count(self.h)
        self.L = L.orc_link_count(self.h)

    @property
    def order(self) -> np.ndarray:
        o = np.empty(self.E, dtype=np.int32)
        oracle_lib().orc_get_order(self.h, _ptr(o, _pi32))
        return o

    def base_durations(self) -> np.ndarray:
        """Durations of links that need no random draw (NaN where a draw is needed)."""
        # Last registration per type wins (README.md:44-45).
        last = {}
        for d in self.dag.dists:
            last[d.activity_type] = d
        types = np.asarray(list(last.keys()), dtype=np.int32)
        is_const = np.asarray([1 if d.kind == 0 else 0 for d in last.values()], dtype=np.int32)
        factor = np.asarray([d.params[0] if d.kind == 0 else 0.0 for d in last.values()], dtype=np.float64)
        out = np.empty(self.L)
        oracle_lib().orc_base_durations(
            self.h, types.size, _ptr(types, _pi32), _ptr(is_const, _pi32), _ptr(factor, _pf64), _ptr(out, _pf64)
        )
        return out

    def propagate(self, durations: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """realized (n, E), cause (n, E) for injected durations (n, L) (any row stride)."""
        d = np.asarray(durations, dtype=np.float64)
        if d.ndim == 1:
            d = d[None, :]
        d = np.ascontiguousarray(d)
        n = d.shape[0]
        r = np.empty((n, self.E))
        c = np.empty((n, self.E), dtype=np.int32)
        oracle_lib().orc_propagate(self.h, _ptr(d, _pf64), d.shape[1], n, _ptr(r, _pf64), _ptr(c, _pi32))
        return r, c

    def propagate_columns(self, durations: np.ndarray, realized_out: np.ndarray | None = None,
                          cause_out: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
        """Event-major form: durations (L, n) -> realized (E, n), cause (E, n).  Rows may have any
        stride (columns contiguous); given outputs (views of one leading dimension) are written in
        place.  The C call releases the GIL, so threads may run it on disjoint column ranges."""
        d = np.asarray(durations, dtype=np.float64)
        if d.ndim != 2 or d.shape[0] != self.L or (d.size and d.strides[1] != 8):
            d = np.ascontiguousarray(d)
        n = d.shape[1]
        ld = d.strides[0] // 8 if d.size else n
        if realized_out is None:
            r = np.empty((self.E, n))
            c = np.empty((self.E, n), dtype=np.int32)
        else:
            r, c = realized_out, cause_out
            if r.shape != (self.E, n) or c.shape != (self.E, n) or r.dtype != np.float64 or c.dtype != np.int32:
                raise ValueError("outputs must be (E, n) float64 and int32")
            if n and (r.strides[1] != 8 or c.strides[1] != 4 or r.strides[0] // 8 != c.strides[0] // 4):
                raise ValueError("outputs need contiguous columns and one leading dimension")
        ldo = (r.strides[0] // 8) if (n and self.E) else max(n, 1)
        oracle_lib().orc_propagate_columns(self.h, d.ctypes.data_as(_pf64), ld, n, r.ctypes.data_as(_pf64),
                                           c.ctypes.data_as(_pi32), ldo)
        return r, c

    def links(self) -> tuple[np.ndarray, np.ndarray]:
        """The compiled link table: (minimal duration [L], activity type [L])."""
        d = np.empty(self.L)
        t = np.empty(self.L, dtype=np.int32)
        oracle_lib().orc_get_links(self.h, _ptr(d, _pf64), _ptr(t, _pi32))
        return d, t

    def link_relative(self) -> np.ndarray:
        """1 for links whose duration is d + X*d (constant, exponential, gamma, empirical relative), 0 for
        d + X (empirical absolute) or d (no distribution)."""
        last = {}
        for d in self.dag.dists:
            last[d.activity_type] = d.kind
        _, types = self.links()
        rel = {t for t, k in last.items() if k != 3}  # 3: empirical absolute
        return np.asarray([1 if int(t) in rel else 0 for t in types], dtype=np.int32)

    def gradient(self, realized, cause, durations, weights, mult1=None):
        """Sums over samples of (J_s, dJ_s/d minimal_duration [L], dJ_s/d earliest [E]), with
        J_s = sum_e w_e max(0, realized_e - earliest_e); inputs sample-major (n, E) / (n, L)."""
        r = np.ascontiguousarray(realized, dtype=np.float64)
        c = np.ascontiguousarray(cause, dtype=np.int32)
        d = np.ascontiguousarray(durations, dtype=np.float64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        m1 = None if mult1 is None else np.ascontiguousarray(mult1, dtype=np.float64)
        rel = self.link_relative()
        n = r.shape[0]
        obj = np.zeros(1)
        gd = np.empty(self.L)
        ge = np.empty(self.E)
        err = C.create_string_buffer(256)
        rc = oracle_lib().orc_gradient(self.h, _ptr(r, _pf64), _ptr(c, _pi32), _ptr(d, _pf64), d.shape[1] if n else 0,
                                       None if m1 is None else _ptr(m1, _pf64), n, _ptr(w, _pf64), _ptr(rel, _pi32),
                                       _ptr(obj, _pf64), _ptr(gd, _pf64), _ptr(ge, _pf64), err, 256)
        if rc:
            raise OracleError(rc, err.value.decode())
        return float(obj[0]), gd, ge

    def __del__(self) -> None:  # pragma: no cover
        try:
            oracle_lib().orc_free(self.h)
        except Exception:
            pass


def philox4x32(ctr: np.ndarray, key: np.ndarray, rounds: int = 10) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.empty_like(ctr)
    oracle_lib().orc_philox4x32_many(rounds, ctr.shape[0], _ptr(ctr, _pu32), _ptr(key, _pu32), _ptr(out, _pu32))
    return out


def philox2x32(ctr: np.ndarray, key: np.ndarray, rounds: int = 10) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 2)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1)
    out = np.empty_like(ctr)
    oracle_lib().orc_philox2x32_many(rounds, ctr.shape[0], _ptr(ctr, _pu32), _ptr(key, _pu32), _ptr(out, _pu32))
    return out


# ---- the compiled reference (oracle/_ref) ------------------------------------
_rlib = None


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib() -> C.CDLL:
    global _rlib
    if _rlib is None:
        if not REF_LIB.exists():
            raise FileNotFoundError(f"{REF_LIB} missing: run `make -C oracle ref` where the reference sources exist")
        L = C.CDLL(str(REF_LIB))
        L.ref_sim_create.restype = _P
        L.ref_sim_create.argtypes = [
            C.c_int, _pf64, _pf64, _pf64, C.c_int, _pi32, _pi32, _pi32, _pf64, _pi32, C.c_int, _pi32, _pi64,
            _pi32, _pi32, C.c_double, C.c_int, _pi32, _pi32, _pf64, _pi64, _pf64, _pf64, C.c_char_p, C.c_int,
        ]
        L.ref_sim_destroy.argtypes = [_P]
        L.ref_sim_node_count.argtypes = [_P]
        L.ref_sim_activity_count.argtypes = [_P]
        L.ref_sim_run.argtypes = [_P, _pi32, C.c_int64, _pf64, _pf64, _pi32]
        L.ref_run_many_timed.restype = C.c_double
        L.ref_run_many_timed.argtypes = [_P, _pi32, C.c_int64, C.c_int]
        _rlib = L
    return _rlib


class RefSim:
    """The reference's own Simulator (compiled from its sources), driven through ctypes."""

    def __init__(self, dag) -> None:
        L = ref_lib()
        tabs_v, tabs_w, off = [], [], [0]
        types, kinds, params = [], [], []
        for d in dag.dists:
            types.append(d.activity_type)
            kinds.append(d.kind)
            params.extend(d.params)
            if d.values is not None:
                tabs_v.extend(np.asarray(d.values, dtype=np.float64))
                tabs_w.extend(np.asarray(d.weights, dtype=np.float64))
            off.append(len(tabs_v))
        arrs = [
            np.ascontiguousarray(dag.earliest, np.float64), np.ascontiguousarray(dag.latest, np.float64),
            np.ascontiguousarray(dag.actual, np.float64), np.ascontiguousarray(dag.act_src, np.int32),
            np.ascontiguousarray(dag.act_dst, np.int32), np.ascontiguousarray(dag.act_idx, np.int32),
            np.ascontiguousarray(dag.act_dur, np.float64), np.ascontiguousarray(dag.act_type, np.int32),
            np.ascontiguousarray(dag.prec_target, np.int32), np.ascontiguousarray(dag.prec_off, np.int64),
            np.ascontiguousarray(dag.prec_src, np.int32), np.ascontiguousarray(dag.prec_edge, np.int32),
            np.asarray(types, np.int32), np.asarray(kinds, np.int32), np.asarray(params or [0.0], np.float64),
            np.asarray(off, np.int64), np.asarray(tabs_v or [0.0], np.float64), np.asarray(tabs_w or [0.0], np.float64),
        ]
        self._keep = arrs
        (ea, la, ac, asr, ads, ai, ad, at, pt, po, ps, pe, ty, ki, pa, of, tv, tw) = arrs
        err = C.create_string_buffer(512)
        self.h = L.ref_sim_create(
            ea.size, _ptr(ea, _pf64), _ptr(la, _pf64), _ptr(ac, _pf64), ai.size, _ptr(asr, _pi32), _ptr(ads, _pi32),
            _ptr(ai, _pi32), _ptr(ad, _pf64), _ptr(at, _pi32), pt.size, _ptr(pt, _pi32), _ptr(po, _pi64),
            _ptr(ps, _pi32), _ptr(pe, _pi32), float(dag.max_delay), ty.size, _ptr(ty, _pi32), _ptr(ki, _pi32),
            _ptr(pa, _pf64), _ptr(of, _pi64), _ptr(tv, _pf64), _ptr(tw, _pf64), err, 512,
        )
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.E = L.ref_sim_node_count(self.h)
        self.L = L.ref_sim_activity_count(self.h)

    def run(self, seeds) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Simulator::run for each seed in order: realized (n,E), durations (n,L), cause (n,E)."""
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        n = s.size
        r = np.empty((n, self.E))
        d = np.empty((n, self.L))
        c = np.empty((n, self.E), dtype=np.int32)
        ref_lib().ref_sim_run(self.h, _ptr(s, _pi32), n, _ptr(r, _pf64), _ptr(d, _pf64), _ptr(c, _pi32))
        return r, d, c

    def run_many_timed(self, seeds, threads: int = 1) -> float:
        s = np.ascontiguousarray(seeds, dtype=np.int32)
        return float(ref_lib().ref_run_many_timed(self.h, _ptr(s, _pi32), s.size, int(threads)))

    def __del__(self) -> None:  # pragma: no cover
        try:
            ref_lib().ref_sim_destroy(self.h)
        except Exception:
            pass


# ---- closed forms of the distributions (_core.cpp:72-141) ---------------------
def cdf_exponential(x, lam: float, max_scale: float = np.inf):
    """CDF of X ~ Exp(mean lam) conditioned on X <= max_scale (_core.cpp:78-90)."""
    x = np.asarray(x, dtype=np.float64)
    base = -np.expm1(-np.clip(x, 0, None) / lam)
    if np.isinf(max_scale):
        return base
    return np.clip(base / -np.expm1(-max_scale / lam), 0.0, 1.0)


def cdf_gamma(x, shape: float, scale: float, max_scale: float = np.inf):
    """CDF of X ~ Gamma(shape, scale) conditioned on X <= max_scale (_core.cpp:92-105)."""
    from scipy.special import gammainc

    x = np.asarray(x, dtype=np.float64)
    base = gammainc(shape, np.clip(x, 0, None) / scale)
    if np.isinf(max_scale):
        return base
    return np.clip(base / gammainc(shape, max_scale / scale), 0.0, 1.0)


def gamma_moments(shape: float, scale: float) -> tuple[float, float, float, float]:
    """Mean, variance, skewness and excess kurtosis of Gamma(shape, scale)."""
    return shape * scale, shape * scale**2, 2.0 / np.sqrt(shape), 6.0 / shape


def pmf_empirical(weights) -> np.ndarray:
    """P(bin i) = w_i / sum(w) (libstdc++ discrete_distribution, used at _core.cpp:121, 136)."""
    w = np.asarray(weights, dtype=np.float64)
    if w.size < 2:
        return np.ones(max(w.size, 1))
    return w / w.sum()


# ---- per-event statistics ---------------------------------------------------------
def pred_sources(dag) -> list[list[int]]:
    """Each event's predecessor sources in input order, from the last precedence entry naming it
    (_core.cpp:232-285: a target's last entry supplies its list)."""
    preds: list[list[int]] = [[] for _ in range(dag.n_events)]
    for i, t in enumerate(dag.prec_target):
        a, b = int(dag.prec_off[i]), int(dag.prec_off[i + 1])
        preds[int(t)] = [int(x) for x in dag.prec_src[a:b]]
    return preds


def event_stats(realized: np.ndarray, cause: np.ndarray | None, preds: list[list[int]], q) -> tuple:
    """Per-event statistics of sample-major results realized[s, e] / cause[s, e], as a consumer of
    run_many computes them on the host (demo/monte_carlo.py:45-63 gathers realized per event):

    * mean[e] = math.fsum(realized[:, e]) / n, the exactly rounded sum; rows whose float sum is not
      finite (inf, NaN) use that sum;
    * quantiles[e, j] = numpy.quantile(realized[:, e], q[j]) (method "linear");
    * counts: per event, [#cause == -1, #cause == source of predecessor 0, ...], a source listed twice
      counted at its first position.
    """
    import math

    realized = np.asarray(realized, dtype=np.float64)
    n, E = realized.shape
    mean = np.empty(E)
    for e in range(E):
        col = realized[:, e]
        plain = float(np.sum(col))
        if math.isfinite(plain):
            try:
                mean[e] = math.fsum(col.tolist()) / n
            except OverflowError:
                mean[e] = plain / n
        else:
            mean[e] = plain / n
    q = np.asarray(q, dtype=np.float64)
    # realized may hold +-inf (infinite durations or max_delay): numpy's linear interpolation
    # between two equal infinities evaluates inf - inf and warns; the device follows the same
    # interpolation formula, so both sides give the same value there
    with np.errstate(invalid="ignore"):
        quant = np.quantile(realized, q, axis=0).T.copy() if q.size else np.empty((E, 0))
    counts = None
    if cause is not None:
        cause = np.asarray(cause)
        # counters per event: cause -1, then one per predecessor position; a source listed twice counts at
        # its first position (its later positions stay 0)
        kmax = max((len(p) for p in preds), default=0)
        first_src = np.full((kmax, E), -2, dtype=np.int64)  # -2: no counter matches (causes are >= -1)
        for e in range(E):
            seen = set()
            for k, src in enumerate(preds[e]):
                if src not in seen:
                    seen.add(src)
                    first_src[k, e] = src
        per_pos = [np.count_nonzero(cause == -1, axis=0)]
        per_pos += [np.count_nonzero(cause == first_src[k][None, :], axis=0) for k in range(kmax)]
        out = []
        for e in range(E):
            out.append(np.asarray([per_pos[k][e] for k in range(1 + len(preds[e]))], dtype=np.int64))
        counts = np.concatenate(out) if out else np.zeros(0, dtype=np.int64)
    return mean, quant, counts
