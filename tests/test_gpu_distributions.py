"""GPU sampler vs the distributions the reference samples (_core.cpp:72-141).

The device draws do not reproduce the reference's Xoshiro256++/libstdc++
streams; they are matched in law.  For every sampling method the device
selects, ~1e6 device draws are compared with the closed-form distribution:
Kolmogorov-Smirnov distance below the alpha = 0.001 critical value
1.95 / sqrt(n), and the first four moments within 6 standard errors.
Empirical tables: every bin's frequency within 5 binomial standard errors of
w_i / sum(w), and zero-weight bins never drawn.

Draws come from activities no predecessor references (each activity is its
own Philox counter), on one event, so 64 activities x 16,384 seeds = 1,048,576
independent draws per case.  Durations are 1 + X (minimal_duration 1.0).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate, stats

import oracle
from mc_dagprop_amd import capi
from mc_dagprop_amd.workloads import (
    EMPIRICAL_ABSOLUTE,
    EMPIRICAL_RELATIVE,
    EXPONENTIAL,
    GAMMA,
    DistSpec,
    FlatDag,
)

pytestmark = pytest.mark.gpu

INF = float("inf")
N_ACTS = 64
N_SEEDS = 16384


def orphan_dag(dist: DistSpec, dmin: float = 1.0) -> FlatDag:
    n = N_ACTS
    return FlatDag(
        earliest=np.zeros(1), latest=np.zeros(1), actual=np.zeros(1),
        act_src=np.zeros(n, np.int32), act_dst=np.zeros(n, np.int32), act_idx=np.arange(n, dtype=np.int32),
        act_dur=np.full(n, dmin), act_type=np.full(n, dist.activity_type, np.int32),
        prec_target=np.zeros(0, np.int32), prec_off=np.zeros(1, np.int64), prec_src=np.zeros(0, np.int32),
        prec_edge=np.zeros(0, np.int32), max_delay=0.0, dists=[dist],
    )


def draws(dist: DistSpec, seeds=None) -> np.ndarray:
    P = capi.Propagator(orphan_dag(dist))
    s = np.arange(N_SEEDS, dtype=np.int32) if seeds is None else np.asarray(seeds, dtype=np.int32)
    _, d, _ = P.run_many(s)
    return np.ascontiguousarray(d).ravel() - 1.0


def moments_of(pdf, lo: float, hi: float):
    """Raw moments E[X^k], k = 1..8, of a density on [lo, hi] (numerical quadrature)."""
    z = integrate.quad(pdf, lo, hi, limit=400)[0]
    return [integrate.quad(lambda x, k=k: x**k * pdf(x) / z, lo, hi, limit=400)[0] for k in range(1, 9)]


def gamma_raw_moments(k: float, theta: float):
    """E[X^j] = theta^j Gamma(k + j) / Gamma(k), j = 1..8 (untruncated gamma)."""
    from scipy.special import gammaln

    return [float(np.exp(j * np.log(theta) + gammaln(k + j) - gammaln(k))) for j in range(1, 9)]


def check_law(x: np.ndarray, cdf, mom, floor: float = 1e-10) -> None:
    """KS + moments.  X is observed through duration = 1 + X, which cannot resolve X < 2^-53 (the
    reference's d + X d output has the same resolution), so the KS test compares the draws above `floor`
    with the law of X conditioned on X > floor; the moments use every draw."""
    n = x.size
    assert np.all(np.isfinite(x)) and np.all(x >= 0)
    f0 = float(cdf(np.array([floor]))[0])
    above = x[x > floor]
    D = stats.kstest(above, lambda t: (cdf(t) - f0) / (1.0 - f0)).statistic
    assert D < 1.95 / np.sqrt(above.size), f"KS distance {D:.5f} >= {1.95 / np.sqrt(above.size):.5f}"
    assert abs(above.size / n - (1.0 - f0)) < 6 * np.sqrt(f0 * (1 - f0) / n) + 1e-12, (above.size / n, 1 - f0)
    # First four raw moments, each within 6 standard errors sqrt((E[X^2k] - E[X^k]^2) / n).
    for j in range(1, 5):
        ej, e2j = mom[j - 1], mom[2 * j - 1]
        se = np.sqrt(max(e2j - ej * ej, 0.0) / n)
        got = float(np.mean(x**j))
        assert abs(got - ej) < 6 * se + 1e-12 * abs(ej), f"E[X^{j}] = {got} vs {ej} (se {se})"


EXP_CASES = [(2.0, INF), (2.0, 3.0), (1000.0, 1.0), (0.5, 0.1)]


@pytest.mark.parametrize("lam,m", EXP_CASES)
def test_exponential_law(gpu, lam: float, m: float) -> None:
    x = draws(DistSpec(1, EXPONENTIAL, (lam, m, INF)))
    assert x.max() <= m
    pdf = lambda t: np.exp(-t / lam) / lam  # noqa: E731
    hi = m if np.isfinite(m) else 60 * lam
    check_law(x, lambda t: oracle.cdf_exponential(t, lam, m), moments_of(pdf, 0.0, hi))


GAMMA_CASES = [
    (2.0, 1.5, INF),    # shape 2: sum of two exponentials
    (0.5, 2.0, INF),    # shape 1/2: squared normal / 2
    (3.7, 0.8, INF),    # Marsaglia-Tsang
    (9.0, 0.5, INF),    # Marsaglia-Tsang, larger shape
    (0.3, 1.0, INF),    # Marsaglia-Tsang with the U^{1/k} boost
    (0.05, 1.0, INF),   # tiny shape
    (1.0, 2.0, 5.0),    # shape 1: truncated exponential inversion
    (2.0, 1.0, 0.5),    # severe truncation: power proposal (test_monte_carlo_extra.py:47-56)
    (5.0, 1.0, 3.0),    # truncated Marsaglia-Tsang by rejection
    (0.5, 1.0, 0.1),    # truncated shape 1/2 by rejection
    (0.3, 2.0, 0.0005), # severe truncation, shape < 1: power proposal
]


@pytest.mark.parametrize("k,theta,m", GAMMA_CASES)
def test_gamma_law(gpu, k: float, theta: float, m: float) -> None:
    x = draws(DistSpec(2, GAMMA, (k, theta, m)))
    assert x.max() <= m
    if np.isfinite(m):
        pdf = lambda t: stats.gamma(k, scale=theta).pdf(t)  # noqa: E731
        mom = moments_of(pdf, 1e-300 if k < 1 else 0.0, m)  # k < 1: integrable singularity at 0
    else:
        mom = gamma_raw_moments(k, theta)
    check_law(x, lambda t: oracle.cdf_gamma(t, k, theta, m), mom)


def _freq_check(vals: np.ndarray, values: np.ndarray, weights: np.ndarray) -> None:
    p = oracle.pmf_empirical(weights)
    n = vals.size
    idx = np.searchsorted(values, vals)
    assert np.all(values[np.clip(idx, 0, len(values) - 1)] == vals), "draw outside the table"
    counts = np.bincount(idx, minlength=len(values))
    se = np.sqrt(n * p * (1 - p))
    assert np.all(counts[p == 0] == 0)
    z = np.abs(counts - n * p)[p > 0] / np.maximum(se[p > 0], 1e-12)
    assert z.max() < 5.0, z.max()


def test_empirical_absolute_frequencies(gpu) -> None:
    rng = np.random.default_rng(2)
    values = 7.03125 + 14.0625 * np.arange(64)
    w = rng.dirichlet(np.ones(64))
    w[[3, 17, 40]] = 0.0
    vals = draws(DistSpec(5, EMPIRICAL_ABSOLUTE, values=values, weights=w)) + 1.0  # d = 1: dur = 1 + v
    _freq_check(vals - 1.0, values, w)


def test_empirical_relative_frequencies(gpu) -> None:
    values = np.array([0.0, 0.25, 0.5, 1.0, 2.0])
    w = np.array([0.0, 3.0, 1.0, 0.0, 6.0])
    vals = draws(DistSpec(6, EMPIRICAL_RELATIVE, values=values, weights=w))
    _freq_check(vals, values, w)


def test_empirical_single_bin_is_constant(gpu) -> None:
    """n = 1 always draws bin 0, whatever its weight (libstdc++ discrete_distribution with < 2 weights)."""
    x = draws(DistSpec(5, EMPIRICAL_ABSOLUTE, values=np.array([42.0]), weights=np.array([0.0])), seeds=range(64))
    assert np.all(x == 42.0)


def test_draws_independent_across_activities_and_seeds(gpu) -> None:
    """Neighbouring activities and seeds give uncorrelated draws (|r| < 5 / sqrt(n))."""
    x = draws(DistSpec(1, EXPONENTIAL, (1.0, INF, INF))).reshape(N_SEEDS, N_ACTS)
    n = x.size
    r_act = np.corrcoef(x[:, :-1].ravel(), x[:, 1:].ravel())[0, 1]
    r_seed = np.corrcoef(x[:-1].ravel(), x[1:].ravel())[0, 1]
    assert abs(r_act) < 5 / np.sqrt(n) and abs(r_seed) < 5 / np.sqrt(n), (r_act, r_seed)


def test_exponential_stream_formula(gpu) -> None:
    """Documented stream: X = -mean log(((w0 >> 11) 2^32 + w1 + 1) 2^-53),
    (w0, w1) = Philox-2x32-10(ctr = (seed, 0), key = a) for activity a."""
    lam = 3.0
    seeds = np.arange(-40, 40, dtype=np.int32)
    x = draws(DistSpec(1, EXPONENTIAL, (lam, INF, INF)), seeds=seeds).reshape(seeds.size, N_ACTS)
    a = np.tile(np.arange(N_ACTS, dtype=np.uint32), seeds.size)
    s = np.repeat(seeds.astype(np.uint32), N_ACTS)
    ctr = np.stack([s, np.zeros_like(s)], axis=1)
    key = a
    w = oracle.philox2x32(ctr, key)
    m = (w[:, 0] >> np.uint32(11)).astype(np.float64) * 2.0**32 + w[:, 1].astype(np.float64)
    expect = (1.0 + (-lam * np.log((m + 1.0) * 2.0**-53))) - 1.0
    np.testing.assert_allclose(x.ravel(), expect, rtol=1e-13, atol=1e-15)
