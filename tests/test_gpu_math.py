"""The samplers' FP64 elementary functions on the device against extended-precision host values.

log_fast, cos2pi_fast and exp_fast (mc_dagprop_amd/csrc/mcdp_math.hpp) feed every exponential and
gamma draw.  The reference uses libstdc++'s std::log / std::exp / std::cos through its
distributions (_core.cpp:78-107); here each device function is checked against the host's 80-bit
long double evaluation over the domain the samplers use, and must stay within a few ulp.
"""

from __future__ import annotations

import numpy as np
import pytest

from mc_dagprop_amd import capi

pytestmark = pytest.mark.gpu

LD = np.longdouble


def ulp_error(got: np.ndarray, ref_ld: np.ndarray) -> np.ndarray:
    ref = ref_ld.astype(np.float64)
    return np.abs(got.astype(LD) - ref_ld) / np.spacing(np.abs(ref)).astype(LD)


def uniforms53(rng: np.random.Generator, n: int) -> np.ndarray:
    """(m + 1) 2^-53 for random 53-bit m: the log argument of the exponential samplers."""
    m = rng.integers(0, 2**53, n, dtype=np.int64)
    return (m.astype(np.float64) + 1.0) * 2.0**-53


def test_log_fast(gpu) -> None:
    assert np.finfo(LD).nmant >= 63, "needs x87 long double on the host"
    rng = np.random.default_rng(11)
    x = np.concatenate([
        uniforms53(rng, 1 << 20),
        np.ldexp(uniforms53(rng, 1 << 16), -rng.integers(0, 60, 1 << 16)),  # tiny uniforms, products of two
        1.0 + rng.uniform(-2**-20, 2**-20, 1 << 14),                         # near 1: log ~ 0
        np.exp(rng.uniform(-700, 700, 1 << 14)),                             # the full normal range
        np.array([2.0**-106, 2.0**-53, 0.5, 1.0 - 2**-53, 1.0, np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 1)]),
    ])
    got = capi.debug_math("log", x)
    ref = np.log(x.astype(LD))
    err = ulp_error(got, ref)
    assert got[np.flatnonzero(x == 1.0)[0]] == 0.0
    assert float(err.max()) <= 2.0, (float(err.max()), x[np.argmax(err)])


def test_cos2pi_fast(gpu) -> None:
    rng = np.random.default_rng(12)
    w = rng.integers(0, 2**32, 1 << 20, dtype=np.int64)
    x = np.concatenate([(w + 0.5) * 2.0**-32, rng.uniform(0, 1, 1 << 16),
                        np.array([0.0, 0.25, 0.5, 0.75, 1.0, 0.125, 0.375])])  # the samplers' angles
    got = capi.debug_math("cos2pi", x)
    # np.pi is pi rounded to double (~1e-16 relative in the angle): use pi to 64 bits instead
    ref = np.cos(2 * _pi_ld() * x.astype(LD))
    abs_err = np.abs(got.astype(LD) - ref)
    # Absolute error (cos is a factor of a normal deviate): the rounding of the angle (<= 1.7e-16 at
    # |a| <= pi/2) plus that of the alternating Taylor sum, whose terms add up to cosh(pi/2) ~ 2.5.
    assert float(abs_err.max()) < 4e-16, (float(abs_err.max()), x[np.argmax(abs_err)])
    assert got[-7] == 1.0 and got[-3] == 1.0 and got[-5] == -1.0
    assert abs(got[-6]) < 1e-16 and abs(got[-4]) < 1e-16


def _pi_ld() -> np.longdouble:
    # atan(1) * 4 in long double: pi to 64 bits
    return np.arctan(LD(1)) * 4


def test_exp_fast(gpu) -> None:
    rng = np.random.default_rng(13)
    x = np.concatenate([-rng.exponential(1.0, 1 << 18), rng.uniform(-708, 0, 1 << 18),
                        np.array([0.0, -1e-300, -np.log(2.0), -708.0])])
    got = capi.debug_math("exp", x)
    ref = np.exp(x.astype(LD))
    err = ulp_error(got, ref)
    assert got[-4] == 1.0
    assert float(err.max()) <= 2.0, (float(err.max()), x[np.argmax(err)])
