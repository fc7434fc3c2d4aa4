"""The rate expressions' logarithm (log_fast in hyperflow_amd/csrc/hf_device.hpp),
restated in numpy: x = f 2^e with f in [sqrt(1/2), sqrt(2)), ln f = 2 atanh(s),
s = (f - 1) / (f + 1), by its series to s^21, and e ln 2 added in a two-part
split. Checked against 40-digit logarithms over the whole double range. CPU
only."""
from decimal import Decimal, getcontext

import numpy as np

LN2_HI = float.fromhex("0x1.62e42ffp-1")
LN2_LO = -4.2009150726810846e-11
INV_LN10 = 0.4342944819032518


def log_fast(x):
    x = np.asarray(x, dtype=np.float64)
    f, e = np.frexp(x)
    low = f < 0.70710678118654752440
    f = np.where(low, f + f, f)
    e = np.where(low, e - 1, e).astype(np.float64)
    s = (f - 1.0) / (f + 1.0)
    z = s * s
    p = np.full_like(s, 2.0 / 21.0)
    for c in (19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = p * z + 2.0 / c
    return e * LN2_HI + (e * LN2_LO + ((s * z) * p + (s + s)))


def exact_ln(x):
    getcontext().prec = 40
    return np.array([float(Decimal(float(v)).ln()) for v in x])


def test_log_is_correct_to_a_few_ulp():
    rng = np.random.default_rng(7)
    x = np.concatenate([10.0 ** rng.uniform(-307, 308, 1500), rng.uniform(0.5, 2.0, 500),
                        rng.uniform(200.0, 4000.0, 500), [5e-324, 2.2e-308, 0.7071067811865476, 1.0 + 2.0 ** -52,
                                                            1.4142135623730951, 1.7e308]])
    ref = exact_ln(x)
    got = log_fast(x)
    scale = np.spacing(np.maximum(np.abs(ref), 1e-300))
    err = np.abs(got - ref) / scale
    assert err.max() <= 2, err.max()
    assert log_fast(np.array([1.0]))[0] == 0.0


def test_log10_of_the_falloff_range():
    """log10 = ln x / ln 10 within 3 ulp over the Troe arguments (Fcent and
    reduced pressures, 1e-300 .. 1e30)."""
    x = 10.0 ** np.random.default_rng(8).uniform(-300, 30, 2000)
    getcontext().prec = 40
    ref = np.array([float(Decimal(float(v)).log10()) for v in x])
    got = log_fast(x) * INV_LN10
    err = np.abs(got - ref) / np.spacing(np.abs(ref))
    assert err.max() <= 3, err.max()
