"""The step-size root of the GPU BDF integrators (root_n in
hyperflow_amd/csrc/hf_device.hpp), restated in numpy: x^(1/n) from a range
reduction x = m 2^(n k), a single-precision estimate of m^(1/n) and two
double-precision Newton steps on r^n = m. The device estimate comes from the
hardware log2/exp2 (relative error below 1e-5); here the estimate is
perturbed by up to 1e-5 either way, so the test covers any estimate within
that bound. CPU only."""
import numpy as np
import pytest


def root_n(x, n, delta):
    x = np.asarray(x, dtype=np.float64)
    rn = np.float32(1.0) / np.float32(n)
    _, e = np.frexp(x)
    k = np.trunc(e.astype(np.float32) * rn).astype(np.int64)
    m = np.ldexp(x, -n * k)
    r = np.exp2(np.log2(m.astype(np.float32)) * rn).astype(np.float64) * (1.0 + delta)
    c = np.float64(rn)
    for _ in range(2):
        r = r + (r * c) * (m / r ** n - 1.0)
    return np.ldexp(r, k)


def true_root(x, n):
    """x^(1/n) to 40 digits, rounded to double."""
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    return np.array([float(Decimal(float(v)) ** (Decimal(1) / Decimal(n))) for v in x])


@pytest.mark.parametrize("n", range(1, 8))
def test_root_is_correct_to_a_few_ulp(n):
    """Within 2 ulp of the exact root over the whole double range (pow(x,
    1.0/n) itself is up to about 100 ulp from it at the ends of the range,
    since 1.0/n is rounded: x^(1/n + e) = x^(1/n) x^e)."""
    rng = np.random.default_rng(n)
    x = np.concatenate([10.0 ** rng.uniform(-300, 300, 500), 10.0 ** rng.uniform(-6, 3, 200),
                        [2.2e-308, 1.0, 2.0 ** n, 3.0 ** n, 1.7e308]])
    ref = true_root(x, n)
    for delta in (-1e-5, 0.0, 1e-5):
        got = root_n(x, n, delta)
        err = np.abs(got - ref) / np.spacing(ref)
        assert err.max() <= 2, (n, delta, err.max())
    # exact powers give exact roots
    assert root_n(np.array([2.0 ** n, 3.0 ** n]), n, 0.0).tolist() == [2.0, 3.0]


def test_a_coarse_estimate_would_not_do():
    """One Newton step from a 1e-5 estimate leaves about 3e-10: the second
    step is what brings the result to pow's."""
    x = np.array([3.7e-12, 0.42, 9.1e17])
    n = 6
    rn = np.float32(1.0) / np.float32(n)
    r = x ** (1.0 / n) * (1.0 + 1e-5)
    one = r + (r * np.float64(rn)) * (x / r ** n - 1.0)
    rel = np.abs(one / x ** (1.0 / n) - 1.0)
    assert rel.max() > 1e-11 and rel.max() < 1e-9
