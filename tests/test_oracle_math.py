"""Pins of the oracle's correctly rounded cube root by exact arithmetic.

cbrt_exact(x) must be the double nearest to x^(1/3): with y = cbrt_exact(x)
and its neighbours y-, y+, the exact cube test
    ((y + y-)/2)^3 < x < ((y + y+)/2)^3
is decided with Python integers/fractions, independently of the C code.
"""
from fractions import Fraction

import numpy as np
import pytest


def _is_nearest(x, y):
    X = Fraction(x)
    lo = Fraction(np.nextafter(y, 0.0))
    hi = Fraction(np.nextafter(y, np.inf))
    Y = Fraction(y)
    m_lo, m_hi = (lo + Y) / 2, (Y + hi) / 2
    return m_lo ** 3 < X < m_hi ** 3


def test_cbrt_exact_is_correctly_rounded(oracle_mod):
    rng = np.random.default_rng(7)
    xs = list(np.exp(rng.uniform(-600, 600, 1500))) + list(rng.uniform(1e-8, 1e-1, 1500))
    xs += [1.0, 2.0, 7.999999999999999, 8.0, 1e-300, 5e-324, 1.7976931348623157e308, 0.125]
    for x in xs:
        y = oracle_mod.cbrt_exact(x)
        if Fraction(y) ** 3 == Fraction(x):
            continue   # exact cube
        assert _is_nearest(x, y), (x, y)


def test_cbrt_exact_sign_and_specials(oracle_mod):
    assert oracle_mod.cbrt_exact(-27.0) == -3.0
    assert oracle_mod.cbrt_exact(0.0) == 0.0
    assert np.signbit(oracle_mod.cbrt_exact(-0.0))
    assert oracle_mod.cbrt_exact(float("inf")) == float("inf")
    assert np.isnan(oracle_mod.cbrt_exact(float("nan")))


def test_libm_cbrt_within_one_ulp(oracle_mod):
    # what separates libm mode from exact mode: glibc's cbrt is faithful
    # (<= 1 ulp), not always correctly rounded
    rng = np.random.default_rng(3)
    xs = rng.uniform(1e-7, 1e-2, 20000)
    ex = np.array([oracle_mod.cbrt_exact(v) for v in xs])
    lm = np.cbrt(xs)
    ulps = np.abs(ex.view(np.int64) - lm.view(np.int64))
    assert ulps.max() <= 1
    assert ulps.mean() < 0.5


@pytest.mark.parametrize("mode", [0, 1])
def test_exact_and_libm_runs_agree_closely(oracle_mod, mode):
    ref = oracle_mod.Oracle(16, cbrt_mode=0)
    ref.run()
    o = oracle_mod.Oracle(16, cbrt_mode=mode)
    o.run()
    e0, e1 = ref.field("e")[0], o.field("e")[0]
    assert abs(e0 - e1) / abs(e0) < 1e-12
