"""Pins the oracle's magnitude preparation (oracle.prepare_magnitudes, which
restates periodfind/_utils.py:12-47) without trusting numpy's float32 path:

- closed forms on the reference's own unit-test input [1, 2, 3]
  (tests/test_periodfind.py:64-89): normalised to 5e-4 .. 0.9995, centred to
  exactly [-1, 0, 1];
- random curves against exact rational arithmetic (fractions.Fraction) with an
  explicit round-to-nearest-even to float32 after every operation, which fixes
  both the operation order and that every step is a float32 operation (a
  float64 intermediate, a dropped term or a swapped operand fails it);
- the special cases numpy defines: a constant curve and a curve holding a NaN
  normalise to all NaN; with an infinite magnitude the span is infinite, so
  the finite points normalise to 5e-4 and the infinite one to NaN.
"""

from fractions import Fraction
import math
import warnings

import numpy as np
import pytest

from oracle import oracle as O


def rn32(x):
    """Round an exact rational to the nearest float32 (ties to even), as a Fraction."""
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    e = max(e, -126)  # subnormal spacing below 2^-126
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(a, ulp)
    if r > ulp / 2 or (r == ulp / 2 and q % 2 == 1):
        q += 1
    return sign * q * ulp


def f32(x):
    return Fraction(float(np.float32(x)))


def test_rn32_matches_known_roundings():
    # The helper itself: exact values stay, halfway cases go to even.
    assert rn32(Fraction(1)) == 1
    assert rn32(Fraction(1) + Fraction(1, 2 ** 24)) == 1  # tie -> even (1.0)
    assert rn32(Fraction(1) + Fraction(3, 2 ** 24)) == 1 + Fraction(1, 2 ** 22)
    assert rn32(Fraction(1, 3)) == Fraction(float(np.float32(1.0 / 3.0)))
    assert rn32(Fraction(999, 1000)) == Fraction(float(np.float32(0.999)))


def test_normalize_closed_form_123():
    r = O.prepare_magnitudes([np.array([1.0, 2.0, 3.0], np.float32)], False, True)[0]
    assert r.dtype == np.float32
    c999, c5 = f32(0.999), f32(5e-4)
    # (m - 1) / 2 is exactly 0, 1/2, 1; each later step rounds to float32.
    want = [rn32(rn32(Fraction(k, 2) * c999) + c5) for k in (0, 1, 2)]
    assert [Fraction(float(v)) for v in r] == want
    assert r[0] == np.float32(5e-4)
    assert 0.0 <= r.min() and r.max() <= 1.0  # the reference's own KAT


def test_center_closed_form_123():
    r = O.prepare_magnitudes([np.array([1.0, 2.0, 3.0], np.float32)], True, False)[0]
    assert r.dtype == np.float32
    np.testing.assert_array_equal(r, np.array([-1.0, 0.0, 1.0], np.float32))
    assert abs(np.mean(r)) < 1e-6  # the reference's own KAT


def test_passthrough_is_identity():
    m = [np.array([1.0, 2.0, 3.0], np.float32)]
    assert O.prepare_magnitudes(m, False, False)[0] is m[0]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_normalize_exact_rationals(seed):
    rng = np.random.default_rng(seed)
    m = (15.0 + rng.normal(0, 0.3, 37)).astype(np.float32)
    r = O.prepare_magnitudes([m], False, True)[0]
    assert r.dtype == np.float32
    mf = [Fraction(float(v)) for v in m]
    lo, hi = min(mf), max(mf)
    span = rn32(hi - lo)
    c999, c5 = f32(0.999), f32(5e-4)
    for i, v in enumerate(mf):
        want = rn32(rn32(rn32(rn32(v - lo) / span) * c999) + c5)
        assert Fraction(float(r[i])) == want, i


def test_center_exact_rationals():
    # Sums of these dyadic values are exact in float32, so the mean is the
    # exact mean rounded once, whatever the summation order.
    rng = np.random.default_rng(7)
    for n in (3, 7, 8, 9, 40, 129):
        m = (rng.integers(-4000, 4000, n) / 64.0 + 12.0).astype(np.float32)
        r = O.prepare_magnitudes([m], True, False)[0]
        assert r.dtype == np.float32
        mf = [Fraction(float(v)) for v in m]
        mean = rn32(sum(mf) / n)
        for i, v in enumerate(mf):
            assert Fraction(float(r[i])) == rn32(v - mean), (n, i)


def test_special_curves():
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        const = O.prepare_magnitudes([np.full(5, 3.0, np.float32)], False, True)[0]
        nanc = O.prepare_magnitudes([np.array([1.0, np.nan, 2.0], np.float32)], False, True)[0]
        infc = O.prepare_magnitudes([np.array([1.0, np.inf, 2.0], np.float32)], False, True)[0]
        cen = O.prepare_magnitudes([np.array([1.0, np.nan, 2.0], np.float32)], True, False)[0]
    assert np.all(np.isnan(const))  # 0 / 0
    assert np.all(np.isnan(nanc))   # min/max propagate NaN
    # (x - 1) / inf is 0 for finite x and NaN for x = inf
    np.testing.assert_array_equal(infc, np.array([5e-4, np.nan, 5e-4], np.float32))
    assert np.all(np.isnan(cen))    # the mean is NaN
