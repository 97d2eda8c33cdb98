"""Identities the HIP kernels rely on to evaluate the reference's formulas
with fewer operations and the same bits.

quality_from_mean (numeric.cpp:302-305) takes mean = min(perim / 3,
3 / perim); the kernels compute `perim < 3 ? perim / 3 : 3 / perim`, one
division.  IEEE division in numpy float64 rounds like the device's, so the
identity is checked here on the values where a rounding could tell the two
apart: around 3 at ulp spacing, across many magnitudes, and at 0, inf, NaN.
"""
import numpy as np


def two_division_min(perim):
    a = perim / 3
    b = 3 / perim
    return np.where(b < a, b, a)  # dmin(a, b) = (b < a) ? b : a


def one_division(perim):
    return np.where(perim < 3, perim / 3, 3 / perim)


def test_perimeter_mean_one_division_matches_two():
    rng = np.random.default_rng(0)
    three = np.float64(3.0)
    near = [three]
    lo, hi = three, three
    for _ in range(4096):
        lo = np.nextafter(lo, 0.0)
        hi = np.nextafter(hi, np.inf)
        near += [lo, hi]
    wide = np.exp(rng.uniform(-700, 700, 200000))
    mid = rng.uniform(0.5, 6.0, 200000)
    special = np.array([0.0, np.inf, 5e-324, 1.7976931348623157e308, 1.0, 9.0])
    x = np.concatenate([np.array(near), wide, mid, special])
    with np.errstate(divide="ignore", over="ignore", under="ignore"):
        a, b = two_division_min(x), one_division(x)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    with np.errstate(invalid="ignore"):
        assert np.isnan(one_division(np.array([np.nan]))).all()
        assert np.isnan(two_division_min(np.array([np.nan]))).all()
