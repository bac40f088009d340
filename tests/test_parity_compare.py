"""The parity comparator itself: an infinity matches only an infinity of the
same sign, a NaN only a NaN, finite values within the relative tolerance."""
import numpy as np

from tests.parity import close

inf, nan = np.inf, np.nan


def test_infinity_against_finite_fails():
    assert not close([inf], [1.0])[0]
    assert not close([1.0], [inf])[0]
    assert not close([-inf], [1e300])[0]


def test_infinities_of_opposite_sign_fail():
    assert not close([inf], [-inf])[0]


def test_equal_infinities_and_nans_pass():
    assert close([inf, -inf, nan, 2.0], [inf, -inf, nan, 2.0])[0]


def test_nan_against_anything_else_fails():
    assert not close([nan], [0.0])[0]
    assert not close([nan], [inf])[0]
    assert not close([0.0], [nan])[0]


def test_relative_tolerance():
    assert close([1.0 + 5e-10], [1.0])[0]
    assert not close([1.0 + 2e-9], [1.0])[0]
    assert close([0.0], [0.0])[0]
    assert not close([1e-300], [0.0])[0]
    ok, err = close([1.0, inf], [1.0 + 1e-10, inf])
    assert ok and 0 < err < 2e-10


def test_shape_mismatch_fails():
    assert not close([1.0, 2.0], [1.0])[0]
