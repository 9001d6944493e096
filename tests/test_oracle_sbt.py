"""Pins of the oracle's spherical Bessel transform (sbt.c:24-213) and the
filtered partial-wave difference built from it (projector.c:114-163).

* forward: g(k) = int r^2 j_l(k r) R(r) dr from f = r R(r); for
  R = r^l exp(-a r^2) the closed form is
  sqrt(pi) k^l exp(-k^2 / 4a) / (2^(l+2) a^(l+3/2));
* the reference's own test_sbt (test_core.py:130-146): g(k) against a
  trapezoid integral of j_l(k r) f r at one k, to 3 decimals;
* inverse: R(r) = (2/pi) int k^2 j_l(k r) g(k) dk, checked on the closed
  form within the algorithm's own error (its k grid starts at
  kmin = sqrt(1e7 c) e^{-(n-1) drho}, so the low-k tail is cut);
* smooth_diffwave: a partial-wave difference whose transform lies well
  below the cutoff comes back unchanged on the sphere grid; a cutoff below
  its transform removes most of it; sdw[0] follows projector.c:143-151.
"""

import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import pyoracle as orc


def log_grid(n=400, r0=1e-3, r1=3.0):
    return r0 * np.exp(np.arange(n) * np.log(r1 / r0) / (n - 1))


def gauss_sbt(k, l, a):
    return np.sqrt(np.pi) * k ** l * np.exp(-k * k / (4 * a)) / (2 ** (l + 2) * a ** (l + 1.5))


@pytest.mark.parametrize("l", [0, 1, 2, 3])
@pytest.mark.parametrize("a,tol", [(4.0, 1e-8), (1.5, 2e-4)])
def test_forward_matches_closed_form(l, a, tol):
    r = log_grid()
    R = r ** l * np.exp(-a * r * r)
    ks, g = orc.sbt_forward(r, r * R, l, lmax=3)
    m = (ks > 0.6) & (ks < 15.0)
    ex = gauss_sbt(ks, l, a)
    assert np.abs(g[m] - ex[m]).max() / np.abs(ex[m]).max() < tol


def test_forward_matches_reference_trapezoid_check():
    """test_core.py:130-146 on a smooth synthetic partial-wave difference."""
    r = log_grid(500, 1e-4, 2.5)
    f = r * (np.exp(-3.0 * r * r) - 0.5 * np.exp(-6.0 * r * r)) * (1 - (r / 2.5) ** 2) ** 3
    ks, g = orc.sbt_forward(r, f, 0, lmax=0, encut=1e6)
    for i in (120, 180, 240):
        integral = np.trapezoid(spherical_jn(0, r * ks[i]) * f * r, r)
        assert abs(integral - g[i]) < 5e-4


@pytest.mark.parametrize("l,a,tol", [(0, 4.0, 3e-3), (0, 8.0, 1.5e-3), (1, 4.0, 1e-4), (2, 4.0, 3e-5),
                                     (2, 8.0, 5e-6)])
def test_inverse_matches_closed_form(l, a, tol):
    r = log_grid()
    ks = orc.sbt_forward(r, r, l, lmax=2)[0]
    rs, R = orc.sbt_inverse(r, gauss_sbt(ks, l, a), l, lmax=2)
    np.testing.assert_allclose(rs, r, rtol=1e-12)
    m = (rs > 0.1) & (rs < 1.0)
    ex = rs ** l * np.exp(-a * rs * rs)
    assert np.abs(R[m] - ex[m]).max() / np.abs(ex[m]).max() < tol


def test_inverse_of_forward_round_trip():
    r = log_grid()
    R = r * np.exp(-5.0 * r * r)
    ks, g = orc.sbt_forward(r, r * R, 1, lmax=1)
    rs, Rb = orc.sbt_inverse(r, g, 1, lmax=1)
    m = (rs > 0.1) & (rs < 1.0)
    assert np.abs(Rb[m] - R[m]).max() / np.abs(R).max() < 1e-4


@pytest.mark.parametrize("l,tol", [(0, 3e-3), (1, 1e-3), (2, 1e-3)])
def test_smooth_diffwave_below_cutoff_is_unchanged(l, tol):
    """Transform of r^l exp(-8 r^2) is ~e^{-k^2/32}: negligible above
    k = sqrt(c * 1e4 eV) = 51 / A, so the filter keeps the function."""
    r = log_grid(400, 1e-3, 2.0)
    R = r ** l * np.exp(-8.0 * r * r)
    npj = 100
    sdw, kg, kw = orc.smooth_diffwave(r, r * R, l, 2, 1e4, npj)
    x = 2.0 / npj * np.arange(npj)
    ex = x ** l * np.exp(-8.0 * x * x)
    # within the inverse transform's own error (l = 0 carries the constant
    # offset of the cut low-k tail), away from r = 0
    assert np.abs(sdw[5:] - ex[5:]).max() < tol
    if l == 0:
        assert sdw[0] == sdw[1]
    else:
        assert sdw[0] == 0.0


def test_smooth_diffwave_cutoff_removes_high_k():
    """A narrow function (transform width ~ 2 sqrt(a) = 20 / A) filtered at
    k < sqrt(c * 20 eV) = 2.3 / A loses most of its amplitude."""
    r = log_grid(400, 1e-3, 2.0)
    R = np.exp(-100.0 * r * r)
    sdw, kg, kw = orc.smooth_diffwave(r, r * R, 0, 0, 20.0, 100)
    assert np.abs(sdw).max() < 0.05
