"""Pins of the oracle's off-site partial-wave overlap (radial.c:116-196,
gaunt.py:17-27) and the library's host implementation against it.

* Wigner 3j symbols equal sympy's exact values;
* two s-type Gaussians: closed form (1/4pi) (pi/(a+b))^(3/2) e^{-ab d^2/(a+b)};
* l <= 2 pairs: brute-force 3D quadrature of int phi_1 conj(phi_2(r - d)) d^3r
  (a sign, index or conjugation mistake fails these);
* the library (pp_offsite_overlap, pawpyc.pyx:151-190 form) equals the
  oracle given the same radial inputs.
"""

import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import pyoracle as orc

sympy = pytest.importorskip("sympy")


def test_wigner3j_matches_sympy():
    from sympy.physics.wigner import wigner_3j
    worst = 0.0
    for j1 in range(4):
        for j2 in range(4):
            for j3 in range(abs(j1 - j2), j1 + j2 + 1):
                for m1 in range(-j1, j1 + 1):
                    for m2 in range(-j2, j2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) > j3:
                            continue
                        worst = max(worst, abs(orc.wigner3j(j1, j2, j3, m1, m2, m3) -
                                               float(wigner_3j(j1, j2, j3, m1, m2, m3))))
    assert worst < 1e-14


def kgrid(n=600, k0=1e-3, k1=60.0):
    return k0 * np.exp(np.arange(n) * np.log(k1 / k0) / (n - 1))


def gauss_sbt(k, l, a):
    return np.sqrt(np.pi) * k ** l * np.exp(-k * k / (4 * a)) / (2 ** (l + 2) * a ** (l + 1.5))


@pytest.mark.parametrize("d", [[0.3, 0.2, -0.1], [0.0, 0.0, 0.0], [1.0, 0.5, 0.2]])
def test_ss_gaussians_closed_form(d):
    a, b = 3.0, 5.0
    k = kgrid()
    d = np.array(d)
    got = orc.offsite_overlap(d, k, gauss_sbt(k, 0, a), k, gauss_sbt(k, 0, b), 0, 0, 0, 0)
    ex = (np.pi / (a + b)) ** 1.5 * np.exp(-a * b * d @ d / (a + b)) / (4 * np.pi)
    assert abs(got - ex) < 1e-7 * ex


def _ylm(l, m, x, y, z):
    r = np.sqrt(x * x + y * y + z * z) + 1e-300
    return sph_harm_y(l, m, np.arccos(np.clip(z / r, -1, 1)), np.arctan2(y, x))


@pytest.fixture(scope="module")
def box():
    h = 0.05
    xs = np.arange(-3.0, 3.0 + 1e-9, h)
    return h, np.meshgrid(xs, xs, xs, indexing="ij")


@pytest.mark.parametrize("l1,m1,l2,m2", [(1, 0, 0, 0), (0, 0, 1, 1), (1, -1, 1, 0), (2, 1, 1, 0), (1, 1, 2, -1),
                                         (2, 2, 2, 1)])
def test_brute_force_quadrature(box, l1, m1, l2, m2):
    h, (X, Y, Z) = box
    a, b = 3.0, 4.0
    d = np.array([0.35, -0.2, 0.25])
    R1 = np.sqrt(X * X + Y * Y + Z * Z)
    phi1 = R1 ** l1 * np.exp(-a * R1 * R1) * _ylm(l1, m1, X, Y, Z)
    Xd, Yd, Zd = X - d[0], Y - d[1], Z - d[2]
    R2 = np.sqrt(Xd * Xd + Yd * Yd + Zd * Zd)
    phi2 = R2 ** l2 * np.exp(-b * R2 * R2) * _ylm(l2, m2, Xd, Yd, Zd)
    want = np.sum(phi1 * np.conj(phi2)) * h ** 3
    k = kgrid()
    got = orc.offsite_overlap(d, k, gauss_sbt(k, l1, a), k, gauss_sbt(k, l2, b), l1, m1, l2, m2)
    assert abs(got - want) < 2e-5 * max(abs(want), 1e-3)


@pytest.mark.parametrize("l1,m1,l2,m2", [(0, 0, 0, 0), (1, -1, 2, 1), (2, 2, 1, 0), (3, -2, 1, 1)])
def test_library_matches_oracle(l1, m1, l2, m2):
    from pawpyseed_amd.core.radial import reciprocal_offsite_wave_overlap, spherical_bessel_transform
    r = 1e-4 * np.exp(np.arange(500) * np.log(2.0 / 1e-4) / 499)
    f1 = r ** (l1 + 1) * np.exp(-2.5 * r * r) * (1 - (r / 2.0) ** 2) ** 3
    f2 = r ** (l2 + 1) * np.exp(-3.5 * r * r) * (1 - (r / 2.0) ** 2) ** 3
    d = np.array([0.4, 0.1, -0.7])
    got = reciprocal_offsite_wave_overlap(d, r, f1, r, f2, l1, m1, l2, m2)
    k1, g1 = orc.sbt_forward(r, f1, l1, lmax=l1, encut=1e5)
    k2, g2 = orc.sbt_forward(r, f2, l2, lmax=l2, encut=1e5)
    kl, gl = spherical_bessel_transform(1e5, l1, r, f1)
    np.testing.assert_allclose(kl, k1, rtol=1e-14)
    assert np.abs(gl - g1).max() < 1e-11 * np.abs(g1).max()
    want = orc.offsite_overlap(d, k1, g1, k2, g2, l1, m1, l2, m2)
    assert abs(got - want) < 1e-10 * max(abs(want), 1e-6)
