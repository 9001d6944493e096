"""Pins of the oracle's radial and angular building blocks.

* spline (utils.c:699-749): equals scipy's CubicSpline with the
  reference's boundary conditions, y'(x0) = secant slope, y''(end) = 0;
* interpolation (utils.c:461-475): scipy evaluation inside the grid, 0
  beyond it;
* spline_integral (utils.c:751-764): scipy's exact integral of the spline;
* Legendre / Y_lm (utils.c:377-449): scipy.special.lpmv / sph_harm_y
  (the reference's own test_Ylm and test_legendre pin the same way);
* PAW partial-wave overlaps (projector.c:507-558): scipy spline integrals.
"""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import lpmv, sph_harm_y

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


def ref_spline(x, y):
    slope = (y[1] - y[0]) / (x[1] - x[0])
    return CubicSpline(x, y, bc_type=((1, slope), (2, 0.0)))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spline_coefficients_match_scipy(seed):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0, 3, 40))
    y = np.sin(2 * x) + 0.1 * rng.standard_normal(40)
    b, c, d = orc.spline_coeff(x, y)
    cs = ref_spline(x, y)
    # scipy: c[0]=d, c[1]=c, c[2]=b per interval
    np.testing.assert_allclose(b[:-1], cs.c[2], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(c[:-1], cs.c[1], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(d[:-1], cs.c[0], rtol=1e-9, atol=1e-8)


def test_natural_start_would_fail():
    """The start condition matters: a natural-start spline differs visibly."""
    sp = synthetic.make_species([1], [1.5], 1.2)
    x, f = sp.projgrid, sp.realprojs[0]
    b, c, d = orc.spline_coeff(x, f)
    nat = CubicSpline(x, f, bc_type="natural")
    assert abs(nat.c[2][0] - b[0]) > 1e-4


def test_projector_interpolation_matches_scipy():
    sp = synthetic.make_species([0, 1, 2], [1.5, 3.0, 1.5], 1.6)
    for f in sp.realprojs:
        x = sp.projgrid
        coeffs = orc.spline_coeff(x, f)
        cs = ref_spline(x, f)
        rs = np.linspace(0.0, x[-1], 777)
        got = np.array([orc.proj_interpolate(r, sp.rmax, x, f, coeffs) for r in rs])
        np.testing.assert_allclose(got, cs(rs), rtol=1e-10, atol=1e-12)
        assert orc.proj_interpolate(x[-1] + 1e-9, sp.rmax, x, f, coeffs) == 0.0


def test_spline_integral_matches_scipy():
    sp = synthetic.make_species([0, 1], [1.5, 3.0], 1.2)
    g = sp.grid
    for a in sp.aewaves:
        y = a * a
        b, c, d = orc.spline_coeff(g, y)
        got = orc.spline_integral(g, y, b, c, d)
        want = ref_spline(g, y).integrate(g[0], g[-1])
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_legendre_matches_scipy():
    xs = np.linspace(-1, 1, 41)
    for l in range(4):
        for m in range(-l, l + 1):
            got = np.array([orc.legendre(l, m, x) for x in xs])
            np.testing.assert_allclose(got, lpmv(m, l, xs), rtol=1e-12, atol=1e-12)


def test_ylm_matches_scipy():
    rng = np.random.default_rng(5)
    th = rng.uniform(0, np.pi, 50)
    ph = rng.uniform(0, 2 * np.pi, 50)
    for l in range(4):
        for m in range(-l, l + 1):
            got = np.array([orc.ylm(l, m, t, p) for t, p in zip(th, ph)])
            np.testing.assert_allclose(got, sph_harm_y(l, m, th, ph), rtol=1e-12, atol=1e-12)


def test_pwave_overlaps_match_scipy():
    sp = synthetic.make_species([0, 0, 1, 1], [1.5, 3.0, 1.5, 3.0], 1.2)
    psov, aeov, diov = orc.pwave_overlaps(sp.grid, sp.ls, sp.aewaves, sp.pswaves)
    g = sp.grid
    for i in range(4):
        for j in range(4):
            if sp.ls[i] != sp.ls[j]:
                assert psov[i, j] == aeov[i, j] == diov[i, j] == 0.0
                continue
            ps = ref_spline(g, sp.pswaves[i] * sp.pswaves[j]).integrate(g[0], g[-1])
            ae = ref_spline(g, sp.aewaves[i] * sp.aewaves[j]).integrate(g[0], g[-1])
            di = ref_spline(g, (sp.aewaves[i] - sp.pswaves[i]) * (sp.aewaves[j] - sp.pswaves[j])).integrate(
                g[0], g[-1])
            assert abs(psov[i, j] - ps) < 1e-12 * max(1, abs(ps))
            assert abs(aeov[i, j] - ae) < 1e-12 * max(1, abs(ae))
            assert abs(diov[i, j] - di) < 1e-12 * max(1, abs(di))
