"""Pins of the oracle's grid-side functions.

* sphere tables (utils.c:590-696): the point set equals a brute-force scan
  of the whole grid for points within (n-1)/n*rmax of the site's nearest
  image; each value equals scipy spline x scipy Y_lm of that vector;
* fft3d (linalg.c:14-45): equals numpy's inverse FFT times N/sqrt(V) and
  an explicit plane-wave sum, on synthetic data and on the reference's
  test_files/WAVECAR (its own fft_check, tests.c:28-80);
* sphere sums (projector.c:223-274): a brute-force numpy evaluation over
  the whole grid, and the Rayleigh closed form for one plane wave.
"""

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import CubicSpline
from scipy.special import spherical_jn, sph_harm_y

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


def spline(sp, ch):
    x, y = sp.projgrid, sp.realprojs[ch]
    return CubicSpline(x, y, bc_type=((1, (y[1] - y[0]) / (x[1] - x[0])), (2, 0.0)))


def brute_site_points(lattice, fftg, coord, R0):
    n0, n1, n2 = fftg
    f = np.stack(np.meshgrid(np.arange(n0) / n0, np.arange(n1) / n1, np.arange(n2) / n2, indexing="ij"),
                 axis=-1).reshape(-1, 3)
    d = f - coord[None, :]
    d -= np.round(d)
    cart = d @ lattice
    r = np.linalg.norm(cart, axis=1)
    sel = np.nonzero(r < R0)[0]
    return sel, cart[sel], r[sel]


def angles(v, r):
    th = np.arccos(np.clip(v[:, 2] / np.where(r > 0, r, 1), -1, 1))
    ph = np.arctan2(v[:, 1], v[:, 0]) % (2 * np.pi)
    return th, ph


@pytest.mark.parametrize("ncell,fftg", [(1, (24, 24, 24)), (1, (30, 20, 25))])
def test_site_tables_match_brute_force(ncell, fftg):
    wl = synthetic.tiny_config(ncell=ncell, fftg=fftg, ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=1.2)
    sp = wl.species
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, sp)
    R0 = (sp.npts - 1) * sp.rmax / sp.npts
    for s in range(len(wl.frac_S)):
        sel, cart, r = brute_site_points(wl.lattice, wl.fftg, wl.frac_S[s], R0)
        a, b = T.site_pt[s], T.site_pt[s + 1]
        got_idx = T.idx[a:b]
        order = np.argsort(got_idx)
        assert np.array_equal(np.sort(got_idx), np.sort(sel))
        # path = unwrapped site->point vector; equals the minimum image here
        brute = dict(zip(sel, cart))
        np.testing.assert_allclose(T.path[a:b], np.array([brute[i] for i in got_idx]), atol=1e-12)
        # values: spline(r) * Y_lm(direction), projector order (channel, m)
        n = b - a
        vals = T.values[T.val_off[s]:T.val_off[s] + T.nproj[s] * n].reshape(T.nproj[s], n)
        path = T.path[a:b]
        rr = np.linalg.norm(path, axis=1)
        th, ph = angles(path, rr)
        p = 0
        for ch, l in enumerate(sp.ls):
            rad = spline(sp, ch)(rr)
            for m in range(-l, l + 1):
                want = rad * sph_harm_y(l, m, th, ph)
                np.testing.assert_allclose(vals[p], want, rtol=1e-10, atol=1e-12)
                p += 1
        assert order is not None


def scatter(G, C, fftg):
    x = np.zeros(tuple(fftg), np.complex128)
    g = np.mod(G, np.array(fftg)[None, :])
    x[g[:, 0], g[:, 1], g[:, 2]] = C
    return x


def test_fft3d_matches_numpy_and_explicit_sum():
    wl = synthetic.tiny_config(encut=120.0, fftg=(18, 20, 15))
    G = wl.ensure_gvecs()[0]
    C = wl.coefficients("R", nband=1)[0][0]
    got = orc.fft3d(wl.lattice, G, C, wl.fftg)
    vol = abs(np.linalg.det(wl.lattice))
    want = np.fft.ifftn(scatter(G, C.astype(np.complex128), wl.fftg)) * np.prod(wl.fftg) / np.sqrt(vol)
    np.testing.assert_allclose(got, want, atol=1e-13 * np.abs(want).max())
    # explicit sum at a few points
    rng = np.random.default_rng(0)
    for _ in range(20):
        i, j, k = (rng.integers(n) for n in wl.fftg)
        f = np.array([i / wl.fftg[0], j / wl.fftg[1], k / wl.fftg[2]])
        v = np.sum(C.astype(np.complex128) * np.exp(2j * np.pi * (G @ f))) / np.sqrt(vol)
        assert abs(got[i, j, k] - v) < 1e-12


def test_fft3d_on_reference_wavecar(wavecar):
    """The reference's fft_check (tests.c:28-80, 1e-5 abs) on test_files/WAVECAR."""
    got = orc.fft3d(wavecar["lattice"], wavecar["G_0"], wavecar["C_0"][0], wavecar["grid_fftg"])
    want = wavecar["grid_k0_b0"]
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    assert np.abs(got - want).max() < 1e-5


def test_ifft3d_matches_numpy():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((12, 10, 9)) + 1j * rng.standard_normal((12, 10, 9))
    np.testing.assert_allclose(orc.ifft3d(x), np.fft.ifftn(x) * x.size, atol=1e-11)


def brute_onto_projector(wl, frac, G, C, kfrac):
    """Whole-grid evaluation of <p|psi> independent of the oracle's tables."""
    sp = wl.species
    vol = abs(np.linalg.det(wl.lattice))
    psi = np.fft.ifftn(scatter(G, C.astype(np.complex128), wl.fftg)) * np.prod(wl.fftg) / np.sqrt(vol)
    psi = psi.ravel()
    dv = vol / np.prod(wl.fftg)
    kc = np.asarray(kfrac) @ synthetic.reciprocal_lattice(wl.lattice)
    R0 = (sp.npts - 1) * sp.rmax / sp.npts
    out = []
    for s in range(len(frac)):
        sel, cart, r = brute_site_points(wl.lattice, wl.fftg, frac[s], R0)
        th, ph = angles(cart, r)
        x = psi[sel] * dv * np.exp(1j * cart @ kc)
        for ch, l in enumerate(sp.ls):
            rad = spline(sp, ch)(r)
            for m in range(-l, l + 1):
                p = rad * sph_harm_y(l, m, th, ph)
                out.append(np.sum(np.conj(p) * x))
    return np.array(out)


@pytest.mark.parametrize("kfrac", [(0.0, 0.0, 0.0), (0.25, -0.25, 0.0), (0.1, 0.3, -0.2)])
def test_onto_projector_matches_brute_force(kfrac):
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), kpts=[kfrac], ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    G = wl.ensure_gvecs()[0]
    C = wl.coefficients("S", nband=1)[0][0]
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, wl.species)
    grid = orc.fft3d(wl.lattice, G, C, wl.fftg)
    got = orc.onto_projector(grid, wl.lattice, kfrac, wl.fftg, T)
    want = brute_onto_projector(wl, wl.frac_S, G, C, kfrac)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


def test_onto_projector_single_plane_wave_closed_form():
    """Rayleigh expansion: one plane wave q = k+G, projector f(r) Y_lm,
    <p|psi> = e^{iG.R}/sqrt(V) 4 pi i^l conj(Y_lm(q^)) int r^2 f(r) j_l(qr) dr
    (exact up to the grid quadrature of the sphere sum)."""
    lattice = np.eye(3) * 4.0
    fftg = np.array([64, 64, 64], np.int32)
    sp = synthetic.make_species([0, 1, 2], [2.0, 2.0, 2.0], 1.4)
    frac = np.array([[0.31, 0.42, 0.27]])
    kfrac = np.array([0.2, -0.1, 0.15])
    Gv = np.array([[1, -2, 1]], np.int32)
    T = orc.site_tables(lattice, fftg, frac, sp)
    grid = orc.fft3d(lattice, Gv, np.array([1.0 + 0j], np.complex64), fftg)
    got = orc.onto_projector(grid, lattice, kfrac, fftg, T)
    rec = synthetic.reciprocal_lattice(lattice)
    q = (Gv[0] + kfrac) @ rec
    qn = np.linalg.norm(q)
    thq = np.arccos(q[2] / qn)
    phq = np.arctan2(q[1], q[0]) % (2 * np.pi)
    R = frac[0] @ lattice
    vol = 64.0
    R0 = (sp.npts - 1) * sp.rmax / sp.npts
    want = []
    for ch, l in enumerate(sp.ls):
        f = spline(sp, ch)
        radial = quad(lambda r: r * r * f(r) * spherical_jn(l, qn * r), 0, R0, limit=200)[0]
        for m in range(-l, l + 1):
            want.append(np.exp(1j * (Gv[0] @ rec) @ R) / np.sqrt(vol) * 4 * np.pi * (1j ** l)
                        * np.conj(sph_harm_y(l, m, thq, phq)) * radial)
    want = np.array(want)
    assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()
