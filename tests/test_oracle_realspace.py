"""Pins of the oracle's all-electron real-space state (density.c:232-327).

* wave_value2 (utils.c:522-545): equals scipy's CubicSpline with the
  reference's end conditions (first interval clamped to the secant slope,
  natural end) divided by r, times scipy's Y_lm, on the log grid of a
  synthetic species, and the exact constant for linear data f = a r;
* realspace_state: a single plane wave is e^{2 pi i (G + k).r} / sqrt(V);
  the full state equals numpy's inverse FFT times N^3 / sqrt(V), times the
  Bloch phase, plus a brute-force sum over every site of the partial-wave
  differences at the minimum-image vector with phase e^{2 pi i k.(r - d)}
  (d the fractional site-to-point vector), on a cubic cell with
  overlapping spheres and on a skewed cell with sites near the faces;
* remove_phase: the same state times e^{-2 pi i k.r}.
"""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import sph_harm_y

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


def log_spline(x, y):
    return CubicSpline(x, y, bc_type=((1, (y[1] - y[0]) / (x[1] - x[0])), (2, 0.0)))


def test_wave_value2_matches_scipy():
    sp = synthetic.make_species([0, 1, 2], [1.5, 3.0, 1.5], 1.2)
    rng = np.random.default_rng(3)
    for ch, l in enumerate(sp.ls):
        f = np.asarray(sp.aewaves[ch]) - np.asarray(sp.pswaves[ch])
        cs = log_spline(sp.grid, f)
        coeffs = orc.spline_coeff(sp.grid, f)
        for _ in range(40):
            v = rng.normal(size=3)
            v *= rng.uniform(sp.grid[0], sp.grid[-1]) / np.linalg.norm(v)
            r = np.linalg.norm(v)
            th, ph = np.arccos(v[2] / r), np.arctan2(v[1], v[0]) % (2 * np.pi)
            for m in range(-l, l + 1):
                want = cs(r) / r * sph_harm_y(l, m, th, ph)
                got = orc.wave_value2(sp.grid, f, coeffs, l, m, v)
                assert abs(got - want) < 1e-10 * max(1.0, abs(want)), (l, m, r)
        # below the first grid point: f[0] / x[0]; beyond the last: 0
        v = np.array([0.0, 0.0, sp.grid[0] * 0.5])
        assert abs(orc.wave_value2(sp.grid, f, coeffs, l, 0, v) - f[0] / sp.grid[0] * sph_harm_y(l, 0, 0.0, 0.0)) < 1e-12
        v = np.array([sp.grid[-1] * 1.01, 0, 0])
        assert orc.wave_value2(sp.grid, f, coeffs, l, 0, v) == 0


def test_wave_value2_linear_data_exact():
    x = 1.3 * np.exp(np.linspace(np.log(1e-3), 0.0, 200))
    a = 0.7
    f = a * x
    coeffs = orc.spline_coeff(x, f)
    np.testing.assert_allclose(coeffs[0][:-1], a, atol=1e-12)
    for v in ([0.3, -0.2, 0.5], [0.0, 0.0, 0.9], [-0.4, 0.0, 0.0], [1e-5, 0, 0]):
        v = np.asarray(v, float)
        r = np.linalg.norm(v)
        th, ph = np.arccos(v[2] / r), np.arctan2(v[1], v[0]) % (2 * np.pi)
        for l in range(3):
            for m in range(-l, l + 1):
                assert abs(orc.wave_value2(x, f, coeffs, l, m, v) - a * sph_harm_y(l, m, th, ph)) < 1e-12


def test_single_plane_wave_closed_form():
    lat = np.array([[5.0, 0, 0], [1.0, 5.5, 0], [0.5, 0.7, 6.0]])
    fftg = (12, 14, 16)
    k = np.array([0.25, -0.125, 0.3])
    G = np.array([[2, -1, 3]], np.int32)
    C = np.array([1.0 + 0j], np.complex64)
    sp = synthetic.make_species([0], [1.5], 1.0)
    W = orc.WaveData([sp])
    x = orc.realspace_state(lat, fftg, k, G, C, W, np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0))
    f = np.stack(np.meshgrid(*[np.arange(n) / n for n in fftg], indexing="ij"), axis=-1)
    want = np.exp(2j * np.pi * (f @ (G[0] + k))) / np.sqrt(abs(np.linalg.det(lat)))
    np.testing.assert_allclose(x, want, atol=1e-13)
    xr = orc.realspace_state(lat, fftg, k, G, C, W, np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0),
                             remove_phase=True)
    np.testing.assert_allclose(xr, want * np.exp(-2j * np.pi * (f @ k)), atol=1e-13)


def brute_state(lat, fftg, k, G, C, species, labels, coords, proj):
    """numpy restatement: FFT by numpy, augmentation by a scan of the whole
    grid against each site's minimum image."""
    n = np.array(fftg)
    grid = np.zeros(fftg, np.complex128)
    for g, c in zip(G, C):
        grid[tuple(np.mod(g, n))] = c
    vol = abs(np.linalg.det(lat))
    x = np.fft.ifftn(grid) * np.prod(n) / np.sqrt(vol)
    f = np.stack(np.meshgrid(*[np.arange(q) / q for q in fftg], indexing="ij"), axis=-1).reshape(-1, 3)
    x = x.ravel() * np.exp(2j * np.pi * (f @ k))
    p0 = 0
    for s, e in enumerate(labels):
        sp = species[e]
        d = f - coords[s]
        d -= np.round(d)
        cart = d @ lat
        r = np.linalg.norm(cart, axis=1)
        sel = np.nonzero(r < sp.grid[-1])[0]
        v, rs = cart[sel], r[sel]
        th = np.arccos(np.clip(v[:, 2] / rs, -1, 1))
        ph = np.arctan2(v[:, 1], v[:, 0]) % (2 * np.pi)
        phase = np.exp(2j * np.pi * ((f[sel] - d[sel]) @ k))
        p = p0
        for ch, l in enumerate(sp.ls):
            fq = np.asarray(sp.aewaves[ch]) - np.asarray(sp.pswaves[ch])
            rad = log_spline(sp.grid, fq)(rs) / rs
            for m in range(-l, l + 1):
                x[sel] += rad * sph_harm_y(l, m, th, ph) * proj[p] * phase
                p += 1
        p0 = p
    return x.reshape(fftg)


@pytest.mark.parametrize("case", ["cubic_overlapping", "skewed_faces"])
def test_realspace_state_brute_force(case):
    rng = np.random.default_rng(11)
    if case == "cubic_overlapping":
        wl = synthetic.tiny_config(ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=1.1, fftg=(20, 20, 20),
                                   kpts=[[0.25, -0.25, 0.1]])
        lat, fftg, k = wl.lattice, tuple(wl.fftg), np.asarray(wl.kpts[0], float)
        species = [wl.species]
        coords = np.asarray(wl.frac_S)
        labels = np.zeros(len(coords), np.int32)
        G = wl.ensure_gvecs()[0]
        C = wl.coefficients("S", nband=1)[0][0]
    else:
        lat = np.array([[5.0, 0, 0], [1.0, 5.5, 0], [0.5, 0.7, 6.0]])
        fftg, k = (20, 22, 24), np.array([0.3, 0.1, -0.2])
        species = [synthetic.make_species([0, 1], [1.5, 3.0], 1.2), synthetic.make_species([2], [2.0], 0.9)]
        coords = np.array([[0.02, 0.5, 0.97], [0.51, 0.99, 0.03], [0.97, 0.01, 0.5]])
        labels = np.array([0, 1, 0], np.int32)
        G = rng.integers(-5, 6, size=(30, 3)).astype(np.int32)
        G = np.unique(G, axis=0)
        C = (0.1 * (rng.normal(size=len(G)) + 1j * rng.normal(size=len(G)))).astype(np.complex64)
    W = orc.WaveData(species)
    P = sum(W.proj_counts[e] for e in labels)
    proj = rng.normal(size=P) + 1j * rng.normal(size=P)
    got = orc.realspace_state(lat, fftg, k, G, C, W, labels, coords, proj)
    want = brute_state(lat, fftg, k, G, C, species, labels, coords, proj)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # the augmentation is not negligible against the pseudo part
    plain = brute_state(lat, fftg, k, G, C, species, labels, coords, 0 * proj)
    assert np.abs(want - plain).max() > 0.1 * np.abs(plain).max()
    gr = orc.realspace_state(lat, fftg, k, G, C, W, labels, coords, proj, remove_phase=True)
    f = np.stack(np.meshgrid(*[np.arange(q) / q for q in fftg], indexing="ij"), axis=-1)
    assert np.abs(gr - want * np.exp(-2j * np.pi * (f @ k))).max() < 1e-10 * np.abs(want).max()
