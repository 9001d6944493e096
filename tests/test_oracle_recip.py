"""Pin of the oracle's aug_recip building block (or_aug_freqs:
projector.c:276-330, 420-453; linalg.c:47-79) against an independent numpy
route: np.add.at accumulation of the site functions on the grid, then
numpy's forward FFT scaled by sqrt(V)/N^3, sampled at the G vectors."""

import numpy as np

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


def numpy_aug_freqs(wl, kfrac, G, T, P):
    fftg = tuple(int(n) for n in wl.fftg)
    x = np.zeros(int(np.prod(fftg)), np.complex128)
    rec = synthetic.reciprocal_lattice(wl.lattice)
    kc = np.asarray(kfrac) @ rec
    for s in range(T.nsites):
        a, b = T.site_pt[s], T.site_pt[s + 1]
        n = b - a
        vals = T.values[T.val_off[s]: T.val_off[s] + T.nproj[s] * n].reshape(T.nproj[s], n)
        coef = P[T.proj_off[s]: T.proj_off[s + 1]]
        ph = np.exp(-1j * (T.path[a:b] @ kc))
        np.add.at(x, T.idx[a:b], (coef @ vals) * ph)
    X = np.fft.fftn(x.reshape(fftg)) * np.sqrt(abs(np.linalg.det(wl.lattice))) / np.prod(fftg)
    g = np.mod(G, fftg)
    return X[g[:, 0], g[:, 1], g[:, 2]]


def test_aug_freqs_matches_numpy():
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 20, 18), kpts=[[0.1, -0.2, 0.3]], ls=(0, 1, 2),
                               alphas=(1.5, 3.0, 2.0))
    G = wl.ensure_gvecs()[0]
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S[:5], wl.species)
    rng = np.random.default_rng(4)
    P = rng.standard_normal(T.ptot) + 1j * rng.standard_normal(T.ptot)
    got = orc.aug_freqs(wl.lattice, wl.fftg, wl.kpts[0], G, T, P)
    want = numpy_aug_freqs(wl, wl.kpts[0], G, T, P)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
