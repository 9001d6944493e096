"""The oracle's MKL transforms (or_fft_use_mkl: libmkl_rt's FFTW3 entry
points, loaded at run time, which the benchmark's CPU baseline uses)
against the oracle's own mixed-radix transform, itself pinned to numpy's
FFT in test_oracle_grid.py: the 3D backward transform on grids with factors
2, 3, 5, 7 and the 108^3-type factor 4 x 27, and a whole setup_projections
(scatter, transform, sphere sums) to 1e-12 relative.  Skipped where
libmkl_rt cannot be loaded."""

import numpy as np
import pytest

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


@pytest.fixture()
def mkl():
    if not orc.fft_use_mkl(True):
        pytest.skip("libmkl_rt not loadable")
    orc.fft_use_mkl(False)
    yield
    orc.fft_use_mkl(False)


@pytest.mark.parametrize("fftg,encut", [((24, 20, 30), 250.0), ((28, 14, 21), 60.0), ((54, 27, 36), 300.0)])
def test_mkl_fft3d_matches_oracle(mkl, fftg, encut):
    wl = synthetic.tiny_config(encut=encut, fftg=fftg, nband=2, kpts=[[0.1, 0.2, 0.3]])
    G = wl.ensure_gvecs()[0]
    C = wl.coefficients("R", nband=2)[0]
    own = [orc.fft3d(wl.lattice, G, C[b], wl.fftg) for b in range(2)]
    assert orc.fft_use_mkl(True)
    via = [orc.fft3d(wl.lattice, G, C[b], wl.fftg) for b in range(2)]
    for a, b in zip(own, via):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()


def test_mkl_setup_projections_matches_oracle(mkl):
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=6, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=6)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32))
    own = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T, nthreads=2)
    assert orc.fft_use_mkl(True)
    via = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T, nthreads=2)
    assert np.abs(own - via).max() <= 1e-12 * np.abs(own).max()
