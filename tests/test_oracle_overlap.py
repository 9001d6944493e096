"""Pins of the oracle's band-overlap and augmentation routines.

* pseudoprojection (pseudoprojector.c:63-90): numpy vdot per (band, k),
  spin-flip index map, Hermitian symmetry, on synthetic data and on the
  reference's test_files/WAVECAR; the FP32-accumulating variant (the
  reference's cdotc) agrees with FP64 to single-precision error;
* compensation_terms O_M (projector.c:850-910): a numpy block-diagonal
  matrix product; with dO = identity on one channel and P_S = P_R it
  reduces to the sum of |P|^2 over that channel.
"""

import numpy as np
import pytest

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic


def test_pseudoprojection_matches_vdot(wavecar):
    nwk, nspin = int(wavecar["nwk"]), int(wavecar["nspin"])
    C = [wavecar[f"C_{i}"] for i in range(nwk * nspin)]
    for flip in (False, True):
        for band in (0, 3):
            got = orc.pseudoprojection(C, C, band, nwk, nspin, flip=flip)
            nb = C[0].shape[0]
            for b in range(nb):
                for k in range(nwk * nspin):
                    kp = (k + nwk) % (2 * nwk) if flip else k
                    want = np.vdot(C[kp][b].astype(np.complex128), C[k][band].astype(np.complex128))
                    assert abs(got[b * nwk * nspin + k] - want) < 1e-12


def test_pseudoprojection_hermitian_and_fp32():
    wl = synthetic.tiny_config(encut=200.0, nband=6, kpts=[[0, 0, 0], [0.25, 0.25, 0]])
    CR = wl.coefficients("R")
    CS = wl.coefficients("S")
    nk = len(CR)
    for b1 in range(3):
        r1 = orc.pseudoprojection(CR, CS, b1, nk, 1)
        for b2 in range(3):
            r2 = orc.pseudoprojection(CS, CR, b2, nk, 1)
            for k in range(nk):
                assert abs(r1[b2 * nk + k] - np.conj(r2[b1 * nk + k])) < 1e-14
    full = orc.pseudoprojection(CR, CS, 2, nk, 1)
    f32 = orc.pseudoprojection(CR, CS, 2, nk, 1, fp32acc=True)
    err = np.abs(full - f32).max()
    assert 0 < err < 1e-6


def block_dO(ls, dO):
    ns, lls, ms = orc.projector_meta(ls)
    same = (lls[:, None] == lls[None, :]) & (ms[:, None] == ms[None, :])
    return np.where(same, dO[ns[:, None], ns[None, :]], 0.0)


@pytest.mark.parametrize("flip", [False, True])
def test_compensation_m_matches_block_product(flip):
    rng = np.random.default_rng(11)
    ls = [0, 0, 1, 1, 2]
    P = sum(2 * l + 1 for l in ls)
    nsites, nwk, nspin, nbR = 5, 2, 2, 4
    nkt = nwk * nspin
    proj_R = rng.standard_normal((nkt, nbR, nsites * P)) + 1j * rng.standard_normal((nkt, nbR, nsites * P))
    proj_S = rng.standard_normal((nkt, nsites * P)) + 1j * rng.standard_normal((nkt, nsites * P))
    dO = rng.standard_normal((5, 5))
    dO = dO + dO.T

    class T:
        pass

    T.ptot = nsites * P
    T.proj_off = np.arange(nsites + 1, dtype=np.int32) * P
    M_R = np.array([0, 2, 3], np.int32)
    M_S = np.array([1, 2, 4], np.int32)
    base = rng.standard_normal(nbR * nkt) + 1j * rng.standard_normal(nbR * nkt)
    got = orc.compensation_terms_m(base.copy(), 0, nwk, nspin, proj_R, T, np.zeros(nsites, np.int32), proj_S, T,
                                   M_R, M_S, [dO], [ls], flip=flip)
    D = block_dO(ls, dO)
    want = base.copy()
    for k in range(nkt):
        kR = (k + nwk) % nkt if flip else k
        for b in range(nbR):
            acc = 0
            for s1, s2 in zip(M_R, M_S):
                acc += np.conj(proj_R[kR, b, s1 * P:(s1 + 1) * P]) @ D @ proj_S[k, s2 * P:(s2 + 1) * P]
            want[b * nkt + k] += acc
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_compensation_m_identity_reduces_to_norm():
    rng = np.random.default_rng(2)
    ls = [0, 1, 2]
    P = 9
    proj = rng.standard_normal((1, 3, 2 * P)) + 1j * rng.standard_normal((1, 3, 2 * P))
    dO = np.zeros((3, 3))
    dO[2, 2] = 1.0  # only the l=2 channel

    class T:
        pass

    T.ptot = 2 * P
    T.proj_off = np.array([0, P, 2 * P], np.int32)
    for b in range(3):
        got = orc.compensation_terms_m(np.zeros(3, complex), b, 1, 1, proj, T, np.zeros(2, np.int32), proj[:, b],
                                       T, np.array([0, 1], np.int32), np.array([0, 1], np.int32), [dO], [ls])
        want = np.sum(np.abs(proj[0, b, 4:9]) ** 2) + np.sum(np.abs(proj[0, b, P + 4:P + 9]) ** 2)
        assert abs(got[b] - want) < 1e-12
