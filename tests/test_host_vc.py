"""Valence/conduction split of Projector.proportion_conduction and the band
window of defect_band_analysis (projector.py:384-503), host logic only.

The projections come from a stub in place of single_band_projection, so
these run without a GPU.  The expected values come from the oracle's loop
restatement (oracle.pyoracle.proportion_conduction); that restatement is
itself pinned by identities independent of its loops:
- spin-resolved, v[s] + c[s] = sum_b,k |res|^2 w_k for any occupations;
- for occupations 0 or 1, the spin sums of the spin-resolved split equal
  nspin times the unresolved split;
- with the 'pseudo' method v + c = 1.
"""

import types

import numpy as np
import pytest

from oracle import pyoracle as orc
from pawpyseed_amd.core.projector import Projector


def _stub(nband=7, nwk=3, nspin=2, method="aug_real", seed=0, binary=False, nband_wf=None):
    rng = np.random.default_rng(seed)
    n = nband * nwk * nspin
    res = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    occ = rng.uniform(0, 1, n)
    if binary:
        occ = (occ > 0.4).astype(float)
    kws = rng.uniform(0.1, 1.0, nwk)
    nb_wf = nband_wf or nband
    occ_wf = np.zeros(nb_wf * nwk * nspin)
    occ_wf[: 4 * nwk * nspin] = 1.0          # bands 0..3 filled at every (k, spin)
    occ_wf[6 * nwk * nspin] = 1.0            # band 6 filled at its first (k, spin) only
    basis = types.SimpleNamespace(nband=nband, nwk=nwk, nspin=nspin, _get_occs=lambda: occ.copy())
    wf = types.SimpleNamespace(nband=nb_wf, nwk=nwk, nspin=nspin, kws=kws, _get_occs=lambda: occ_wf.copy(),
                               _get_energy_list=lambda bands: {b: [[float(b), 0.0]] for b in bands})
    pr = Projector.__new__(Projector)
    pr.method, pr.basis, pr.wf = method, basis, wf
    pr.single_band_projection = lambda band_num, **kw: res * (1.0 + band_num)
    return pr, res, occ, kws, occ_wf


@pytest.mark.parametrize("spinpol", [False, True])
@pytest.mark.parametrize("method", ["aug_real", "pseudo"])
def test_proportion_conduction_matches_restatement(spinpol, method):
    pr, res, occ, kws, _ = _stub(method=method)
    v, c = pr.proportion_conduction(2, spinpol=spinpol)
    wv, wc = orc.proportion_conduction(res * 3.0, occ, kws, 7, 3, 2, spinpol=spinpol, pseudo=method == "pseudo")
    assert np.allclose(v, wv, rtol=1e-12, atol=0) and np.allclose(c, wc, rtol=1e-12, atol=0)
    if spinpol:
        assert isinstance(v, list) and len(v) == 2
    if method == "pseudo":
        assert np.allclose(np.asarray(v) + np.asarray(c), 1.0, rtol=1e-13)


def test_restatement_identities():
    nband, nwk, nspin = 7, 3, 2
    _, res, occ, kws, _ = _stub(binary=True, seed=4)
    v, c = orc.proportion_conduction(res, occ, kws, nband, nwk, nspin, spinpol=True)
    a = np.abs(res.reshape(nband, nspin, nwk)) ** 2 * kws[None, None, :]
    assert np.allclose(np.asarray(v) + np.asarray(c), a.sum(axis=(0, 2)), rtol=1e-13)
    v1, c1 = orc.proportion_conduction(res, occ, kws, nband, nwk, nspin, spinpol=False)
    assert abs(sum(v) - nspin * v1) < 1e-12 * abs(v1) and abs(sum(c) - nspin * c1) < 1e-12 * abs(c1)
    # fractional occupations: the unresolved split thresholds at 0.5
    _, res, occ, kws, _ = _stub(seed=5)
    v, c = orc.proportion_conduction(res, occ, kws, nband, nwk, nspin, spinpol=True)
    v1, c1 = orc.proportion_conduction(res, occ, kws, nband, nwk, nspin, spinpol=False)
    assert abs(sum(v) - nspin * v1) > 1e-3 * abs(v1)


@pytest.mark.parametrize("kw", [dict(), dict(num_below_ef=2, num_above_ef=1), dict(vbmband=3, num_below_ef=10),
                                dict(band_list=[5, 1]), dict(analyze_all=True), dict(num_above_ef=50)])
def test_defect_band_window(kw):
    pr, res, occ, kws, occ_wf = _stub(nband=9, nband_wf=8)
    got = pr.defect_band_analysis(**kw)
    args = dict(num_below_ef=kw.get("num_below_ef", 20), num_above_ef=kw.get("num_above_ef", 20),
                vbmband=kw.get("vbmband"), band_list=kw.get("band_list"), analyze_all=kw.get("analyze_all", False))
    want = orc.analysis_bands(occ_wf, 8, nband_basis=9, **args)
    assert list(got) == want
    for b in want:
        assert got[b] == pr.proportion_conduction(b)
    # the VBM is the last band filled at the first (k, spin): band 6 here
    if not kw:
        assert want == list(range(0, 8))
    res2, en = pr.defect_band_analysis(return_energies=True, band_list=[2])
    assert list(res2) == [2] and en == {2: [[2.0, 0.0]]}
    with pytest.raises(ValueError):
        pr.defect_band_analysis(num_below_ef=-1)
