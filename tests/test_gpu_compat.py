"""End-to-end parity through the reference-named entry points
(setup_projections, pseudoprojection, compensation_terms, fft3d) and the
reference's Python API (Wavefunction, Projector.single_band_projection),
against the oracle's restatement of the same path:

res = pseudoprojection (pseudoprojector.c:63-90)
      + compensation_terms O_M, O_R, O_S, O_N (projector.c:850-963),
with <p|psi> from setup_projections (projector.c:560-602) and the
wave projections / off-site overlaps of overlap_setup_real
(projector.c:604-720).
Tolerance: 1e-9 relative (normwise) for the FP64 path.
"""

import ctypes as ct

import numpy as np
import pytest

from helpers import normwise, oracle_projections
from oracle import pyoracle as orc
from pawpyseed_amd import synthetic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core import _lib
    _lib.lib().pp_set_precision(_lib.PP_FP64)
    return True


def _pair(nband=8, kpts=((0, 0, 0), (0.25, -0.25, 0.0))):
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=nband, kpts=[list(k) for k in kpts],
                               ls=(0, 0, 1, 2), alphas=(1.5, 3.0, 1.5, 2.0))
    CR = wl.coefficients("R")
    CS = wl.coefficients("S")
    return wl, CR, CS


def _oracle_single_band(wl, CR, CS, band, flip=False):
    nwk, nspin = len(wl.kpts), 1
    PR, TR = oracle_projections(wl, wl.frac_R, CR)
    PS, TS = oracle_projections(wl, wl.frac_S, CS)
    res = orc.pseudoprojection(CR, CS, band, nwk, nspin, flip=flip)
    res = orc.compensation_terms_m(res, band, nwk, nspin, PR, TR, np.zeros(len(wl.frac_R), np.int32),
                                   PS[:, band], TS, wl.M_R, wl.M_S, [wl.delta_o], [wl.species.ls], flip=flip)
    return res, PR, PS


def test_projector_single_band_projection(gpu):
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair()
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    pr = Projector(wf, basis, site_lists=(wl.M_R, wl.M_S))
    for band in (0, 5):
        got = pr.single_band_projection(band)
        want, PR, PS = _oracle_single_band(wl, CR, CS, band)
        err = normwise(got, want)
        print("single_band_projection band", band, "normwise", err)
        assert err < 1e-9
    # per-band projections held by the library
    for k in range(len(wl.kpts)):
        assert normwise(wf.get_projections(3, k), PS[k, 3]) < 1e-9
        assert normwise(basis.get_projections(7, k), PR[k, 7]) < 1e-9
    # pseudo method = pseudoprojection only
    pr2 = Projector(wf, basis, method="pseudo")
    got = pr2.single_band_projection(2)
    assert normwise(got, orc.pseudoprojection(CR, CS, 2, len(wl.kpts), 1)) < 1e-12
    with pytest.raises(ValueError):
        pr.single_band_projection(wf.nband)


def test_defect_band_analysis_runs(gpu):
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=6, kpts=((0, 0, 0),))
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    pr = Projector(wf, basis, site_lists=(wl.M_R, wl.M_S))
    out = pr.defect_band_analysis(band_list=[0, 1, 2])
    assert sorted(out) == [0, 1, 2]
    for b, (v, c) in out.items():
        res = pr.single_band_projection(b)
        assert abs((v + c) - np.sum(np.abs(res) ** 2)) < 1e-10


def _grid_encut(wl):
    abc = np.linalg.norm(wl.lattice, axis=1)
    return float(np.max((np.pi * np.asarray(wl.fftg) / abc) ** 2 / 0.262))


def _oracle_all_terms(wl, CR, CS, band, M_R, M_S, N_R, N_S, N_RS):
    """pseudoprojection + O_M + O_R + O_S + O_N (projector.c:850-963) from
    the oracle, with the filtered partial waves and off-site overlaps of
    projector.c:114-163, 604-720."""
    nwk = len(wl.kpts)
    Gs = wl.ensure_gvecs()
    PR, TR = oracle_projections(wl, wl.frac_R, CR)
    PS, TS = oracle_projections(wl, wl.frac_S, CS)
    ge = _grid_encut(wl)
    smooth = orc.smooth_species(wl.species, ge)
    res = orc.pseudoprojection(CR, CS, band, nwk, 1)
    res = orc.compensation_terms_m(res, band, nwk, 1, PR, TR, np.zeros(len(wl.frac_R), np.int32), PS[:, band], TS,
                                   np.asarray(M_R, np.int32), np.asarray(M_S, np.int32), [wl.delta_o],
                                   [wl.species.ls])
    if len(N_R) or len(N_S):
        N_R, N_S = np.asarray(N_R, np.int32), np.asarray(N_S, np.int32)
        WT_S = orc.site_tables(wl.lattice, wl.fftg, wl.frac_R[N_R], smooth)
        WP_S = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, CS, WT_S)
        WT_R = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S[N_S], smooth)
        WP_R = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, CR, WT_R)
        res = orc.compensation_terms_rs(res, nwk, 1, PR, TR, N_R, WP_S[:, band], WT_S, WP_R, WT_R, PS[:, band], TS,
                                        N_S)
    if len(N_RS):
        rr, ss = [p[0] for p in N_RS], [p[1] for p in N_RS]
        ovs, dcs = orc.offsite_blocks(wl.species, wl.species, wl.lattice, wl.frac_R, wl.frac_S, rr, ss, ge)
        res = orc.compensation_terms_n(res, nwk, 1, PR, TR, PS[:, band], TS, rr, ss, ovs, dcs, wl.kpts)
    return res


@pytest.mark.parametrize("kpts", [((0, 0, 0),), ((0, 0, 0), (0.25, -0.25, 0.0))])
def test_projector_offsite_pairs(gpu, kpts):
    """All four augmentation terms with explicit site lists, including N_RS
    pairs of unmatched sites (O_N, projector.c:940-959, with the off-site
    overlaps of radial.c:116-196), against the oracle."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=5, kpts=kpts)
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    M = [0, 1, 2, 3]
    N_R, N_S = [4, 5, 6, 7], [7, 4, 6, 5]
    N_RS = [(4, 4), (5, 5), (6, 6), (7, 7), (4, 6)]
    pr = Projector(wf, basis, site_lists=(M, M, N_R, N_S, N_RS))
    for band in (0, 3):
        got = pr.single_band_projection(band)
        want = _oracle_all_terms(wl, CR, CS, band, M, M, N_R, N_S, N_RS)
        no_on = _oracle_all_terms(wl, CR, CS, band, M, M, N_R, N_S, [])
        assert np.abs(want - no_on).max() > 1e-6 * np.abs(want).max()  # the O_N term is not negligible
        err = normwise(got, want)
        print("O_M + O_R + O_S + O_N band", band, "normwise", err, flush=True)
        assert err < 1e-9


def test_projector_default_site_lists(gpu):
    """make_site_lists (projector.py:115-160) on a jittered structure: no
    site within 0.02 A, so every term but O_M is used."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=4, kpts=((0, 0, 0),))
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    pr = Projector(wf, basis)
    M_R, M_S, N_R, N_S, N_RS = pr.make_site_lists()
    assert len(N_R) == len(wl.frac_R) and len(N_RS) >= len(N_R)
    got = pr.single_band_projection(1)
    want = _oracle_all_terms(wl, CR, CS, 1, M_R, M_S, N_R, N_S, N_RS)
    assert normwise(got, want) < 1e-9


def test_identity_projection_diagonal(gpu):
    """Projecting a structure onto itself with dO = 0: the overlap matrix is
    the pseudo overlap, whose diagonal is 1 for normalised bands."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, _ = _pair(nband=5, kpts=((0, 0, 0),))
    wl.delta_o = np.zeros_like(wl.delta_o)
    a = wavefunction_from_workload(wl, "R", coeffs=CR)
    b = wavefunction_from_workload(wl, "R", coeffs=CR)
    pr = Projector(a, b, site_lists=(wl.M_R, wl.M_R))
    for band in range(5):
        res = pr.single_band_projection(band)
        assert abs(res[band] - 1.0) < 1e-6
        assert np.abs(np.delete(res, band)).max() < 0.5


def test_compat_fft3d(gpu):
    from pawpyseed_amd.core._lib import c_double_p, c_int_p, check, lib, ptr
    wl = synthetic.tiny_config(encut=120.0, fftg=(18, 20, 15))
    G = np.ascontiguousarray(wl.ensure_gvecs()[0], np.int32)
    C = np.ascontiguousarray(wl.coefficients("R", nband=1)[0][0])
    x = np.zeros(int(np.prod(wl.fftg)), np.complex128)
    lat = np.ascontiguousarray(wl.lattice, np.float64)
    fg = np.ascontiguousarray(wl.fftg, np.int32)
    kp = np.zeros(3)
    gb = np.zeros(6, np.int32)
    check(lib().fft3d(x.ctypes.data_as(ct.c_void_p), ptr(gb, c_int_p), ptr(lat, c_double_p), ptr(kp, c_double_p),
                      ptr(G, c_int_p), C.ctypes.data_as(ct.c_void_p), len(G), ptr(fg, c_int_p)), "fft3d")
    want = orc.fft3d(wl.lattice, G, C, wl.fftg).ravel()
    assert normwise(x, want) < 1e-12


def test_compensation_terms_rejects_unmatched_lists(gpu):
    """N_RS and N_R lists need a preceding overlap_setup_real with the
    same lists."""
    from pawpyseed_amd.core import wavefunction_from_workload
    from pawpyseed_amd.core._lib import PP_ERR_ARG, lib
    wl, CR, CS = _pair(nband=3, kpts=((0, 0, 0),))
    a = wavefunction_from_workload(wl, "R", coeffs=CR, setup_projectors=True)
    b = wavefunction_from_workload(wl, "S", coeffs=CS, setup_projectors=True)
    res = np.zeros(3, np.complex128)
    one = np.array([0], np.int32)
    ip = one.ctypes.data_as(ct.POINTER(ct.c_int))
    st = lib().compensation_terms(res.ctypes.data_as(ct.c_void_p), 0, b.wf_ptr, a.wf_ptr, 0, 0, 0, 1, None, None,
                                  None, None, ip, ip, None, None, None, None, None, 0)
    assert st == PP_ERR_ARG
    st = lib().compensation_terms(res.ctypes.data_as(ct.c_void_p), 0, b.wf_ptr, a.wf_ptr, 0, 1, 0, 0, None, None,
                                  ip, None, None, None, None, None, None, None, None, 0)
    assert st == PP_ERR_ARG
    assert b"overlap_setup_real" in lib().pp_last_error()
    assert not np.any(res)


@pytest.mark.parametrize("kpts", [((0, 0, 0),), ((0, 0, 0), (0.25, -0.25, 0.0))])
def test_projector_unmatched_sites_or_os(gpu, kpts):
    """O_M over the matched sites plus O_R/O_S (projector.c:604-672,
    912-938) over the unmatched ones, through Projector and the
    reference-named overlap_setup_real / compensation_terms, against the
    oracle: the filtered partial waves (SBT, projector.c:114-163) as sphere
    tables of radius wave_rmax on the other structure's grid, projected by
    the oracle's setup_projections."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=6, kpts=kpts)
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    M = np.arange(4, dtype=np.int32)
    N_R = np.array([4, 5, 6, 7], np.int32)
    N_S = np.array([7, 4, 6, 5], np.int32)
    pr = Projector(wf, basis, site_lists=(M, M, N_R, N_S, []))
    nwk = len(wl.kpts)
    PR, TR = oracle_projections(wl, wl.frac_R, CR)
    PS, TS = oracle_projections(wl, wl.frac_S, CS)
    smooth = orc.smooth_species(wl.species, _grid_encut(wl))
    Gs = wl.ensure_gvecs()
    WT_S = orc.site_tables(wl.lattice, wl.fftg, wl.frac_R[N_R], smooth)
    WP_S = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, CS, WT_S)
    WT_R = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S[N_S], smooth)
    WP_R = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, CR, WT_R)
    assert np.abs(WP_S).max() > 1e-3 * np.abs(PS).max()  # the term is not negligible
    for band in (0, 4):
        got = pr.single_band_projection(band)
        want = orc.pseudoprojection(CR, CS, band, nwk, 1)
        want = orc.compensation_terms_m(want, band, nwk, 1, PR, TR, np.zeros(len(wl.frac_R), np.int32),
                                        PS[:, band], TS, M, M, [wl.delta_o], [wl.species.ls])
        want = orc.compensation_terms_rs(want, nwk, 1, PR, TR, N_R, WP_S[:, band], WT_S, WP_R, WT_R, PS[:, band],
                                         TS, N_S)
        err = normwise(got, want)
        print("O_M + O_R + O_S band", band, "normwise", err, flush=True)
        assert err < 1e-9




def test_projector_aug_recip(gpu):
    """method='aug_recip' (projector.c:420-453, 727-848, 965-1077): the
    augmentation functions of each structure's unmatched sites as plane-wave
    coefficients (GPU expansion + forward FFT), their overlaps with the other
    structure's coefficients, plus O_M and O_N, against the oracle
    (or_aug_freqs is pinned against numpy's FFT)."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=5, kpts=((0, 0, 0), (0.25, -0.25, 0.0)))
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS)
    M = [0, 1, 2, 3]
    N_R, N_S = [4, 5, 6, 7], [7, 4, 6, 5]
    N_RS = [(4, 4), (6, 6)]
    pr = Projector(wf, basis, method="aug_recip", site_lists=(M, M, N_R, N_S, N_RS))
    nwk = len(wl.kpts)
    Gs = wl.ensure_gvecs()
    PR, TR = oracle_projections(wl, wl.frac_R, CR)
    PS, TS = oracle_projections(wl, wl.frac_S, CS)
    smooth = orc.smooth_species(wl.species, _grid_encut(wl))
    TsR = orc.site_tables(wl.lattice, wl.fftg, wl.frac_R[N_R], smooth)
    TsS = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S[N_S], smooth)
    offR = np.array([TR.proj_off[s] for s in N_R] + [0], np.int32)
    offS = np.array([TS.proj_off[s] for s in N_S] + [0], np.int32)
    CA_R = [[orc.aug_freqs(wl.lattice, wl.fftg, wl.kpts[k], Gs[k], TsR, PR[k, b], offR) for b in range(5)]
            for k in range(nwk)]
    CA_S = [[orc.aug_freqs(wl.lattice, wl.fftg, wl.kpts[k], Gs[k], TsS, PS[k, b], offS) for b in range(5)]
            for k in range(nwk)]
    assert np.abs(CA_R[0][0]).max() > 1e-4 * np.abs(CR[0][0]).max()
    for band in (0, 3):
        got = pr.single_band_projection(band)
        want = orc.pseudoprojection(CR, CS, band, nwk, 1)
        for bR in range(5):
            for k in range(nwk):
                # projector.c:1008-1022: cdotc(CA_R, C_S) + cdotc(C_R, CA_S)
                want[bR * nwk + k] += np.vdot(CA_R[k][bR], CS[k][band].astype(np.complex128))
                want[bR * nwk + k] += np.vdot(CR[k][bR].astype(np.complex128), CA_S[k][band])
        want = orc.compensation_terms_m(want, band, nwk, 1, PR, TR, np.zeros(len(wl.frac_R), np.int32), PS[:, band],
                                        TS, np.array(M, np.int32), np.array(M, np.int32), [wl.delta_o],
                                        [wl.species.ls])
        rr, ss = [p[0] for p in N_RS], [p[1] for p in N_RS]
        ovs, dcs = orc.offsite_blocks(wl.species, wl.species, wl.lattice, wl.frac_R, wl.frac_S, rr, ss,
                                      _grid_encut(wl))
        want = orc.compensation_terms_n(want, nwk, 1, PR, TR, PS[:, band], TS, rr, ss, ovs, dcs, wl.kpts)
        err = normwise(got, want)
        print("aug_recip band", band, "normwise", err, flush=True)
        assert err < 1e-9
    # the reference keeps float complex CAs on every band
    assert basis.nband == 5


def test_noncollinear_setup_projections(gpu):
    """Spinor bands (onto_projector_ncl, projector.c:359-392): each
    component is projected on its own, into up_projections and
    down_projections; the band overlap of two spinor wavefunctions is the
    dot over both components (pseudoprojector.c:85 over num_waves = 2 W).
    Projection between structures refuses spinors (projector.py:74-75)."""
    from pawpyseed_amd.core import (CoreRegion, NCLWavefunction, PAWpyError, Projector, Pseudopotential, Structure,
                                    Wavefunction)
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=6, kpts=[[0, 0, 0], [0.25, -0.25, 0.0]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    Gs = wl.ensure_gvecs()
    Cu, Cd = wl.coefficients("S"), wl.coefficients("R")
    C2 = [np.concatenate([Cu[k], Cd[k]], axis=1) for k in range(len(Gs))]
    struct = Structure(wl.lattice, ["C"] * len(wl.frac_S), wl.frac_S)
    cr = CoreRegion({"C": Pseudopotential.from_species(wl.species)})
    wf = NCLWavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, C2, wl.encut, setup_projectors=True)
    PU, _ = oracle_projections(wl, wl.frac_S, Cu)
    PD, _ = oracle_projections(wl, wl.frac_S, Cd)
    for k in range(len(Gs)):
        for b in (0, 5):
            got = wf.get_projections(b, k)
            assert got.shape == (2, PU.shape[-1])
            eu, ed = normwise(got[0], PU[k, b]), normwise(got[1], PD[k, b])
            print("spinor projections k", k, "band", b, "up", eu, "down", ed, flush=True)
            assert eu < 1e-9 and ed < 1e-9
    C2b = [np.concatenate([Cd[k], Cu[k]], axis=1) for k in range(len(Gs))]
    wf2 = NCLWavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, C2b, wl.encut)
    got = wf.pseudoprojection(3, wf2)
    want = orc.pseudoprojection(C2b, C2, 3, len(Gs), 1)
    assert normwise(got, want) < 1e-12
    with pytest.raises(PAWpyError):
        Projector(wf, wf2)
    with pytest.raises(PAWpyError):
        Wavefunction(struct, wf._pwf, cr, wl.fftg)


def test_noncollinear_reference_wavecar(gpu, tmp_path, wavecar_ncl):
    """Spinor bands of the reference's noncollinear WAVECAR (Ga4, fixture
    subset), read through the WAVECAR reader: up and down projections of
    synthetic l = 0..2 projectors at the CONTCAR sites against the oracle
    on each half of the coefficients."""
    from pawpyseed_amd.core import CoreRegion, NCLWavefunction, Pseudopotential, Structure
    from test_abi import write_wavecar
    f = wavecar_ncl
    nk = int(f["nk_stored"])
    Cs = [f[f"C_{k}"] for k in range(nk)]
    Gs = [f[f"G_{k}"] for k in range(nk)]
    frac = np.array([[0.1560427174376784, 0.8439572825631387, 0.9166540580595424],
                     [0.3439572825623216, 0.6560427174368613, 0.4166540580595424],
                     [0.8439572825623216, 0.1560427174368613, 0.0833459419404576],
                     [0.6560427174376784, 0.3439572825631387, 0.5833459419404576]])
    p = str(tmp_path / "WAVECAR")
    write_wavecar(p, f["lattice"], float(f["encut"]), f["kpts"], Gs, Cs)
    sp = synthetic.make_species([0, 1, 2], [1.5, 2.0, 2.5], 1.2)
    struct = Structure(f["lattice"], ["Ga"] * 4, frac)
    cr = CoreRegion({"Ga": Pseudopotential.from_species(sp)})
    kws = np.full(nk, 1.0 / nk)
    wf = NCLWavefunction.from_wavecar(struct, p, cr, (30, 30, 30), f["kpts"], kws, setup_projectors=True)
    wl = synthetic.Workload("ga4ncl", f["lattice"], frac, frac, sp, np.array([30, 30, 30], np.int32),
                            float(f["encut"]), f["kpts"], kws, Cs[0].shape[0], np.eye(3), np.arange(4),
                            np.arange(4), gvecs=Gs)
    W = [len(g) for g in Gs]
    PU, _ = oracle_projections(wl, frac, [Cs[k][:, :W[k]] for k in range(nk)])
    PD, _ = oracle_projections(wl, frac, [Cs[k][:, W[k]:] for k in range(nk)])
    for k in range(nk):
        for b in range(Cs[0].shape[0]):
            got = wf.get_projections(b, k)
            eu, ed = normwise(got[0], PU[k, b]), normwise(got[1], PD[k, b])
            assert eu < 1e-9 and ed < 1e-9, (k, b, eu, ed)
    print("reference spinor WAVECAR: projections of", nk, "k x", Cs[0].shape[0], "bands match", flush=True)


def test_projector_own_pwave_overlaps(gpu):
    """O_M with the partial-wave overlap matrices get_projector_list
    integrates itself (make_pwave_overlap_matrices, projector.c:507-558), no
    channel matrix put in their place: against the oracle's O_M with
    dO = aeov - psov from or_pwave_overlaps (pinned against scipy)."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=6)
    basis = wavefunction_from_workload(wl, "R", coeffs=CR, paw_overlap=False)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS, paw_overlap=False)
    sp = wl.species
    psov, aeov, _ = orc.pwave_overlaps(sp.grid, sp.ls, sp.aewaves, sp.pswaves)
    dO = aeov - psov
    assert np.abs(dO).max() > 1e-3  # the synthetic partial waves differ
    nwk = len(wl.kpts)
    PR, TR = oracle_projections(wl, wl.frac_R, CR)
    PS, TS = oracle_projections(wl, wl.frac_S, CS)
    pr = Projector(wf, basis, site_lists=(wl.M_R, wl.M_S))
    for band in (1, 4):
        got = pr.single_band_projection(band)
        want = orc.pseudoprojection(CR, CS, band, nwk, 1)
        want = orc.compensation_terms_m(want, band, nwk, 1, PR, TR, np.zeros(len(wl.frac_R), np.int32), PS[:, band],
                                        TS, wl.M_R, wl.M_S, [dO], [sp.ls])
        with_dO = orc.compensation_terms_m(orc.pseudoprojection(CR, CS, band, nwk, 1), band, nwk, 1, PR, TR,
                                           np.zeros(len(wl.frac_R), np.int32), PS[:, band], TS, wl.M_R, wl.M_S,
                                           [wl.delta_o], [sp.ls])
        assert normwise(with_dO, want) > 1e-4  # the workload's channel matrix would not pass
        err = normwise(got, want)
        print("own partial-wave overlaps, band", band, "normwise", err, flush=True)
        assert err < 1e-9


@pytest.mark.parametrize("method", ["aug_real", "pseudo"])
def test_proportion_conduction_with_occupations(gpu, method):
    """The valence/conduction split (projector.py:384-435) of GPU
    projections with fractional and 0/1 basis occupations, spin-resolved and
    not, against the oracle's restatement of the loops on the same
    projections, which are themselves checked against the oracle."""
    from pawpyseed_amd.core import Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=8)
    nwk = len(wl.kpts)
    rng = np.random.default_rng(17)
    occ = np.clip(rng.uniform(-0.3, 1.3, (nwk, 8)), 0.0, 1.0)
    occ[:, :3] = 1.0
    basis = wavefunction_from_workload(wl, "R", coeffs=CR, occs=occ)
    wf = wavefunction_from_workload(wl, "S", coeffs=CS, occs=occ[:, ::-1].copy())
    pr = Projector(wf, basis, method=method, **({} if method == "pseudo" else dict(site_lists=(wl.M_R, wl.M_S))))
    occs_basis = basis._get_occs()
    assert np.allclose(occs_basis, occ.T.ravel())
    for band in (0, 6):
        res = pr.single_band_projection(band)
        want_res = orc.pseudoprojection(CR, CS, band, nwk, 1)
        if method != "pseudo":
            want_res, _, _ = _oracle_single_band(wl, CR, CS, band)
        assert normwise(res, want_res) < 1e-9
        for spinpol in (False, True):
            v, c = pr.proportion_conduction(band, spinpol=spinpol)
            wv, wc = orc.proportion_conduction(res, occs_basis, wf.kws, 8, nwk, 1, spinpol=spinpol,
                                               pseudo=method == "pseudo")
            assert np.allclose(v, wv, rtol=1e-12, atol=1e-15) and np.allclose(c, wc, rtol=1e-12, atol=1e-15)
            assert np.all(np.asarray(v) > 0) and np.all(np.asarray(c) > 0)
    out = pr.defect_band_analysis(num_below_ef=1, num_above_ef=2)
    # the projected structure's last band filled at its first k: occ reversed, so band 7 - 0 ... wf bands 5..7 full
    wf_occ = wf._get_occs().reshape(8, -1)[:, 0]
    vbm = int(np.flatnonzero(wf_occ > 0.5)[-1])
    assert sorted(out) == list(range(max(vbm - 1, 0), min(vbm + 2, 7) + 1))


def test_setup_multiple_projections(gpu):
    """The batch driver (projector.py:311-382) over in-memory wavefunctions:
    every yielded Projector's single_band_projection equals a Projector
    built on its own; a wavefunction on another k-point grid is refused,
    or skipped and counted with ignore_errors."""
    from pawpyseed_amd.core import PAWpyError, Projector, wavefunction_from_workload
    wl, CR, CS = _pair(nband=5)
    basis = wavefunction_from_workload(wl, "R", coeffs=CR)
    wfs = {"S": wavefunction_from_workload(wl, "S", coeffs=CS),
           "S2": wavefunction_from_workload(wl, "S", coeffs=[c * np.exp(0.7j) for c in CS])}
    got = dict(Projector.setup_multiple_projections(basis, wfs, site_lists=(wl.M_R, wl.M_S)))
    assert sorted(got) == ["S", "S2"]
    for name, pr in got.items():
        alone = Projector(wfs[name], basis, site_lists=(wl.M_R, wl.M_S))
        for band in (0, 4):
            assert normwise(pr.single_band_projection(band), alone.single_band_projection(band)) < 1e-12
    want, _, _ = _oracle_single_band(wl, CR, CS, 2)
    assert normwise(got["S"].single_band_projection(2), want) < 1e-9
    other, _, _ = _pair(nband=5, kpts=((0.1, 0, 0),))
    bad = {"S": wfs["S"], "odd": wavefunction_from_workload(other, "S", coeffs=other.coefficients("S"))}
    with pytest.raises(PAWpyError):
        list(Projector.setup_multiple_projections(basis, bad, site_lists=(wl.M_R, wl.M_S)))
    kept = list(Projector.setup_multiple_projections(basis, bad, ignore_errors=True, site_lists=(wl.M_R, wl.M_S)))
    assert [n for n, _ in kept] == ["S"]
