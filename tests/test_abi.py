"""C-ABI checks that need no GPU: the library loads, exports every entry
point declared in include/*.h, and its host-only entry points (WAVECAR
reader, in-memory constructor, accessors, get_projector_list) behave as
the reference's."""

import ctypes as ct
import glob
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic
from pawpyseed_amd.core import _lib
from pawpyseed_amd.core._lib import c_double_p, lib, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "sizeof", "defined"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    L = lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) <= names


def test_version_and_error_plumbing():
    L = lib()
    assert b"gfx950" in L.pp_version()
    assert L.pp_set_precision(7) == _lib.PP_ERR_ARG
    assert b"unknown precision" in L.pp_last_error()
    assert L.pp_set_precision(_lib.PP_FP64) == 0


def write_wavecar(path, lattice, encut, kpts, Gs, coeffs, nspin=1, occ=None, energies=None):
    """Single-precision WAVECAR in the record layout read by reader.c:129-315."""
    nwk = len(kpts)
    nband = coeffs[0].shape[0]
    maxw = max(c.shape[1] for c in coeffs)
    nrecl = max(8 * maxw, 8 * (4 + 3 * nband), 8 * 12)
    nrecl = (nrecl + 7) // 8 * 8
    nrec = 2 + nwk * nspin * (1 + nband)
    buf = np.zeros(nrec * nrecl, np.uint8)

    def put(rec, arr):
        b = np.asarray(arr).tobytes()
        buf[rec * nrecl:rec * nrecl + len(b)] = np.frombuffer(b, np.uint8)

    put(0, np.array([nrecl, nspin, 45200], np.float64))
    put(1, np.concatenate([[nwk, nband, encut], np.asarray(lattice, np.float64).ravel()]))
    for iwk in range(nwk * nspin):
        irec = iwk * (1 + nband)
        head = np.zeros(4 + 3 * nband)
        head[0] = coeffs[iwk].shape[1]
        head[1:4] = kpts[iwk % nwk]
        for b in range(nband):
            head[4 + 3 * b] = 0.0 if energies is None else energies[iwk][b]
            head[6 + 3 * b] = 0.0 if occ is None else occ[iwk][b]
        put(irec + 2, head)
        for b in range(nband):
            put(irec + 3 + b, np.asarray(coeffs[iwk][b], np.complex64))
    buf.tofile(path)


def test_reader_roundtrip(tmp_path):
    wl = synthetic.tiny_config(nband=3, kpts=[[0, 0, 0], [0.25, 0.25, 0.0]])
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("R")
    occ = [[1, 1, 0], [1, 0.5, 0]]
    en = [[-1, 0, 1], [-2, 0.5, 3]]
    p = str(tmp_path / "WAVECAR")
    write_wavecar(p, wl.lattice, wl.encut, wl.kpts, Gs, Cs, occ=occ, energies=en)
    from pawpyseed_amd.core.wavefunction import PWFPointer
    pw = PWFPointer.from_wavecar(p, wl.kpts, wl.kws)
    L = lib()
    assert L.get_nband(pw.ptr) == 3 and L.get_nwk(pw.ptr) == 2 and L.get_nspin(pw.ptr) == 1
    assert L.get_encut(pw.ptr) == wl.encut
    assert L.get_occ(pw.ptr, 1, 1, 0) == 0.5 and L.get_energy(pw.ptr, 2, 1, 0) == 3
    L.free_pswf(pw.ptr)


def test_reader_on_reference_wavecar(wavecar):
    path = "/root/reference/test_files/WAVECAR"
    if not os.path.exists(path):
        pytest.skip("reference fixture not present on this machine")
    w = np.array([0.5, 0.5])
    L = lib()
    p = L.read_wavefunctions(path.encode(), ptr(w, c_double_p))
    assert p
    assert (L.get_nband(p), L.get_nwk(p), L.get_nspin(p)) == (24, 2, 2)
    occ = L.get_occ(p, 3, 1, 1)
    assert occ == wavecar["occ_3"][3]
    L.free_pswf(p)


def test_reader_spinor_wavecar(tmp_path):
    """A WAVECAR whose plane-wave count is twice the G vectors inside the
    cutoff is a spinor one (reader.c:274-281): is_ncl set, the G list kept
    twice."""
    wl = synthetic.tiny_config(nband=2, kpts=[[0, 0, 0], [0.25, 0.25, 0.0]])
    Gs = wl.ensure_gvecs()
    Cu, Cd = wl.coefficients("S"), wl.coefficients("R")
    C2 = [np.concatenate([Cu[k], Cd[k]], axis=1) for k in range(2)]
    p = str(tmp_path / "WAVECAR")
    write_wavecar(p, wl.lattice, wl.encut, wl.kpts, Gs, C2)
    from pawpyseed_amd.core.wavefunction import PWFPointer
    pw = PWFPointer.from_wavecar(p, wl.kpts, wl.kws)
    L = lib()
    assert L.is_ncl(pw.ptr) == 1 and L.get_nband(pw.ptr) == 2
    L.free_pswf(pw.ptr)
    # the same arrays given directly: a G list repeated is a spinor one
    pw = PWFPointer.from_arrays(wl.lattice, wl.kpts, wl.kws, [np.concatenate([g, g]) for g in Gs], C2, wl.encut)
    assert L.is_ncl(pw.ptr) == 1
    L.free_pswf(pw.ptr)
    pw = PWFPointer.from_arrays(wl.lattice, wl.kpts, wl.kws, Gs, Cu, wl.encut)
    assert L.is_ncl(pw.ptr) == 0
    L.free_pswf(pw.ptr)


def test_reader_spinor_fixture(tmp_path, wavecar_ncl):
    """The spinor records of the reference's noncollinear WAVECAR (fixture
    subset), written back in the WAVECAR layout and read: spinor flag, the
    enumerated G count is half the stored plane-wave count, occupations."""
    f = wavecar_ncl
    nk = int(f["nk_stored"])
    Cs = [f[f"C_{k}"] for k in range(nk)]
    Gs = [f[f"G_{k}"] for k in range(nk)]
    assert all(2 * len(Gs[k]) == int(f[f"nplane_{k}"]) == Cs[k].shape[1] for k in range(nk))
    p = str(tmp_path / "WAVECAR")
    write_wavecar(p, f["lattice"], float(f["encut"]), f["kpts"], Gs, Cs, occ=[f[f"occ_{k}"] for k in range(nk)],
                  energies=[f[f"energy_{k}"] for k in range(nk)])
    w = np.full(nk, 1.0 / nk)
    L = lib()
    ptr_ = L.read_wavefunctions(p.encode(), ptr(w, c_double_p))
    assert ptr_
    assert L.is_ncl(ptr_) == 1 and L.get_nband(ptr_) == Cs[0].shape[0] and L.get_nwk(ptr_) == nk
    assert L.get_occ(ptr_, 2, 1, 0) == f["occ_1"][2]
    L.free_pswf(ptr_)


def test_reader_on_reference_ncl_wavecar():
    """The reference's noncollinear fixture (test_files/noncollinear:
    4 k-points, 24 spinor bands, 520 eV)."""
    path = "/root/reference/test_files/noncollinear/WAVECAR"
    if not os.path.exists(path):
        pytest.skip("reference fixture not present on this machine")
    w = np.full(4, 0.25)
    L = lib()
    p = L.read_wavefunctions(path.encode(), ptr(w, c_double_p))
    assert p
    assert L.is_ncl(p) == 1
    assert (L.get_nband(p), L.get_nwk(p), L.get_nspin(p)) == (24, 4, 1)
    assert L.get_encut(p) == 520.0
    L.free_pswf(p)


def test_reader_errors():
    L = lib()
    w = np.array([1.0])
    assert not L.read_wavefunctions(b"/nonexistent/WAVECAR", ptr(w, c_double_p))
    assert b"cannot open" in L.pp_last_error()


def test_get_projector_list_overlaps_match_oracle():
    """projector.c:507-558 overlap matrices from get_projector_list equal
    the oracle's spline integrals (themselves pinned against scipy)."""
    from pawpyseed_amd.core.ppot import ppot_array
    sp = synthetic.make_species([0, 0, 1, 1, 2], [1.5, 3.0, 1.5, 3.0, 2.0], 1.3)
    clabels = np.array([0, len(sp.ls), len(sp.realprojs[0]), len(sp.grid)], np.int32)
    ls = np.array(sp.ls, np.int32)
    args = [np.ascontiguousarray(a, np.float64) for a in
            (sp.grid, np.concatenate(sp.realprojs), np.concatenate(sp.aewaves), np.concatenate(sp.pswaves),
             [sp.rmax])]
    L = lib()
    pl = L.get_projector_list(1, clabels.ctypes.data_as(_lib.c_int_p), ls.ctypes.data_as(_lib.c_int_p),
                              *[ptr(a, c_double_p) for a in args], 100.0)
    assert pl
    pp = ppot_array(pl, 1)[0]
    n = len(sp.ls)
    assert pp.num_projs == n and pp.total_projs == sum(2 * l + 1 for l in sp.ls) and pp.lmax == 2
    psov, aeov, diov = orc.pwave_overlaps(sp.grid, sp.ls, sp.aewaves, sp.pswaves)
    got_ps = np.ctypeslib.as_array(pp.pspw_overlap_matrix, shape=(n * n,)).reshape(n, n)
    got_ae = np.ctypeslib.as_array(pp.aepw_overlap_matrix, shape=(n * n,)).reshape(n, n)
    got_di = np.ctypeslib.as_array(pp.diff_overlap_matrix, shape=(n * n,)).reshape(n, n)
    np.testing.assert_allclose(got_ps, psov, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(got_ae, aeov, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(got_di, diov, rtol=1e-12, atol=1e-15)
    # projector grid x_j = j rmax / n and the spline of utils.c:699-749
    x = np.ctypeslib.as_array(pp.proj_grid, shape=(len(sp.projgrid),))
    np.testing.assert_allclose(x, sp.projgrid, rtol=0, atol=1e-15)
    b, c, d = orc.spline_coeff(sp.projgrid, sp.realprojs[2])
    f = pp.funcs[2]
    np.testing.assert_allclose([f.proj_spline[0][i] for i in range(10)], b[:10], rtol=1e-10, atol=1e-13)
    # projector.c:114-163: transforms and filtered partial-wave differences
    # (grid_encut 100 eV) equal the oracle's
    nw, npj = len(sp.grid), len(sp.projgrid)
    kgrid = np.ctypeslib.as_array(pp.kwave_grid, shape=(nw,))
    sg = np.ctypeslib.as_array(pp.smooth_grid, shape=(npj,))
    np.testing.assert_allclose(sg, sp.grid[-1] / npj * np.arange(npj), rtol=0, atol=1e-15)
    assert abs(pp.wave_rmax - sp.grid[-1]) < 1e-15
    for ch, l in enumerate(sp.ls):
        so, ko, wo = orc.smooth_diffwave(sp.grid, sp.aewaves[ch] - sp.pswaves[ch], l, max(sp.ls), 100.0, npj)
        fc = pp.funcs[ch]
        np.testing.assert_allclose(kgrid, ko, rtol=1e-14)
        kw = np.ctypeslib.as_array(fc.kwave, shape=(nw,))
        sd = np.ctypeslib.as_array(fc.smooth_diffwave, shape=(npj,))
        assert np.abs(kw - wo).max() < 1e-11 * np.abs(wo).max()
        assert np.abs(sd - so).max() < 1e-10 * np.abs(so).max()
    L.free_ppot_list(ct.c_void_p(pl), 1)


def test_pswf_from_arrays_accessors():
    from pawpyseed_amd.core.wavefunction import PWFPointer
    wl = synthetic.tiny_config(nband=4, kpts=[[0, 0, 0]])
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("R") * 2
    occ = np.array([[1.0, 1.0, 0.2, 0.0], [1.0, 0.0, 0.0, 0.0]])
    pw = PWFPointer.from_arrays(wl.lattice, wl.kpts, wl.kws, Gs * 2, Cs, wl.encut, nspin=2, occs=occ)
    L = lib()
    assert L.get_nspin(pw.ptr) == 2 and L.get_nwk(pw.ptr) == 1
    assert L.get_occ(pw.ptr, 2, 0, 0) == 0.2 and L.get_occ(pw.ptr, 0, 0, 1) == 1.0
    L.free_pswf(pw.ptr)


def test_write_volumetric_format(tmp_path):
    """density.c:461-477: values x[i][j][k] * scale with z slowest and x
    fastest, printf "%E   " each, a newline after every fifth value; the
    VASP-style header of wavefunction.py:480-510 in front."""
    from pawpyseed_amd.core.wavefunction import write_volumetric
    dim = (3, 4, 2)
    x = np.arange(24, dtype=np.float64).reshape(dim) * 0.5 - 3.0
    p = tmp_path / "v.vasp"
    write_volumetric(str(p), x, dim, 2.0)
    txt = p.read_text()
    lines = txt.split("\n")
    assert all(len(l.split()) == 5 for l in lines[:4]) and len(lines[4].split()) == 4
    vals = [float(v) for v in txt.split()]
    want = [x[i, j, k] * 2.0 for k in range(2) for j in range(4) for i in range(3)]
    assert vals == want
    assert txt.split()[0] == "%E" % want[0] and "%E   " % want[1] in txt
