"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``) and exposes
numpy-level wrappers.  Only tests/, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` import this module; the product package never
does.
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ct.POINTER(ct.c_double)
_f = ct.POINTER(ct.c_float)
_i = ct.POINTER(ct.c_int)
_l = ct.POINTER(ct.c_long)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ct.CDLL(path)
        L = _LIB
        L.or_spline_integral.restype = ct.c_double
        L.or_proj_interpolate.restype = ct.c_double
        L.or_legendre.restype = ct.c_double
        L.or_legendre.argtypes = [ct.c_int, ct.c_int, ct.c_double]
        L.or_ylm.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double, _d]
        L.or_proj_interpolate.argtypes = [ct.c_double, ct.c_double, ct.c_int, _d, _d, _d, _d, _d]
        L.or_site_capacity.argtypes = [_d, _i, ct.c_double]
        L.or_site_table.argtypes = [_d, _i, _d, ct.c_double, ct.c_int, _d, ct.c_int, _i, _d, _d, _d, _d,
                                    ct.c_int, _i, _d, _d]
        L.or_wigner3j.restype = ct.c_double
        L.or_wigner3j.argtypes = [ct.c_int] * 6
        L.or_sbtfac.restype = ct.c_double
        L.or_sbtfac.argtypes = [ct.c_int] * 5
        L.or_wave_interpolate.restype = ct.c_double
        L.or_fft_use_mkl.restype = ct.c_int
        L.or_fft_use_mkl.argtypes = [ct.c_int]
    return _LIB


def fft_use_mkl(enable=True):
    """Run the oracle's 3D backward transforms through MKL (libmkl_rt, loaded
    at run time) instead of its own; True when MKL is in use."""
    return bool(lib().or_fft_use_mkl(1 if enable else 0))


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- splines
def spline_coeff(x, y):
    x = _c(x, np.float64)
    y = _c(y, np.float64)
    n = len(x)
    b, c, d = (np.zeros(n) for _ in range(3))
    lib().or_spline_coeff(_p(x, _d), _p(y, _d), ct.c_int(n), _p(b, _d), _p(c, _d), _p(d, _d))
    return b, c, d


def spline_integral(x, a, b, c, d):
    x, a, b, c, d = (_c(v, np.float64) for v in (x, a, b, c, d))
    return lib().or_spline_integral(_p(x, _d), _p(a, _d), _p(b, _d), _p(c, _d), _p(d, _d), ct.c_int(len(x)))


def proj_interpolate(r, rmax, x, f, coeffs):
    x = _c(x, np.float64)
    f = _c(f, np.float64)
    b, c, d = (_c(v, np.float64) for v in coeffs)
    return lib().or_proj_interpolate(float(r), float(rmax), len(x), _p(x, _d), _p(f, _d), _p(b, _d), _p(c, _d),
                                     _p(d, _d))


def legendre(l, m, x):
    return lib().or_legendre(int(l), int(m), float(x))


def ylm(l, m, theta, phi):
    out = np.zeros(2)
    lib().or_ylm(int(l), int(m), float(theta), float(phi), _p(out, _d))
    return complex(out[0], out[1])


# ---------------------------------------------------------------- sphere tables
class SiteTables:
    """CSR sphere tables of a whole structure (all sites, one species set)."""

    def __init__(self, site_pt, idx, path, values, val_off, nproj, proj_off, meta):
        self.site_pt = site_pt
        self.idx = idx
        self.path = path
        self.values = values          # complex128, per site [P_s, n_s] flattened
        self.val_off = val_off
        self.nproj = nproj
        self.proj_off = proj_off      # length nsites + 1
        self.meta = meta              # per species (n, l, m) arrays

    @property
    def nsites(self):
        return len(self.nproj)

    @property
    def ptot(self):
        return int(self.proj_off[-1])


def species_spline(sp):
    """(x, f[nchan, n], b, c, d) spline data of a species' projectors."""
    x = np.asarray(sp.projgrid, np.float64)
    f = np.stack([np.asarray(p, np.float64) for p in sp.realprojs])
    bcd = [spline_coeff(x, fi) for fi in f]
    b = np.stack([t[0] for t in bcd])
    c = np.stack([t[1] for t in bcd])
    d = np.stack([t[2] for t in bcd])
    return x, f, b, c, d


def projector_meta(ls):
    ns, lls, ms = [], [], []
    for ch, l in enumerate(ls):
        for m in range(-l, l + 1):
            ns.append(ch)
            lls.append(l)
            ms.append(m)
    return np.array(ns, np.int32), np.array(lls, np.int32), np.array(ms, np.int32)


def site_tables(lattice, fftg, frac, species, labels=None):
    """utils.c:590-696 for every site; species is one Species or a list
    indexed by ``labels``."""
    L = lib()
    lattice = _c(lattice, np.float64)
    fftg = _c(fftg, np.int32)
    frac = _c(frac, np.float64)
    sps = species if isinstance(species, (list, tuple)) else [species]
    labels = np.zeros(len(frac), np.int32) if labels is None else _c(labels, np.int32)
    spl = [species_spline(sp) for sp in sps]
    idxs, paths, vals, npts, nproj = [], [], [], [], []
    for s in range(len(frac)):
        sp = sps[labels[s]]
        x, f, b, c, d = spl[labels[s]]
        ls = _c(sp.ls, np.int32)
        P = int(sum(2 * l + 1 for l in sp.ls))
        cap = L.or_site_capacity(_p(lattice, _d), _p(fftg, _i), ct.c_double(sp.rmax))
        idx = np.zeros(cap, np.int32)
        path = np.zeros(3 * cap)
        values = np.zeros(2 * P * cap)
        coord = _c(frac[s], np.float64)
        n = L.or_site_table(_p(lattice, _d), _p(fftg, _i), _p(coord, _d), ct.c_double(sp.rmax), len(x),
                            _p(x, _d), len(ls), _p(ls, _i), _p(f, _d), _p(b, _d), _p(c, _d), _p(d, _d), cap,
                            _p(idx, _i), _p(path, _d), _p(values, _d))
        assert n >= 0
        v = values.view(np.complex128).reshape(P, cap)[:, :n]
        idxs.append(idx[:n])
        paths.append(path[: 3 * n].reshape(n, 3))
        vals.append(np.ascontiguousarray(v).ravel())
        npts.append(n)
        nproj.append(P)
    site_pt = np.concatenate([[0], np.cumsum(npts)]).astype(np.int32)
    val_off = np.concatenate([[0], np.cumsum([len(v) for v in vals])[:-1]]).astype(np.int64)
    proj_off = np.concatenate([[0], np.cumsum(nproj)]).astype(np.int32)
    meta = [projector_meta(sp.ls) for sp in sps]
    return SiteTables(site_pt, np.concatenate(idxs).astype(np.int32), np.concatenate(paths),
                      np.concatenate(vals), val_off, np.array(nproj, np.int32), proj_off, meta)


# ---------------------------------------------------------------- FFT
def fft3d(lattice, G, C, fftg):
    """linalg.c:14-45 for one band: c64 coefficients -> c128 grid [N0,N1,N2]."""
    fftg = _c(fftg, np.int32)
    G = _c(G, np.int32)
    C = _c(C, np.complex64)
    x = np.zeros(int(np.prod(fftg)), np.complex128)
    lib().or_fft3d(_p(x, _d), _p(_c(lattice, np.float64), _d), _p(G, _i), _p(C, _f), ct.c_int(len(G)),
                   _p(fftg, _i))
    return x.reshape(tuple(fftg))


def ifft3d(x):
    """Unscaled backward DFT (e^{+i}) of a c128 grid."""
    y = np.ascontiguousarray(x, np.complex128).copy()
    fftg = np.array(y.shape, np.int32)
    lib().or_ifft3d_inplace(_p(y, _d), _p(fftg, _i))
    return y


# ---------------------------------------------------------------- sphere sums
def onto_projector(grid, lattice, kfrac, fftg, T: SiteTables):
    grid = _c(grid, np.complex128).ravel()
    out = np.zeros(T.ptot, np.complex128)
    lib().or_onto_projector(_p(grid, _d), _p(_c(lattice, np.float64), _d), _p(_c(kfrac, np.float64), _d),
                            _p(_c(fftg, np.int32), _i), ct.c_int(T.nsites), _p(T.site_pt, _i), _p(T.nproj, _i),
                            _p(T.val_off, _l), _p(T.proj_off, _i), _p(T.idx, _i), _p(T.path, _d),
                            _p(T.values, _d), _p(out, _d))
    return out


def setup_projections(lattice, fftg, kpts, Gs, Cs, T: SiteTables, pairs=None, nthreads=0):
    """projector.c:560-602 hot loop.  Cs: per k complex64 [nband, W_k].
    Returns c128 [nk, nband, ptot] (rows not in ``pairs`` stay zero)."""
    nk = len(Gs)
    nband = Cs[0].shape[0]
    ldc = max(c.shape[1] for c in Cs)
    C = np.zeros((nk, nband, ldc), np.complex64)
    for k in range(nk):
        C[k, :, : Cs[k].shape[1]] = Cs[k]
    nw = np.array([len(g) for g in Gs], np.int32)
    Gc = [_c(g, np.int32) for g in Gs]
    Gp = (ct.POINTER(ct.c_int) * nk)(*[_p(g, _i) for g in Gc])
    kp = _c(kpts, np.float64)
    out = np.zeros((nk, nband, T.ptot), np.complex128)
    if pairs is not None:
        pr = _c(pairs, np.int32).ravel()
        npairs = len(pr) // 2
        prp = _p(pr, _i)
    else:
        npairs, prp = 0, None
    lib().or_setup_projections(_p(_c(lattice, np.float64), _d), _p(_c(fftg, np.int32), _i), ct.c_int(nk),
                               _p(kp, _d), _p(nw, _i), Gp, _p(C, _f), ct.c_long(ldc), ct.c_int(nband),
                               ct.c_int(T.nsites), _p(T.site_pt, _i), _p(T.nproj, _i), _p(T.val_off, _l),
                               _p(T.proj_off, _i), ct.c_int(T.ptot), _p(T.idx, _i), _p(T.path, _d),
                               _p(T.values, _d), ct.c_int(npairs), prp, _p(out, _d), ct.c_int(nthreads))
    return out


# ---------------------------------------------------------------- overlaps
def pseudoprojection(C_ref, C_proj, band, nwk, nspin, flip=False, fp32acc=False):
    """pseudoprojector.c:63-90.  C_ref/C_proj: per k-index lists of
    complex64 [nband, W_k] (k-index = k + spin*nwk).  Returns the
    reference-ordered c128 vector res[b*nkt + k]."""
    nkt = nwk * nspin
    nbr = C_ref[0].shape[0]
    nbp = C_proj[0].shape[0]
    ldc = max(c.shape[1] for c in C_ref)
    A = np.zeros((nkt, nbr, ldc), np.complex64)
    B = np.zeros((nkt, nbp, ldc), np.complex64)
    for k in range(nkt):
        A[k, :, : C_ref[k].shape[1]] = C_ref[k]
        B[k, :, : C_proj[k].shape[1]] = C_proj[k]
    nw = np.array([c.shape[1] for c in C_ref], np.int32)
    out = np.zeros(nbr * nkt, np.complex128)
    lib().or_pseudoprojection(_p(out, _d), _p(A, _f), _p(B, _f), ct.c_long(ldc), nbr, nbp, nwk, nspin,
                              _p(nw, _i), int(band), int(bool(flip)), int(bool(fp32acc)))
    return out


def compensation_terms_m(overlap, band, nwk, nspin, proj_R, T_R: SiteTables, labels_R, proj_S_band, T_S,
                         M_R, M_S, species_dO, species_ls, flip=False):
    """O_M of projector.c:850-910 added into ``overlap`` (c128 [nbR*nkt]).

    proj_R: c128 [nkt, nbR, ptotR]; proj_S_band: c128 [nkt, ptotS] (the
    projected band's rows); species_dO: list of channel matrices
    (aepw - pspw) per species; species_ls: list of channel l lists."""
    nkt = nwk * nspin
    nbR = proj_R.shape[1]
    metas = [projector_meta(ls) for ls in species_ls]
    meta_n = np.concatenate([m[0] for m in metas]).astype(np.int32)
    meta_l = np.concatenate([m[1] for m in metas]).astype(np.int32)
    meta_m = np.concatenate([m[2] for m in metas]).astype(np.int32)
    sp_meta_off = np.concatenate([[0], np.cumsum([len(m[0]) for m in metas])[:-1]]).astype(np.int32)
    sp_nchan = np.array([len(ls) for ls in species_ls], np.int32)
    dO = np.concatenate([np.asarray(d, np.float64).ravel() for d in species_dO])
    sp_do_off = np.concatenate([[0], np.cumsum([np.asarray(d).size for d in species_dO])[:-1]]).astype(np.int32)
    ov = _c(overlap, np.complex128)
    pr = _c(proj_R, np.complex128)
    ps = _c(proj_S_band, np.complex128)
    M_R = _c(M_R, np.int32)
    M_S = _c(M_S, np.int32)
    lab = _c(labels_R, np.int32)
    lib().or_compensation_terms_m(_p(ov, _d), int(band), nwk, nspin, nbR, _p(pr, _d), T_R.ptot,
                                  _p(T_R.proj_off, _i), _p(lab, _i), _p(ps, _d), T_S.ptot, _p(T_S.proj_off, _i),
                                  len(M_R), _p(M_R, _i), _p(M_S, _i), _p(sp_nchan, _i), _p(sp_meta_off, _i),
                                  _p(meta_n, _i), _p(meta_l, _i), _p(meta_m, _i), _p(sp_do_off, _i), _p(dO, _d),
                                  int(bool(flip)))
    return ov


def pwave_overlaps(grid, ls, aewaves, pswaves):
    grid = _c(grid, np.float64)
    n = len(grid)
    nch = len(ls)
    ae = _c(np.stack(aewaves), np.float64)
    ps = _c(np.stack(pswaves), np.float64)
    lsa = _c(ls, np.int32)
    psov, aeov, diov = (np.zeros(nch * nch) for _ in range(3))
    lib().or_pwave_overlaps(_p(grid, _d), n, nch, _p(lsa, _i), _p(ae, _d), _p(ps, _d), _p(psov, _d),
                            _p(aeov, _d), _p(diov, _d))
    return psov.reshape(nch, nch), aeov.reshape(nch, nch), diov.reshape(nch, nch)


def sbt_forward(r, f, l, lmax=None, encut=1e7, enbuf=0.0):
    """sbt.c:98-153.  Returns (k grid, g(k))."""
    r = _c(r, np.float64)
    f = _c(f, np.float64)
    n = len(r)
    ks, g = np.zeros(n), np.zeros(n)
    lib().or_sbt_forward(ct.c_double(encut), ct.c_double(enbuf), int(l if lmax is None else lmax), n, _p(r, _d),
                         _p(f, _d), int(l), _p(ks, _d), _p(g, _d))
    return ks, g


def sbt_inverse(r, g, l, lmax=None, encut=1e7, enbuf=0.0):
    """sbt.c:155-213 with the descriptor of the r grid.  Returns (r grid, R(r))."""
    r = _c(r, np.float64)
    g = _c(g, np.float64)
    n = len(r)
    rs, R = np.zeros(n), np.zeros(n)
    lib().or_sbt_inverse(ct.c_double(encut), ct.c_double(enbuf), int(l if lmax is None else lmax), n, _p(r, _d),
                         _p(g, _d), int(l), _p(rs, _d), _p(R, _d))
    return rs, R


def smooth_diffwave(wave_grid, diffwave, l, lmax, grid_encut, np_proj):
    """projector.c:114-163.  Returns (sdw on j*wave_rmax/np_proj, k grid, kwave)."""
    wg = _c(wave_grid, np.float64)
    dw = _c(diffwave, np.float64)
    n = len(wg)
    sdw, kg, kw = np.zeros(np_proj), np.zeros(n), np.zeros(n)
    lib().or_smooth_diffwave(n, _p(wg, _d), _p(dw, _d), int(l), int(lmax), ct.c_double(grid_encut), int(np_proj),
                             _p(sdw, _d), _p(kg, _d), _p(kw, _d))
    return sdw, kg, kw


def compensation_terms_rs(overlap, nwk, nspin, proj_R, T_R, N_R, wp_S_band, WT_S, wp_R, WT_R, proj_S_band, T_S, N_S,
                          flip=False):
    """O_R + O_S of projector.c:912-938 added into ``overlap`` (c128
    [nbR*nkt]).  proj_R, wp_R: c128 [nkt, nbR, ptot]; wp_S_band,
    proj_S_band: c128 [nkt, ptot] (the projected band's rows).  T_*/WT_*:
    SiteTables (only ptot and proj_off are read); WT_S holds the N_R sites in
    list order, WT_R the N_S sites."""
    nbR = proj_R.shape[1]
    ov = _c(overlap, np.complex128)
    pr = _c(proj_R, np.complex128)
    ws = _c(wp_S_band, np.complex128)
    wr = _c(wp_R, np.complex128)
    ps = _c(proj_S_band, np.complex128)
    nr = _c(N_R, np.int32)
    ns = _c(N_S, np.int32)
    lib().or_compensation_terms_rs(_p(ov, _d), nwk, nspin, nbR, _p(pr, _d), T_R.ptot, _p(T_R.proj_off, _i), len(nr),
                                   _p(nr, _i), _p(ws, _d), WT_S.ptot, _p(WT_S.proj_off, _i), _p(wr, _d), WT_R.ptot,
                                   _p(WT_R.proj_off, _i), _p(ps, _d), T_S.ptot, _p(T_S.proj_off, _i), len(ns),
                                   _p(ns, _i), int(bool(flip)))
    return ov


def smooth_species(sp, grid_encut):
    """The species' filtered partial-wave differences (projector.c:114-163)
    as a projector set on the smooth grid j * wave_rmax / npts with radius
    wave_rmax (the tables of smooth_pw_values, projector.c:211-221)."""
    import dataclasses
    wave_rmax = float(sp.grid[-1])
    lmax = max(sp.ls)
    vals = [smooth_diffwave(sp.grid, np.asarray(ae) - np.asarray(ps), l, lmax, grid_encut, sp.npts)[0]
            for l, ae, ps in zip(sp.ls, sp.aewaves, sp.pswaves)]
    return dataclasses.replace(sp, rmax=wave_rmax, realprojs=vals,
                               projgrid=np.arange(sp.npts) * wave_rmax / sp.npts)


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().or_wigner3j(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3))


def sbtfac(l1, l2, l3, m1, m2):
    return lib().or_sbtfac(int(l1), int(l2), int(l3), int(m1), int(m2))


def offsite_overlap(dcoord, k1, f1, k2, f2, l1, m1, l2, m2):
    """radial.c:116-196 for transforms f1, f2 on log k grids k1, k2."""
    k1, f1, k2, f2 = (_c(a, np.float64) for a in (k1, f1, k2, f2))
    b1, c1, d1 = spline_coeff(k1, f1)
    b2, c2, d2 = spline_coeff(k2, f2)
    dc = _c(dcoord, np.float64)
    out = np.zeros(2)
    lib().or_offsite_overlap(_p(dc, _d), _p(k1, _d), _p(f1, _d), _p(b1, _d), _p(c1, _d), _p(d1, _d), len(k1),
                             _p(k2, _d), _p(f2, _d), _p(b2, _d), _p(c2, _d), _p(d2, _d), len(k2), int(l1), int(m1),
                             int(l2), int(m2), _p(out, _d))
    return complex(out[0], out[1])


def compensation_terms_n(overlap, nwk, nspin, proj_R, T_R, proj_S_band, T_S, N_RS_R, N_RS_S, ovs, dcoords, kpts,
                         flip=False):
    """O_N of projector.c:940-959 added into ``overlap`` (c128 [nbR*nkt]);
    ovs: list of complex [P_R_s, P_S_s] matrices (the conj of the off-site
    overlaps, as overlap_setup_real stores them); dcoords [n, 3]; kpts
    [nkt, 3] fractional."""
    nbR = proj_R.shape[1]
    ov = _c(overlap, np.complex128)
    pr = _c(proj_R, np.complex128)
    ps = _c(proj_S_band, np.complex128)
    flat = _c(np.concatenate([np.asarray(o, np.complex128).ravel() for o in ovs]) if len(ovs) else np.zeros(1),
              np.complex128)
    off = _c(np.concatenate([[0], np.cumsum([np.asarray(o).size for o in ovs])]), np.int64)
    dc = _c(np.asarray(dcoords, np.float64).reshape(-1, 3) if len(ovs) else np.zeros((1, 3)), np.float64)
    kp = _c(kpts, np.float64)
    nr = _c(N_RS_R if len(ovs) else [0], np.int32)
    ns = _c(N_RS_S if len(ovs) else [0], np.int32)
    lib().or_compensation_terms_n(_p(ov, _d), nwk, nspin, nbR, _p(pr, _d), T_R.ptot, _p(T_R.proj_off, _i),
                                  _p(ps, _d), T_S.ptot, _p(T_S.proj_off, _i), len(ovs), _p(nr, _i), _p(ns, _i),
                                  _p(flat, _d), off.ctypes.data_as(ct.POINTER(ct.c_long)), _p(dc, _d), _p(kp, _d),
                                  int(bool(flip)))
    return ov


def min_cart_path(coord, center, lattice):
    """utils.c:51-73: minimum-image cartesian vector coord - center (fractional inputs) and its length."""
    c = _c(coord, np.float64)
    z = _c(center, np.float64)
    lat = _c(lattice, np.float64)
    path = np.zeros(3)
    r = ct.c_double(0.0)
    lib().or_min_cart_path(_p(c, _d), _p(z, _d), _p(lat, _d), _p(path, _d), ct.byref(r))
    return path, r.value


def offsite_blocks(species_R, species_S, lattice, frac_R, frac_S, N_RS_R, N_RS_S, grid_encut):
    """projector.c:674-712: per pair the [P_R, P_S] block conj(overlap) of the
    partial-wave differences (transforms from smooth_diffwave's kwave) and
    the minimum-image vectors from the R site to the S site."""
    def kw(sp):
        lmax = max(sp.ls)
        return [smooth_diffwave(sp.grid, np.asarray(ae) - np.asarray(ps), l, lmax, grid_encut, sp.npts)[1:]
                for l, ae, ps in zip(sp.ls, sp.aewaves, sp.pswaves)]
    kR, kS = kw(species_R), kw(species_S)
    ovs, dcs = [], []
    for s1, s2 in zip(N_RS_R, N_RS_S):
        dc, _ = min_cart_path(frac_S[s2], frac_R[s1], lattice)
        rows = []
        for (kg1, kw1), l1 in zip(kR, species_R.ls):
            for m1 in range(-l1, l1 + 1):
                row = []
                for (kg2, kw2), l2 in zip(kS, species_S.ls):
                    for m2 in range(-l2, l2 + 1):
                        row.append(np.conj(offsite_overlap(dc, kg1, kw1, kg2, kw2, l1, m1, l2, m2)))
                rows.append(row)
        ovs.append(np.array(rows))
        dcs.append(dc)
    return ovs, np.array(dcs)


def aug_freqs(lattice, fftg, kfrac, G, T: SiteTables, P, proj_off_P=None):
    """projector.c:276-330, 420-453 with linalg.c:47-79: the plane-wave
    coefficients (c128 [nw]) of sum_s sum_p P_p p_p(r) e^{-i k.path} over the
    tables' sites; P holds the band's projections, table site s at
    proj_off_P[s] (default: T.proj_off)."""
    lattice = _c(lattice, np.float64)
    fftg = _c(fftg, np.int32)
    kf = _c(kfrac, np.float64)
    G = _c(G, np.int32)
    P = _c(P, np.complex128)
    po = _c(T.proj_off if proj_off_P is None else proj_off_P, np.int32)
    vals = _c(T.values, np.complex128)
    CA = np.zeros(len(G), np.complex128)
    lib().or_aug_freqs(_p(CA, _d), _p(lattice, _d), _p(kf, _d), _p(fftg, _i), _p(G, _i), len(G), T.nsites,
                       _p(_c(T.site_pt, np.int32), _i), _p(_c(T.nproj, np.int32), _i),
                       _c(T.val_off, np.int64).ctypes.data_as(ct.POINTER(ct.c_long)), _p(_c(T.idx, np.int32), _i),
                       _p(_c(T.path, np.float64), _d), _p(vals, _d), _p(P, _d), _p(po, _i))
    return CA


# ---------------------------------------------------------------- real-space states
def wave_value2(x, f, coeffs, l, m, pos):
    """utils.c:522-545 for one channel spline (x, f, (b, c, d))."""
    x, f = _c(x, np.float64), _c(f, np.float64)
    b, c, d = (_c(v, np.float64) for v in coeffs)
    out = np.zeros(2)
    lib().or_wave_value2(ct.c_int(len(x)), _p(x, _d), _p(f, _d), _p(b, _d), _p(c, _d), _p(d, _d), ct.c_int(l),
                         ct.c_int(m), _p(_c(pos, np.float64), _d), _p(out, _d))
    return complex(out[0], out[1])


class WaveData:
    """Per-species diffwave splines (projector.c:84-112: diffwave = ae - ps on
    the log wave grid) packed for ``or_realspace_state``."""

    def __init__(self, species, diffwaves=None):
        self.wn = np.array([len(sp.grid) for sp in species], np.int32)
        self.woff = np.concatenate([[0], np.cumsum(self.wn)[:-1]]).astype(np.int64)
        self.wg = np.concatenate([np.asarray(sp.grid, np.float64) for sp in species])
        self.chan_off = np.concatenate([[0], np.cumsum([len(sp.ls) for sp in species])]).astype(np.int32)
        self.chan_l = np.concatenate([np.asarray(sp.ls, np.int32) for sp in species]).astype(np.int32)
        self.wmax = int(self.wn.max())
        nch = int(self.chan_off[-1])
        self.f, self.b, self.c, self.d = (np.zeros((nch, self.wmax)) for _ in range(4))
        ch = 0
        for e, sp in enumerate(species):
            for q in range(len(sp.ls)):
                fq = (np.asarray(diffwaves[e][q], np.float64) if diffwaves is not None else
                      np.asarray(sp.aewaves[q], np.float64) - np.asarray(sp.pswaves[q], np.float64))
                n = len(sp.grid)
                self.f[ch, :n] = fq
                bb, cc, dd = spline_coeff(sp.grid, fq)
                self.b[ch, :n], self.c[ch, :n], self.d[ch, :n] = bb, cc, dd
                ch += 1
        self.proj_counts = [sum(2 * l + 1 for l in sp.ls) for sp in species]


def realspace_state(lattice, fftg, kfrac, G, C, W: WaveData, labels, coords, proj, remove_phase=False):
    """density.c:232-310 for one band: fft3d of the c64 coefficients, the
    Bloch phase, and the partial-wave differences weighted by ``proj``
    (c128 [ptot], sites in order, projectors channel-major, m = -l..l).
    With ``remove_phase`` the result is multiplied by e^{-2 pi i k.r}
    afterwards (density.c:312-326)."""
    labels = _c(labels, np.int32)
    coords = _c(coords, np.float64).reshape(-1, 3)
    nsites = len(labels)
    counts = [W.proj_counts[e] for e in labels]
    proj_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    proj = _c(proj, np.complex128)
    assert len(proj) >= proj_off[-1]
    x = np.ascontiguousarray(fft3d(lattice, G, C, fftg)).ravel()
    fg = _c(fftg, np.int32)
    kf = _c(kfrac, np.float64)
    lib().or_realspace_state(_p(x, _d), _p(_c(lattice, np.float64), _d), _p(fg, _i), _p(kf, _d),
                             ct.c_int(nsites), _p(labels, _i), _p(coords, _d), _p(W.wn, _i), _p(W.woff, _l),
                             _p(W.wg, _d), _p(W.chan_off, _i), _p(W.chan_l, _i), ct.c_int(W.wmax), _p(W.f, _d),
                             _p(W.b, _d), _p(W.c, _d), _p(W.d, _d), _p(proj_off, _i), _p(proj, _d))
    if remove_phase:
        lib().or_remove_phase(_p(x, _d), _p(fg, _i), _p(kf, _d))
    return x.reshape(tuple(fg))


def proportion_conduction(res, occs, kws, nband, nwk, nspin, spinpol=False, pseudo=False):
    """projector.py:384-435 restated loop by loop: the valence (v) and
    conduction (c) weight of one projected band from its projections
    res[b * nwk * nspin + s * nwk + k] onto the basis states, whose
    occupations occs share that ordering (pawpyc.pyx:532-538)."""
    res = np.asarray(res)
    occs = np.asarray(occs, np.float64)
    if spinpol:
        v, c = [0.0] * nspin, [0.0] * nspin
        for b in range(nband):
            for s in range(nspin):
                for k in range(nwk):
                    i = (b * nspin + s) * nwk + k
                    p = abs(res[i]) ** 2
                    v[s] += p * occs[i] * kws[k]
                    c[s] += p * (1.0 - occs[i]) * kws[k]
    else:
        v = c = 0.0
        for i in range(nband * nwk * nspin):
            p = abs(res[i]) ** 2 * kws[i % nwk] / nspin
            if occs[i] > 0.5:
                v += p
            else:
                c += p
    if pseudo:
        if spinpol:
            t = [a + b for a, b in zip(v, c)]
            v, c = [a / x for a, x in zip(v, t)], [b / x for b, x in zip(c, t)]
        else:
            t = v + c
            v, c = v / t, c / t
    return v, c


def analysis_bands(occs_wf, nband, num_below_ef, num_above_ef, vbmband=None, band_list=None, analyze_all=False,
                   nband_basis=None):
    """projector.py:477-494: the bands defect_band_analysis looks at; occs_wf
    in get_occs order (band-major over the wavefunction's k-points)."""
    if analyze_all:
        return list(range(nband_basis if nband_basis is not None else nband))
    if band_list:
        return list(band_list)
    nk = len(occs_wf) // nband
    vbm = 0
    for i in range(nband):
        if occs_wf[i * nk] > 0.5:
            vbm = i
    if vbmband is not None:
        vbm = vbmband
    return list(range(max(vbm - num_below_ef, 0), min(vbm + num_above_ef, nband - 1) + 1))
