"""Projector: band projections of one structure onto another's bands.

Mirrors pawpyseed/core/projector.py (Projector, make_site_lists,
setup_overlap, single_band_projection, proportion_conduction,
defect_band_analysis) with the 'aug_real', 'aug_recip', 'realspace' and
'pseudo' methods.  The
pseudo-wavefunction overlap and every augmentation term run in
libpawpy_amd.so: the on-site O_M term, the O_R/O_S terms of sites that
differ between the structures and the off-site O_N term of unmatched sites
with overlapping augmentation spheres (overlap_setup_real,
compensation_terms).  'realspace' integrates both all-electron states on a
real-space grid (project_realspace_state).
"""

from __future__ import annotations

import numpy as np

from ._lib import PAWpyError, c_int_p as _int_p, ptr as _ptr


def _arr(a):
    """int32 array as a C pointer, NULL when empty (pawpyc.pyx:666-671)."""
    return _ptr(a, _int_p) if len(a) else None


class Projector:
    METHODS = ["pseudo", "realspace", "aug_recip", "aug_real"]

    def __init__(self, wf, basis, unsym_basis=False, unsym_wf=False, method="aug_real", site_lists=None):
        self.method = method
        if method == "pseudo":
            self._single_band_projection = self._single_band_projection_pseudo
        elif method == "aug_real":
            self._single_band_projection = self._single_band_projection_aug_real
        elif method == "aug_recip":
            self._single_band_projection = self._single_band_projection_aug_recip
        elif method == "realspace":
            self._single_band_projection = self._single_band_projection_realspace
        else:
            raise PAWpyError("method not recognized for Projector")
        if wf.ncl or basis.ncl:
            raise PAWpyError("Projection not supported for noncollinear case!")
        if unsym_basis or unsym_wf:
            raise PAWpyError("k-point desymmetrisation is not provided by this build")
        # projector.py:97-104
        if np.linalg.norm(basis.kpts - wf.kpts) > 1e-10:
            raise PAWpyError("k-point grids for projection are not matched.")
        if np.linalg.norm(basis.kws - wf.kws) > 1e-10:
            raise PAWpyError("k-point weights for projection are not matched.")
        if not (wf.structure.lattice == basis.structure.lattice):
            raise PAWpyError("Need the lattice to be the same for projections, and they are not")
        self.wf = wf
        self.basis = basis
        self._site_lists = site_lists
        if method != "pseudo":
            basis.check_c_projectors()
            wf.check_c_projectors()
        if "aug" in method:
            self.setup_overlap()

    def make_site_lists(self):
        """projector.py:115-160: M_R/M_S identical sites (same element,
        within 0.02 A), N_R/N_S the rest, N_RS pairs of unmatched sites whose
        augmentation spheres overlap."""
        ref, cur = self.basis.structure, self.wf.structure
        d = ref.distance_matrix_to(cur)
        same_el = np.array(ref.species)[:, None] == np.array(cur.species)[None, :]
        mr, ms = np.nonzero((d <= 0.02) & same_el)
        M_R, M_S = mr.tolist(), ms.tolist()
        N_R = [i for i in range(len(ref)) if i not in set(M_R)]
        N_S = [j for j in range(len(cur)) if j not in set(M_S)]
        N_RS = []
        for i in N_R:
            for j in N_S:
                rmax1 = self.basis.cr.pps[ref.species[i]].rmax
                rmax2 = self.wf.cr.pps[cur.species[j]].rmax
                if d[i, j] < rmax1 + rmax2:
                    N_RS.append((i, j))
        return M_R, M_S, N_R, N_S, N_RS

    def setup_overlap(self):
        """projector.py:162-187 and pawpyc.pyx:651-683.  ``site_lists`` may
        be (M_R, M_S) (on-site term only) or (M_R, M_S, N_R, N_S, N_RS)."""
        if self._site_lists is not None:
            sl = list(self._site_lists)
            M_R, M_S = sl[0], sl[1]
            N_R, N_S, N_RS = (sl[2], sl[3], sl[4]) if len(sl) == 5 else ([], [], [])
        else:
            M_R, M_S, N_R, N_S, N_RS = self.make_site_lists()
        self.M_R = np.asarray(M_R, np.int32)
        self.M_S = np.asarray(M_S, np.int32)
        self.N_R = np.asarray(N_R, np.int32)
        self.N_S = np.asarray(N_S, np.int32)
        N_RS = [tuple(p) for p in N_RS]
        self.N_RS_R = np.asarray([p[0] for p in N_RS], np.int32)
        self.N_RS_S = np.asarray([p[1] for p in N_RS], np.int32)
        self.site_cat = [self.M_R, self.M_S, self.N_R, self.N_S, self.N_RS_R, self.N_RS_S]
        recip = self.method == "aug_recip"
        if recip or len(self.N_R) or len(self.N_S) or len(self.N_RS_R):
            from ._lib import c_double_p, check, lib, ptr
            fn = lib().overlap_setup_recip if recip else lib().overlap_setup_real
            check(fn(self.basis.wf_ptr, self.wf.wf_ptr, ptr(self.basis.nums, _int_p), ptr(self.wf.nums, _int_p),
                     ptr(self.basis.coords, c_double_p), ptr(self.wf.coords, c_double_p), _arr(self.N_R),
                     _arr(self.N_S), _arr(self.N_RS_R), _arr(self.N_RS_S), len(self.N_R), len(self.N_S),
                     len(self.N_RS_R)), "overlap_setup_recip" if recip else "overlap_setup_real")

    def _single_band_projection_pseudo(self, band_num, flip_spin=False):
        return self.wf.pseudoprojection(band_num, self.basis, flip_spin)

    def _single_band_projection_realspace(self, band_num, dim=None):
        """projector.py:199-208: both all-electron states on a real-space
        grid (default: twice the wavefunction's grid) and their overlaps
        V/N^3 sum_r conj(psi_R) psi_S (project_realspace_state,
        density.c:205-230, on the GPU)."""
        if dim is None:
            dim = self.wf.dim * 2
        return self._realspace_projection(band_num, dim)

    def _realspace_projection(self, band_num, dim):
        """pawpyc.pyx:723-735 (with the grid passed as int32)."""
        import ctypes as ct

        from ._lib import c_double_p, check, lib, ptr
        res = np.zeros(self.basis.nband * self.basis.nwk * self.basis.nspin, np.complex128)
        dimv = np.ascontiguousarray(self.wf.dimv if dim is None else dim, np.int32)
        check(lib().project_realspace_state(res.ctypes.data_as(ct.c_void_p), int(band_num), self.wf.wf_ptr,
                                            self.basis.wf_ptr, ptr(dimv, _int_p), ptr(self.wf.nums, _int_p),
                                            ptr(self.wf.coords, c_double_p), ptr(self.basis.nums, _int_p),
                                            ptr(self.basis.coords, c_double_p)), "project_realspace_state")
        return res

    def _single_band_projection_aug_real(self, band_num, flip_spin=False):
        res = self.wf.pseudoprojection(band_num, self.basis, flip_spin)
        self._add_augmentation_terms(res, band_num, flip_spin)
        return res

    def _single_band_projection_aug_recip(self, band_num, flip_spin=False):
        """projector.py:225-236: pseudoprojection + compensation_terms_recip."""
        res = self.wf.pseudoprojection(band_num, self.basis, flip_spin)
        self._add_augmentation_terms(res, band_num, flip_spin, recip=True)
        return res

    def _add_augmentation_terms(self, res, band_num, flip_spin, recip=False):
        """pawpyc.pyx:685-702 -> compensation_terms (O_M, O_R, O_S)."""
        import ctypes as ct

        from ._lib import c_double_p, check, lib, ptr
        fn = lib().compensation_terms_recip if recip else lib().compensation_terms
        check(fn(res.ctypes.data_as(ct.c_void_p), int(band_num), self.wf.wf_ptr, self.basis.wf_ptr, len(self.M_R),
                 len(self.N_R), len(self.N_S), len(self.N_RS_R), _arr(self.M_R), _arr(self.M_S), _arr(self.N_R),
                 _arr(self.N_S), _arr(self.N_RS_R), _arr(self.N_RS_S), ptr(self.wf.nums, _int_p),
                 ptr(self.wf.coords, c_double_p), ptr(self.basis.nums, _int_p), ptr(self.basis.coords, c_double_p),
                 ptr(self.wf.dimv, _int_p), int(bool(flip_spin))),
              "compensation_terms_recip" if recip else "compensation_terms")

    def single_band_projection(self, band_num, **kwargs):
        """projector.py:238-272: overlaps <basis; b,k,s | wf; band_num,k,s>
        ordered res[b*nwk*nspin + s*nwk + k]."""
        if band_num >= self.wf.nband or band_num < 0:
            raise ValueError("Band index out of range (0-indexed)")
        return self._single_band_projection(band_num, **kwargs)

    def proportion_conduction(self, band_num, spinpol=False):
        """projector.py:384-435."""
        basis = self.basis
        nband, nwk, nspin = basis.nband, basis.nwk, basis.nspin
        occs = basis._get_occs()
        res = self.single_band_projection(band_num)
        if spinpol:
            c, v = np.zeros(nspin), np.zeros(nspin)
            for b in range(nband):
                for s in range(nspin):
                    ind = b * nspin + s
                    prop = np.absolute(res[ind * nwk:(ind + 1) * nwk]) ** 2
                    c[s] += np.dot(prop, (1 - occs[ind * nwk:(ind + 1) * nwk]) * self.wf.kws)
                    v[s] += np.dot(prop, occs[ind * nwk:(ind + 1) * nwk] * self.wf.kws)
        else:
            c, v = 0, 0
            for i in range(nband * nwk * nspin):
                if occs[i] > 0.5:
                    v += np.absolute(res[i]) ** 2 * self.wf.kws[i % nwk] / nspin
                else:
                    c += np.absolute(res[i]) ** 2 * self.wf.kws[i % nwk] / nspin
        if self.method == "pseudo":
            t = v + c
            v /= t
            c /= t
        if spinpol:
            v, c = v.tolist(), c.tolist()
        return v, c

    def defect_band_analysis(self, num_below_ef=20, num_above_ef=20, spinpol=False, return_energies=False,
                             vbmband=None, band_list=None, analyze_all=False):
        """projector.py:437-503."""
        if num_below_ef < 0 or num_above_ef < 0:
            raise ValueError("num_above_ef and num_below_ef must both be nonnegative.")
        nband = self.basis.nband
        occs = self.wf._get_occs()
        if analyze_all:
            totest = list(range(nband))
        elif band_list:
            totest = band_list[:]
        else:
            vbm = 0
            for i in range(self.wf.nband):
                if occs[i * self.wf.nwk * self.wf.nspin] > 0.5:
                    vbm = i
            if vbmband is not None:
                vbm = vbmband
            lo, hi = max(vbm - num_below_ef, 0), min(vbm + num_above_ef, self.wf.nband - 1)
            totest = list(range(lo, hi + 1))
        results = {b: self.proportion_conduction(b, spinpol=spinpol) for b in totest}
        if return_energies:
            return results, self.wf._get_energy_list(totest)
        return results
