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

    @staticmethod
    def setup_multiple_projections(basis, wfs, method="aug_real", ignore_errors=False, site_lists=None):
        """The batch driver of projector.py:311-382 over wavefunctions already
        in memory (the reference reads each from a VASP directory; here the
        caller passes the Wavefunction objects, as a list or a {name: wf}
        mapping).  Yields [name, Projector] for each, all projected onto
        ``basis``; with ``ignore_errors`` a wavefunction whose projector
        cannot be set up is skipped and counted, else PAWpyError names it.
        ``site_lists`` (optional, per name) is passed to Projector."""
        items = wfs.items() if isinstance(wfs, dict) else enumerate(wfs)
        basis.check_c_projectors()
        errors, made = 0, 0
        for name, wf in items:
            try:
                sl = site_lists.get(name) if isinstance(site_lists, dict) else site_lists
                pr = Projector(wf, basis, method=method, site_lists=sl)
            except Exception as e:
                if not ignore_errors:
                    raise PAWpyError(f"Unable to set up wavefunction {name}\nGot the following error:\n{e}")
                errors += 1
                continue
            made += 1
            yield [name, pr]
        if not made:
            raise PAWpyError("Could not generate any projector setups")
        print("Number of errors:", errors)

    def proportion_conduction(self, band_num, spinpol=False):
        """Valence and conduction weight of band ``band_num`` of the
        projected structure in the basis (projector.py:384-435).

        The projections are viewed as a (band, spin, k) array and weighted
        by |<basis|wf>|^2 w_k.  Spin-resolved, every basis state counts
        toward valence with its occupation f and toward conduction with
        1 - f; otherwise a state is valence when f > 0.5 and the sums are
        divided by the spin count.  The 'pseudo' method normalises v + c to
        one.  Returns floats, or per-spin lists when ``spinpol``."""
        basis = self.basis
        shape = (basis.nband, basis.nspin, basis.nwk)
        amp2 = np.abs(np.asarray(self.single_band_projection(band_num))).reshape(shape) ** 2
        weighted = amp2 * np.asarray(self.wf.kws, np.float64)[None, None, :]
        occ = np.asarray(basis._get_occs(), np.float64).reshape(shape)
        if spinpol:
            v = np.einsum("bsk,bsk->s", weighted, occ)
            c = np.einsum("bsk,bsk->s", weighted, 1.0 - occ)
        else:
            filled = occ > 0.5
            v = float(weighted[filled].sum()) / basis.nspin
            c = float(weighted[~filled].sum()) / basis.nspin
        if self.method == "pseudo":
            total = v + c
            v, c = v / total, c / total
        if spinpol:
            return np.asarray(v, np.float64).tolist(), np.asarray(c, np.float64).tolist()
        return v, c

    def _analysis_bands(self, num_below_ef, num_above_ef, vbmband, band_list, analyze_all):
        """Bands defect_band_analysis looks at (projector.py:477-494): all of
        them, an explicit list, or a window around the valence band maximum,
        i.e. the highest band whose first (k, spin) state has f > 0.5."""
        if analyze_all:
            return list(range(self.basis.nband))
        if band_list:
            return list(band_list)
        nb = self.wf.nband
        first_state_occ = np.asarray(self.wf._get_occs(), np.float64).reshape(nb, -1)[:, 0]
        occupied = np.flatnonzero(first_state_occ > 0.5)
        vbm = int(occupied[-1]) if occupied.size else 0
        if vbmband is not None:
            vbm = vbmband
        return list(range(max(vbm - num_below_ef, 0), min(vbm + num_above_ef, nb - 1) + 1))

    def defect_band_analysis(self, num_below_ef=20, num_above_ef=20, spinpol=False, return_energies=False,
                             vbmband=None, band_list=None, analyze_all=False):
        """proportion_conduction of each band picked by _analysis_bands, as
        {band: (v, c)} (projector.py:437-503); with ``return_energies`` also
        the bands' energies and occupations."""
        if min(num_below_ef, num_above_ef) < 0:
            raise ValueError("num_above_ef and num_below_ef must both be nonnegative.")
        bands = self._analysis_bands(num_below_ef, num_above_ef, vbmband, band_list, analyze_all)
        results = {b: self.proportion_conduction(b, spinpol=spinpol) for b in bands}
        return (results, self.wf._get_energy_list(bands)) if return_energies else results
