"""Wavefunction: the reference's Python entry object, over the HIP C-ABI.

Mirrors pawpyseed/core/wavefunction.py (Wavefunction, CoreRegion,
Pseudopotential attribute names) and the Cython classes of
pawpyseed/core/pawpyc.pyx:197-555 (PWFPointer, PseudoWavefunction,
CWavefunction), calling libpawpy_amd.so through ctypes.  pymatgen is not
required: a structure is anything with ``lattice.matrix``, ``frac_coords``
and per-site species labels (``Structure`` below is a minimal one).
"""

from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import _lib
from ._lib import PAWpyError, as_f64, as_i32, c_double_p, c_int_p, check, lib, ptr


class Lattice:
    def __init__(self, matrix):
        self.matrix = np.asarray(matrix, np.float64).reshape(3, 3)

    @property
    def abc(self):
        return tuple(np.linalg.norm(self.matrix, axis=1))

    def __eq__(self, other):
        return np.allclose(self.matrix, other.matrix, atol=1e-10)


class Structure:
    """Minimal periodic structure: lattice rows, fractional coordinates,
    one species string per site."""

    def __init__(self, lattice, species, frac_coords):
        self.lattice = lattice if isinstance(lattice, Lattice) else Lattice(lattice)
        self.species = list(species)
        self.frac_coords = np.asarray(frac_coords, np.float64).reshape(-1, 3)

    def __len__(self):
        return len(self.species)

    def distance_matrix_to(self, other):
        """Minimum-image distances between the sites of self and other."""
        d = self.frac_coords[:, None, :] - other.frac_coords[None, :, :]
        d -= np.round(d)
        cart = d @ self.lattice.matrix
        best = np.linalg.norm(cart, axis=-1)
        # check neighbouring images for skewed cells
        for shift in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
            if not shift.any():
                continue
            c2 = (d + shift) @ self.lattice.matrix
            best = np.minimum(best, np.linalg.norm(c2, axis=-1))
        return best


class Pseudopotential:
    """Projector set of one element with the attribute names of the
    reference's POTCAR-derived class (wavefunction.py:17-137): ``ls``,
    ``realprojs`` on ``projgrid`` (= arange(n) * rmax / n), ``rmax``, and the
    partial waves ``aewaves``/``pswaves`` on the logarithmic ``grid``."""

    def __init__(self, ls, realprojs, rmax, grid, aewaves, pswaves):
        self.ls = list(int(l) for l in ls)
        self.realprojs = [np.asarray(p, np.float64) for p in realprojs]
        self.rmax = float(rmax)
        n = len(self.realprojs[0])
        self.ndata = n
        self.projgrid = np.arange(n) * self.rmax / n
        self.grid = np.asarray(grid, np.float64)
        self.aewaves = [np.asarray(a, np.float64) for a in aewaves]
        self.pswaves = [np.asarray(a, np.float64) for a in pswaves]

    @classmethod
    def from_species(cls, sp):
        return cls(sp.ls, sp.realprojs, sp.rmax, sp.grid, sp.aewaves, sp.pswaves)


class CoreRegion:
    """Element symbol -> Pseudopotential (wavefunction.py:143-166)."""

    def __init__(self, pps: dict):
        self.pps = dict(pps)


class PWFPointer:
    """Owner of a pswf_t* plus its k-points and weights (pawpyc.pyx:197-267)."""

    def __init__(self, ptr_value, kpts, weights, band_props=None):
        if not ptr_value:
            raise PAWpyError("NULL pswf_t: " + lib().pp_last_error().decode(errors="replace"))
        self.ptr = ct.c_void_p(ptr_value)
        self.kpts = np.asarray(kpts, np.float64).reshape(-1, 3)
        self.weights = np.asarray(weights, np.float64)
        self.band_props = np.zeros(4) if band_props is None else np.asarray(band_props)

    @classmethod
    def from_wavecar(cls, filename, kpts, weights, band_props=None):
        if not os.path.isfile(filename):
            raise FileNotFoundError(f"File {filename} does not exist.")
        w = as_f64(weights)
        p = lib().read_wavefunctions(str(filename).encode(), ptr(w, c_double_p))
        return cls(p, kpts, weights, band_props)

    @classmethod
    def from_arrays(cls, lattice, kpts, weights, Gs, coeffs, encut, nspin=1, occs=None, energies=None):
        """Gs: per k-index int arrays [W_k, 3]; coeffs: per k-index complex64
        [nband, W_k]; k-index = k + spin * nwk."""
        nkt = len(coeffs)
        nwk = nkt // nspin
        nband = coeffs[0].shape[0]
        nplane = as_i32([c.shape[1] for c in coeffs])
        G = as_i32(np.concatenate([np.asarray(g).reshape(-1, 3) for g in Gs]))
        C = np.ascontiguousarray(np.concatenate([np.asarray(c, np.complex64).ravel() for c in coeffs]))
        occ = as_f64(np.zeros((nkt, nband)) if occs is None else occs)
        en = as_f64(np.zeros((nkt, nband)) if energies is None else energies)
        lat = as_f64(lattice)
        kp = as_f64(kpts)
        w = as_f64(weights)
        p = lib().pp_pswf_from_arrays(nspin, nwk, nband, float(encut), ptr(lat, c_double_p), ptr(kp, c_double_p),
                                      ptr(w, c_double_p), ptr(nplane, c_int_p), ptr(G, c_int_p),
                                      C.ctypes.data_as(ct.c_void_p), ptr(occ, c_double_p), ptr(en, c_double_p))
        return cls(p, kpts, weights)


class Wavefunction:
    """PAW wavefunction of a structure (wavefunction.py:169-430 and the
    CWavefunction/PseudoWavefunction Cython bases)."""

    def __init__(self, struct, pwf: PWFPointer, cr: CoreRegion, dim, symprec=1e-4, setup_projectors=False):
        # checked before the object takes the pointer, so a refused one is
        # not freed here
        if lib().is_ncl(pwf.ptr) > 0:
            raise PAWpyError("Pseudowavefunction is noncollinear! Call NCLWavefunction(...) instead")
        self._setup(struct, pwf, cr, dim, symprec)
        if setup_projectors:
            self.check_c_projectors()

    def _setup(self, struct, pwf, cr, dim, symprec):
        self.band_props = pwf.band_props.copy()
        self._pwf = pwf
        self.wf_ptr = pwf.ptr
        self.kpts = pwf.kpts.copy()
        self.kws = pwf.weights.copy()
        L = lib()
        self.ncl = L.is_ncl(self.wf_ptr) > 0
        self.nband = L.get_nband(self.wf_ptr)
        self.nwk = L.get_nwk(self.wf_ptr)
        self.nspin = L.get_nspin(self.wf_ptr)
        self.encut = L.get_encut(self.wf_ptr)
        self.structure = struct
        self.symprec = symprec
        self.cr = cr
        self.dim = np.array(dim).astype(np.int32)
        if len(self.dim) != 3:
            raise PAWpyError("Grid dimensions must be length 3")
        self.projector_owner = 0

    # ------------------------------------------------------------ factories
    @staticmethod
    def from_arrays(struct, cr, dim, kpts, kws, Gs, coeffs, encut, nspin=1, occs=None, energies=None,
                    setup_projectors=False):
        pwf = PWFPointer.from_arrays(struct.lattice.matrix, kpts, kws, Gs, coeffs, encut, nspin, occs, energies)
        return Wavefunction(struct, pwf, cr, dim, setup_projectors=setup_projectors)

    @staticmethod
    def from_wavecar(struct, wavecar, cr, dim, kpts, kws, setup_projectors=False):
        pwf = PWFPointer.from_wavecar(wavecar, kpts, kws)
        return Wavefunction(struct, pwf, cr, dim, setup_projectors=setup_projectors)

    # ------------------------------------------------------------ checks
    def check_band_index(self, b):
        if b < 0 or b >= self.nband:
            raise ValueError("Invalid band {}. Should be in range [{}, {}]".format(b, 0, self.nband - 1))

    def check_kpoint_index(self, k):
        if k < 0 or k >= self.nwk:
            raise ValueError("Invalid kpoint index {}. Should be in range [{}, {}]".format(k, 0, self.nwk - 1))

    def check_spin_index(self, s):
        if s < 0 or s >= self.nspin:
            raise ValueError("Spin must be 0 for non-spin-polarized or 0 or 1 for spin-polarized.")

    def check_bks_spec(self, b, k, s):
        self.check_band_index(b)
        self.check_kpoint_index(k)
        self.check_spin_index(s)

    # ------------------------------------------------------------ projectors
    def _element_labels(self):
        labels = {}
        for i, e in enumerate(self.cr.pps):
            labels[e] = i
        return labels

    def _make_c_projectors(self):
        """wavefunction.py:386-415 + pawpyc.pyx:352-412: element labels,
        projector arrays, get_projector_list, then setup_projections."""
        labels = self._element_labels()
        pps = [self.cr.pps[e] for e in labels]
        nums = as_i32([labels[s] for s in self.structure.species])
        coords = as_f64(self.structure.frac_coords.ravel())
        self.num_sites = len(nums)
        self.num_elems = len(pps)
        abc = np.array(self.structure.lattice.abc)
        grid_encut = float(np.max((np.pi * self.dim / abc) ** 2 / 0.262))
        clabels = as_i32(np.concatenate([[i, len(pp.ls), len(pp.realprojs[0]), len(pp.grid)]
                                         for i, pp in enumerate(pps)]))
        ls = as_i32(np.concatenate([pp.ls for pp in pps]))
        wgrids = as_f64(np.concatenate([pp.grid for pp in pps]))
        projectors = as_f64(np.concatenate([np.concatenate(pp.realprojs) for pp in pps]))
        aewaves = as_f64(np.concatenate([np.concatenate(pp.aewaves) for pp in pps]))
        pswaves = as_f64(np.concatenate([np.concatenate(pp.pswaves) for pp in pps]))
        rmaxs = as_f64([pp.rmax for pp in pps])
        L = lib()
        pl = L.get_projector_list(self.num_elems, ptr(clabels, c_int_p), ptr(ls, c_int_p), ptr(wgrids, c_double_p),
                                  ptr(projectors, c_double_p), ptr(aewaves, c_double_p), ptr(pswaves, c_double_p),
                                  ptr(rmaxs, c_double_p), grid_encut)
        if not pl:
            raise PAWpyError("get_projector_list failed: " + L.pp_last_error().decode(errors="replace"))
        self.nums = nums
        self.coords = coords
        self.update_dimv(self.dim)
        if getattr(self, "_paw_override", None) is not None:
            self._inject_paw_overlaps(pl, pps)
        # setup_projections takes ownership of the ppot list (projector.c:570)
        check(L.setup_projections(self.wf_ptr, ct.c_void_p(pl), self.num_elems, self.num_sites,
                                  ptr(self.dimv, c_int_p), ptr(nums, c_int_p), ptr(coords, c_double_p)),
              "setup_projections")
        self.projector_owner = 1

    def set_paw_overlap(self, element, delta_o):
        """Use ``delta_o`` (channel x channel) as aepw - pspw for ``element``
        instead of the partial-wave integrals (written into the ppot_t
        overlap matrices after get_projector_list)."""
        if getattr(self, "_paw_override", None) is None:
            self._paw_override = {}
        self._paw_override[element] = np.asarray(delta_o, np.float64)

    def _inject_paw_overlaps(self, pl, pps):
        from .ppot import ppot_array
        arr = ppot_array(pl, len(pps))
        for i, e in enumerate(self._element_labels()):
            if e in self._paw_override:
                d = self._paw_override[e]
                n = arr[i].num_projs
                if d.shape != (n, n):
                    raise PAWpyError(f"delta_o for {e} must be {n}x{n}")
                for a in range(n):
                    for b in range(n):
                        arr[i].aepw_overlap_matrix[a * n + b] = d[a, b]
                        arr[i].pspw_overlap_matrix[a * n + b] = 0.0

    def check_c_projectors(self):
        if not self.projector_owner:
            self._make_c_projectors()

    def get_projections(self, band, k, spin=0):
        """<p_i|psi_{band,k,spin}> for all projectors of all sites (c128);
        for a spinor wavefunction shape (2, P): up, then down component."""
        self.check_bks_spec(band, k, spin)
        ncomp = 2 if self.ncl else 1
        out = np.zeros(ncomp * int(self._ptot()), np.complex128)
        check(lib().pp_get_projections(self.wf_ptr, band, k + spin * self.nwk, out.ctypes.data_as(ct.c_void_p)),
              "pp_get_projections")
        return out.reshape(2, -1) if self.ncl else out

    def _ptot(self):
        labels = self._element_labels()
        return sum(sum(2 * l + 1 for l in self.cr.pps[s].ls) for s in self.structure.species) if labels else 0

    # ------------------------------------------------------------ real space
    def update_dim(self, dim):
        """wavefunction.py:245-247: grid of later real-space calls."""
        self.dim = np.array(dim, dtype=np.int32)
        self.update_dimv(dim)

    def update_dimv(self, dim):
        """pawpyc.pyx:414-419: the state grid and the doubled density grid."""
        self.dimv = as_i32(dim)
        self.fdimv = as_i32(np.asarray(dim) * 2)
        self.gridsize = int(np.prod(self.dimv))
        self.fgridsize = int(np.prod(self.fdimv))

    def get_state_realspace(self, b, k, s, dim=None, remove_phase=False):
        """All-electron state of band b, k-point k, spin s on the grid dim
        (default: the projection grid) as a complex array [N0, N1, N2]
        (wavefunction.py:431-447, density.c:232-310, on the GPU)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim))
        return self._get_realspace_state(b, k, s, remove_phase)

    def get_state_realspace_density(self, b, k, s, dim=None):
        """|psi|^2 of one all-electron state on the doubled grid (dim // 2 * 2)
        (wavefunction.py:449-465, density.c:28-39)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim) // 2)
        return self._get_realspace_state_density(b, k, s)

    def get_realspace_density(self, dim=None, bands=None):
        """All-electron charge density sum_k,b w_k occ (2 / nspin) |psi|^2 on
        the doubled grid (wavefunction.py:467-479, density.c:158-179).  The
        reference accepts ``bands`` here without using it; this one passes
        it on (an int or a list: only those bands, weighted w_k / nspin, as
        pawpyc.pyx:462-482 does for the files)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim) // 2)
        return self._get_realspace_density(bands)

    def _rs_check(self, b, k, s):
        if b < 0 or b >= self.nband:
            raise ValueError("Invalid band choice")
        if k < 0 or k >= self.nwk:
            raise ValueError("Invalid k-point choice")
        if s < 0 or s >= self.nspin:
            raise ValueError("Invalid spin choice")

    def _get_realspace_state(self, b, k, s, remove_phase=False):
        """pawpyc.pyx:425-439."""
        self._rs_check(b, k, s)
        res = np.zeros(self.gridsize, np.complex128)
        L = lib()
        check(L.realspace_state(res.ctypes.data_as(ct.c_void_p), b, k + s * self.nwk, self.wf_ptr,
                                ptr(self.dimv, c_int_p), ptr(self.nums, c_int_p), ptr(self.coords, c_double_p)),
              "realspace_state")
        if remove_phase:
            check(L.remove_phase(res.ctypes.data_as(ct.c_void_p), k + s * self.nwk, self.wf_ptr,
                                 ptr(self.dimv, c_int_p)), "remove_phase")
        return res.reshape(tuple(self.dimv))

    def _get_realspace_state_density(self, b, k, s):
        """pawpyc.pyx:441-453."""
        self._rs_check(b, k, s)
        res = np.zeros(self.fgridsize, np.float64)
        check(lib().ae_state_density(ptr(res, c_double_p), b, k + s * self.nwk, self.wf_ptr,
                                     ptr(self.fdimv, c_int_p), ptr(self.nums, c_int_p),
                                     ptr(self.coords, c_double_p)), "ae_state_density")
        return res.reshape(tuple(self.fdimv))

    def _get_realspace_density(self, bands=None):
        """pawpyc.pyx:455-494."""
        res = np.zeros(self.fgridsize, np.float64)
        L = lib()
        args = (self.wf_ptr, ptr(self.fdimv, c_int_p), ptr(self.nums, c_int_p), ptr(self.coords, c_double_p))
        if bands is None:
            check(L.ae_chg_density(ptr(res, c_double_p), *args), "ae_chg_density")
        else:
            blist = [bands] if isinstance(bands, (int, np.integer)) else list(bands)
            for b in blist:
                if b < 0 or b >= self.nband:
                    raise ValueError("Invalid band choice")
                for k in range(self.nwk * self.nspin):
                    work = np.zeros(self.fgridsize, np.float64)
                    check(L.ae_state_density(ptr(work, c_double_p), int(b), k, *args), "ae_state_density")
                    res += work * self.kws[k % self.nwk] / self.nspin
        return res.reshape(tuple(self.fdimv))

    def _convert_to_vasp_volumetric(self, filename, dim):
        """wavefunction.py:480-510: a POSCAR-style header and the grid size
        in front of the values."""
        latt = np.asarray(self.structure.lattice.matrix, np.float64)
        lines = filename + "\n" + "   1.00000000000000\n"
        for r in range(3):
            lines += " %12.6f%12.6f%12.6f\n" % tuple(latt[r])
        syms, counts = [], []
        for sp in self.structure.species:
            if syms and syms[-1] == str(sp):
                counts[-1] += 1
            else:
                syms.append(str(sp))
                counts.append(1)
        lines += "".join("%5s" % x for x in syms) + "\n"
        lines += "".join("%6d" % x for x in counts) + "\n"
        lines += "Direct\n"
        for fc in self.structure.frac_coords:
            lines += "%10.6f%10.6f%10.6f\n" % tuple(fc)
        lines += " \n"
        with open(filename) as f:
            nums = f.read()
        with open(filename, "w") as f:
            f.write(lines + "%d %d %d\n" % (dim[0], dim[1], dim[2]) + nums)

    def write_state_realspace(self, b, k, s, fileprefix="", dim=None, scale=1, remove_phase=False):
        """Real and imaginary parts of one state, in two VASP-style
        volumetric files (wavefunction.py:512-549)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim))
        base = "%sB%dK%dS%d" % (fileprefix, b, k, s)
        f1, f2 = "%s_REAL.vasp" % base, "%s_IMAG.vasp" % base
        res = self._get_realspace_state(b, k, s, remove_phase)
        write_volumetric(f1, np.real(res), self.dimv, scale)
        write_volumetric(f2, np.imag(res), self.dimv, scale)
        self._convert_to_vasp_volumetric(f1, self.dim)
        self._convert_to_vasp_volumetric(f2, self.dim)
        return res

    def write_density_realspace(self, filename="PYAECCAR.vasp", dim=None, scale=1, bands=None):
        """The all-electron charge density in a VASP-style volumetric file
        (wavefunction.py:551-580)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim) // 2)
        res = self._get_realspace_density(bands)
        write_volumetric(filename, res, self.fdimv, scale)
        self._convert_to_vasp_volumetric(filename, self.dim * 2)
        return res

    # ------------------------------------------------------------ overlaps
    def pseudoprojection(self, band_num, basis, flip_spin=False):
        """<psibt_n1k|psit_n2k> for all n1, k of basis and n2 = band_num of
        self (pawpyc.pyx:311-325)."""
        res = np.zeros(basis.nband * basis.nwk * basis.nspin, dtype=np.complex128)
        check(lib().pseudoprojection(res.ctypes.data_as(ct.c_void_p), basis.wf_ptr, self.wf_ptr, int(band_num),
                                     int(bool(flip_spin))), "pseudoprojection")
        return res

    def _get_occs(self):
        L = lib()
        p = L.get_occs(self.wf_ptr)
        n = self.nband * self.nwk * self.nspin
        out = np.ctypeslib.as_array(p, shape=(n,)).copy()
        L.free_ptr(ct.cast(p, ct.c_void_p))
        return out

    def _get_energy_list(self, bands):
        for b in bands:
            if b < 0 or b >= self.nband:
                raise ValueError("Invalid band choice")
        L = lib()
        out = {}
        for b in bands:
            out[b] = []
            for s in range(self.nspin):
                for k in range(self.nwk):
                    out[b].append([L.get_energy(self.wf_ptr, b, k, s), L.get_occ(self.wf_ptr, b, k, s)])
        return out

    def free(self):
        if getattr(self, "wf_ptr", None):
            lib().free_pswf(self.wf_ptr)
            self.wf_ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def write_volumetric(filename, x, dim, scale):
    """density.c:461-477: values x[i][j][k] * scale, z slowest, "%E   " each,
    five to a line."""
    v = np.asarray(x, np.float64).reshape(tuple(int(n) for n in dim)).transpose(2, 1, 0).ravel() * scale
    with open(filename, "w") as f:
        for r in range(0, len(v), 5):
            row = v[r:r + 5]
            f.write("".join("%E   " % t for t in row))
            if len(row) == 5:
                f.write("\n")


def wavefunction_from_workload(wl, which="R", nband=None, coeffs=None, setup_projectors=False, element="C",
                               paw_overlap=True, occs=None):
    """Wavefunction of one structure of a synthetic workload (tests, smoke,
    bench).  With ``paw_overlap`` the workload's channel matrix replaces the
    partial-wave overlap difference; without it the matrices are those
    get_projector_list integrates from the species' partial waves.  ``occs``:
    optional occupations [nk][nband]."""
    frac = wl.frac_R if which == "R" else wl.frac_S
    struct = Structure(wl.lattice, [element] * len(frac), frac)
    cr = CoreRegion({element: Pseudopotential.from_species(wl.species)})
    Gs = wl.ensure_gvecs()
    C = coeffs if coeffs is not None else wl.coefficients(which, nband=nband)
    wf = Wavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, C, wl.encut, occs=occs)
    if paw_overlap:
        wf.set_paw_overlap(element, wl.delta_o)
    if setup_projectors:
        wf.check_c_projectors()
    return wf
