"""Noncollinear (spinor) wavefunctions: the reference's NCLWavefunction
(noncollinear.py:4-31 over the CNCLWavefunction Cython base).

Each band holds two components of W plane-wave coefficients (up, then
down) over one list of W G vectors.  setup_projections projects the two
components separately (onto_projector_ncl, projector.c:359-392) into
band->up_projections and band->down_projections; get_projections returns
both.  Projection between structures (Projector) stays collinear-only, as
in the reference (projector.py:74-75)."""

import ctypes as ct

import numpy as np

from ._lib import PAWpyError, c_double_p, c_int_p, check, lib, ptr
from .wavefunction import PWFPointer, Wavefunction, write_volumetric


class NCLWavefunction(Wavefunction):
    def __init__(self, struct, pwf: PWFPointer, cr, dim, symprec=1e-4, setup_projectors=False):
        if lib().is_ncl(pwf.ptr) == 0:
            raise PAWpyError("Pseudowavefunction is collinear! Call Wavefunction(...) instead")
        self._setup(struct, pwf, cr, dim, symprec)
        if setup_projectors:
            self.check_c_projectors()

    @staticmethod
    def from_arrays(struct, cr, dim, kpts, kws, Gs, coeffs, encut, occs=None, energies=None,
                    setup_projectors=False):
        """Gs: per k int [W_k, 3]; coeffs: per k complex64 [nband, 2 W_k]
        (up component, then down)."""
        G2 = [np.concatenate([np.asarray(g).reshape(-1, 3)] * 2) for g in Gs]
        pwf = PWFPointer.from_arrays(struct.lattice.matrix, kpts, kws, G2, coeffs, encut, 1, occs, energies)
        return NCLWavefunction(struct, pwf, cr, dim, setup_projectors=setup_projectors)

    @staticmethod
    def from_wavecar(struct, wavecar, cr, dim, kpts, kws, setup_projectors=False):
        pwf = PWFPointer.from_wavecar(wavecar, kpts, kws)
        return NCLWavefunction(struct, pwf, cr, dim, setup_projectors=setup_projectors)

    # ------------------------------------------------------------ real space
    def get_realspace_density(self, dim=None):
        """Charge density of both spinor components on the state grid
        (noncollinear.py:149-171 with pawpyc.pyx:582-588, density.c:181-203)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim))
        return self._get_realspace_density()

    def _get_realspace_state(self, b, k, s, remove_phase=False):
        """pawpyc.pyx:563-580: the (up, down) components of one state."""
        self._rs_check(b, k, s)
        res = np.zeros(2 * self.gridsize, np.complex128)
        L = lib()
        check(L.ncl_realspace_state(res.ctypes.data_as(ct.c_void_p), b, k + s * self.nwk, self.wf_ptr,
                                    ptr(self.dimv, c_int_p), ptr(self.nums, c_int_p),
                                    ptr(self.coords, c_double_p)), "ncl_realspace_state")
        if remove_phase:
            for half in (res[:self.gridsize], res[self.gridsize:]):
                check(L.remove_phase(half.ctypes.data_as(ct.c_void_p), k + s * self.nwk, self.wf_ptr,
                                     ptr(self.dimv, c_int_p)), "remove_phase")
        shape = tuple(self.dimv)
        return res[:self.gridsize].reshape(shape), res[self.gridsize:].reshape(shape)

    def _get_realspace_density(self, bands=None):
        """pawpyc.pyx:582-588 (bands is not used for spinors)."""
        res = np.zeros(self.gridsize, np.float64)
        check(lib().ncl_ae_chg_density(ptr(res, c_double_p), self.wf_ptr, ptr(self.dimv, c_int_p),
                                       ptr(self.nums, c_int_p), ptr(self.coords, c_double_p)), "ncl_ae_chg_density")
        return res.reshape(tuple(self.dimv))

    def get_state_realspace_density(self, b, k, s, dim=None):
        raise NotImplementedError("the reference has no spinor state density (density.c:15-25 is commented out)")

    def write_state_realspace(self, b, k, s, fileprefix="", dim=None, scale=1, remove_phase=False):
        """Real and imaginary parts of both components in four files
        (noncollinear.py:98-147)."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim))
        base = "%sB%dK%dS%d" % (fileprefix, b, k, s)
        names = ["%s_UP_REAL.vasp" % base, "%s_UP_IMAG.vasp" % base, "%s_DOWN_REAL.vasp" % base,
                 "%s_DOWN_IMAG.vasp" % base]
        res0, res1 = self._get_realspace_state(b, k, s, remove_phase)
        for name, part in zip(names, (np.real(res0), np.imag(res0), np.real(res1), np.imag(res1))):
            write_volumetric(name, part, self.dimv, scale)
            self._convert_to_vasp_volumetric(name, self.dim)
        return res0, res1

    def write_density_realspace(self, filename="PYAECCAR.vasp", dim=None, scale=1):
        """noncollinear.py:149-173."""
        self.check_c_projectors()
        if dim is not None:
            self.update_dim(np.array(dim))
        res = self._get_realspace_density()
        write_volumetric(filename, res, self.dimv, scale)
        self._convert_to_vasp_volumetric(filename, self.dim)
        return res
