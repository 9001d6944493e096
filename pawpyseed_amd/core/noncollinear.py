"""Noncollinear (spinor) wavefunctions: the reference's NCLWavefunction
(noncollinear.py:4-31 over the CNCLWavefunction Cython base).

Each band holds two components of W plane-wave coefficients (up, then
down) over one list of W G vectors.  setup_projections projects the two
components separately (onto_projector_ncl, projector.c:359-392) into
band->up_projections and band->down_projections; get_projections returns
both.  Projection between structures (Projector) stays collinear-only, as
in the reference (projector.py:74-75)."""

import numpy as np

from ._lib import PAWpyError, lib
from .wavefunction import PWFPointer, Wavefunction


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
