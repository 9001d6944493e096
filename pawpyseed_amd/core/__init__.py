"""MI355X projection path of pawpyseed: the reference's Python API
(Wavefunction, Projector) over libpawpy_amd.so."""

from ._lib import PAWpyError, lib  # noqa: F401
from .noncollinear import NCLWavefunction  # noqa: F401
from .projector import Projector  # noqa: F401
from .wavefunction import (CoreRegion, Lattice, Pseudopotential, PWFPointer, Structure,  # noqa: F401
                           Wavefunction, wavefunction_from_workload)
