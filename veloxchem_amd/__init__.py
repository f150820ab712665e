"""MI355X-native two-electron Fock build with VeloxChem's FockDriver interface.

The hot path -- screening, electron-repulsion integrals, J/K digestion --
runs in libvlxfock.so (HIP, gfx950).  This package holds the host-side
mirror of the reference's Python surface: Molecule, MolecularBasis,
Matrix/make_matrix/mat_t, T4CScreener and FockDriver.
"""

from .fockdriver import FockDriver
from .matrix import Matrices, Matrix, SubMatrix, make_matrix, mat_t
from .molecularbasis import MolecularBasis
from .molecule import Molecule, bohr_in_angstrom
from .screener import T4CScreener


def mpi_master():
    return 0


__all__ = ['FockDriver', 'Matrices', 'Matrix', 'SubMatrix', 'make_matrix', 'mat_t', 'MolecularBasis',
           'Molecule', 'bohr_in_angstrom', 'T4CScreener', 'mpi_master']
