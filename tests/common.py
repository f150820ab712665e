"""Shared inputs of the tests (molecules of the reference's own tests)."""
from pathlib import Path

import numpy as np

from veloxchem_amd import MolecularBasis, Molecule, SubMatrix

GOLDEN = Path(__file__).resolve().parent / 'golden'

# tests/test_fock_driver.py:48-57 (STO-3G water)
H2O_STO3G = """O 0.000000000000 0.000000000000 0.000000000000
H 0.000000000000 0.740848095288 0.582094932012
H 0.000000000000 -0.740848095288 0.582094932012"""

# tests/test_fock_driver.py:17-28 (water dimer)
H2O_DIMER = """O -1.464  0.099  0.300
H -1.956  0.624 -0.340
H -1.797 -0.799  0.206
O  1.369  0.146 -0.395
H  1.894  0.486  0.335
H  0.451  0.163 -0.083"""


def load(name):
    return np.load(GOLDEN / name)


def sto3g_water():
    mol = Molecule.read_str(H2O_STO3G, 'angstrom')
    return mol, MolecularBasis.read(mol, 'sto-3g')


def dimer(basis):
    mol = Molecule.read_str(H2O_DIMER, 'angstrom')
    return mol, MolecularBasis.read(mol, basis)


def reference_equal(a, b):
    """SubMatrix equality of the reference: |a-b| < max(1e-12, 1e-12 max(|a|,|b|))."""
    n, m = np.asarray(a).shape
    return SubMatrix([0, 0, n, m], a) == SubMatrix([0, 0, n, m], b)
