"""Molecule: element labels and Cartesian coordinates held in bohr.

Mirrors the parts of VeloxChem's ``Molecule`` that the Fock build consumes
(src/pymodule/molecule.py:268 ``read_str``, src/moldata/Molecule.cpp): atom
order is input order, coordinates are stored in bohr, and Angstrom input is
divided by the CODATA 2018 Bohr radius 0.529177210903 Angstrom
(src/general/Codata.cpp:33-38).
"""

import numpy as np

BOHR_IN_ANGSTROM = 0.529177210903

ELEMENTS = ['H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne']


def bohr_in_angstrom():
    return BOHR_IN_ANGSTROM


class Molecule:
    """Atoms (labels, nuclear charges) and coordinates in bohr."""

    def __init__(self, labels, coordinates, units='angstrom'):
        labels = [str(lab).strip().capitalize() for lab in labels]
        for lab in labels:
            if lab not in ELEMENTS:
                raise ValueError(f'Molecule: unsupported element {lab!r}')
        coords = np.array(coordinates, dtype=np.float64).reshape(-1, 3)
        if coords.shape[0] != len(labels):
            raise ValueError('Molecule: labels and coordinates differ in length')
        units = units.lower()
        if units in ('angstrom', 'angs', 'a'):
            coords = coords / BOHR_IN_ANGSTROM
        elif units not in ('bohr', 'au', 'a.u.'):
            raise ValueError(f'Molecule: unknown units {units!r}')
        self._labels = labels
        self._coords = np.ascontiguousarray(coords)
        self._charge = 0

    @staticmethod
    def read_str(mol_str, units='angstrom'):
        labels, coords = [], []
        for line in mol_str.strip().splitlines():
            tok = line.split()
            if not tok:
                continue
            labels.append(tok[0])
            coords.append([float(x) for x in tok[1:4]])
        return Molecule(labels, coords, units)

    read_molecule_string = read_str

    @staticmethod
    def read_xyz_string(xyz_str):
        lines = [ln for ln in xyz_str.strip().splitlines()]
        natoms = int(lines[0].split()[0])
        return Molecule.read_str('\n'.join(lines[2:2 + natoms]), 'angstrom')

    def number_of_atoms(self):
        return len(self._labels)

    def get_labels(self):
        return list(self._labels)

    def get_element_ids(self):
        return [ELEMENTS.index(lab) + 1 for lab in self._labels]

    def get_coordinates_in_bohr(self):
        return self._coords.copy()

    def get_coordinates_in_angstrom(self):
        return self._coords * BOHR_IN_ANGSTROM

    def set_charge(self, charge):
        self._charge = int(charge)

    def get_charge(self):
        return self._charge

    def number_of_electrons(self):
        return sum(self.get_element_ids()) - self._charge

    def set_multiplicity(self, multiplicity):
        self._multiplicity = int(multiplicity)

    def get_multiplicity(self):
        return getattr(self, '_multiplicity', 1)

    def number_of_alpha_electrons(self):
        return (self.number_of_electrons() + self.get_multiplicity() - 1) // 2

    def number_of_beta_electrons(self):
        return (self.number_of_electrons() - self.get_multiplicity() + 1) // 2

    def nuclear_repulsion_energy(self):
        z = np.array(self.get_element_ids(), dtype=np.float64)
        r = self._coords
        e = 0.0
        for i in range(len(z)):
            for j in range(i):
                e += z[i] * z[j] / np.linalg.norm(r[i] - r[j])
        return e
