"""Molecular basis: normalised contracted shells and the reference AO order.

Follows VeloxChem's semantics exactly, because the Fock matrix is only
comparable element by element when the AO basis is identical:

* contraction normalisation: ``CBasisFunction::normalize`` / ``_rescale`` /
  ``_overlap`` (src/orbdata/BasisFunction.cpp:129-143, 172-267);
* AO order: grouped by angular momentum L ascending, then by spherical
  component m, then by contracted function in atom order and in file order
  within an atom; AO index = offset_L + m * n_L + k
  (src/orbdata/MolecularBasis.cpp:400-410, 452-470);
* spherical component order p = (y, z, x), d = (xy, yz, z2, xz, x2-y2),
  f as src/general/SphericalMomentum.hpp:80-134.

The shell table produced by :meth:`MolecularBasis.shell_arrays` is the
structure-of-arrays input of the native Fock library.
"""

import json
import math
from pathlib import Path

import numpy as np

_BASIS_JSON = Path(__file__).with_name('basis') / 'basis_sets.json'
_BASIS_CACHE = None

_ALIASES = {
    'STO-3G': 'STO-3G',
    'DEF2-SVP': 'DEF2-SVP',
    'DEF2-SVPD': 'DEF2-SVPD',
    'DEF2-TZVP': 'DEF2-TZVP',
    'DEF2-QZVP': 'DEF2-QZVP',
    'AO-START-GUESS': 'AO-START-GUESS',
}

MAX_ANGULAR_MOMENTUM = 4  # basis data and one-electron integrals; the Fock build stops at f


def _basis_library():
    global _BASIS_CACHE
    if _BASIS_CACHE is None:
        _BASIS_CACHE = json.loads(_BASIS_JSON.read_text())
    return _BASIS_CACHE


def normalize_contraction(lang, exps, coefs):
    """Normalised contraction coefficients of one shell.

    Restates src/orbdata/BasisFunction.cpp:129-143 (normalize), :172-215
    (_rescale, per-L Cartesian factor) and :217-267 (_overlap, self overlap
    of the contracted function summed over the triangle i <= j).
    """
    exps = [float(e) for e in exps]
    norms = [float(c) for c in coefs]
    if len(exps) == 1:
        norms[0] = 1.0
    fpi = 2.0 / math.pi
    for i, a in enumerate(exps):
        norms[i] *= (a * fpi) ** 0.75
        if lang == 1:
            norms[i] *= 2.0 * math.sqrt(a)
        elif lang == 2:
            norms[i] *= 2.0 / math.sqrt(3.0) * a
        elif lang == 3:
            norms[i] *= 4.0 / math.sqrt(15.0) * a * math.sqrt(a)
        elif lang == 4:
            norms[i] *= 2.0 / math.sqrt(105.0) * a * a
        elif lang != 0:
            raise ValueError(f'normalize: L={lang} not supported')
    fact = 0.0
    for i in range(len(exps)):
        for j in range(i, len(exps)):
            fab = 1.0 / (exps[i] + exps[j])
            fovl = math.pi * fab
            fovl = norms[i] * norms[j] * fovl * math.sqrt(fovl)
            if i != j:
                fovl *= 2.0
            if lang == 0:
                fact += fovl
            elif lang == 1:
                fact += 0.5 * fab * fovl
            elif lang == 2:
                fact += 3.0 * fab * fab * fovl
            elif lang == 3:
                fact += 7.5 * fab * fab * fab * fovl
            elif lang == 4:
                fact += 420.0 * fab ** 4 * fovl
    fact = 1.0 / math.sqrt(fact)
    return [n * fact for n in norms]


class MolecularBasis:
    """Per-atom lists of normalised shells ``(L, exponents, norms)``."""

    def __init__(self, atom_shells, label):
        self._atom_shells = atom_shells
        self._label = label
        self._build_ao_map()

    @staticmethod
    def read(molecule, basis_name, basis_path='.', ostream=None):
        name = basis_name.upper()
        if name not in _ALIASES:
            raise ValueError(f'MolecularBasis: basis set {basis_name!r} is not packaged')
        lib = _basis_library()[_ALIASES[name]]
        atom_shells = []
        for lab in molecule.get_labels():
            key = lab.upper()
            if key not in lib:
                raise ValueError(f'MolecularBasis: no {name} basis for {lab}')
            shells = []
            for sh in lib[key]:
                if sh['l'] > MAX_ANGULAR_MOMENTUM:
                    raise ValueError('MolecularBasis: shells above g are not supported')
                shells.append((sh['l'], list(sh['exps']),
                               normalize_contraction(sh['l'], sh['exps'], sh['coefs'])))
            atom_shells.append(shells)
        basis = MolecularBasis(atom_shells, name)
        basis._coords = molecule.get_coordinates_in_bohr()
        basis._elem_ids = list(molecule.get_element_ids())
        return basis

    def reduce_to_valence_basis(self):
        """Keeps the shells with L up to the element's valence L (0 for H..Be, 1 for B..Ne).

        Restates CAtomBasis::reduce_to_valence_basis (src/orbdata/AtomBasis.cpp:135-152)
        with chem_elem::max_angular_momentum (src/moldata/ChemicalElement.cpp:70-81).
        """
        shells = []
        for z, atom in zip(self._elem_ids, self._atom_shells):
            mang = 0 if z < 5 else 1
            shells.append([sh for sh in atom if sh[0] <= mang])
        basis = MolecularBasis(shells, self._label + '(Valence)')
        basis._coords = self._coords
        basis._elem_ids = list(self._elem_ids)
        return basis

    def number_of_atom_basis_functions(self, atom, lang):
        """Contracted functions of angular momentum lang on one atom."""
        return sum(1 for sh in self._atom_shells[atom] if sh[0] == lang)

    def atom_ao_indices(self):
        """AO indices of each atom, in AO order (L, then component, then atom, then function).

        Same order as SadGuessDriver.get_ao_indices_of_atoms
        (src/pymodule/sadguessdriver.py:817-839).
        """
        natoms = len(self._atom_shells)
        out = [[] for _ in range(natoms)]
        ao = 0
        for lang in range(self.max_angular_momentum() + 1):
            for _ in range(2 * lang + 1):
                for a in range(natoms):
                    for _ in range(self.number_of_atom_basis_functions(a, lang)):
                        out[a].append(ao)
                        ao += 1
        return out

    def get_label(self):
        return self._label

    def _build_ao_map(self):
        # n_L: number of contracted functions of each L in the molecule
        self._n_l = [0] * (MAX_ANGULAR_MOMENTUM + 1)
        for shells in self._atom_shells:
            for lang, _, _ in shells:
                self._n_l[lang] += 1
        self._offset_l = [0] * (MAX_ANGULAR_MOMENTUM + 2)
        for lang in range(MAX_ANGULAR_MOMENTUM + 1):
            self._offset_l[lang + 1] = self._offset_l[lang] + self._n_l[lang] * (2 * lang + 1)
        self._nao = self._offset_l[MAX_ANGULAR_MOMENTUM + 1]
        # function index k of each shell within its L, atom order then file order
        counters = [0] * (MAX_ANGULAR_MOMENTUM + 1)
        self._shell_k = []
        for iatom, shells in enumerate(self._atom_shells):
            for lang, _, _ in shells:
                self._shell_k.append(counters[lang])
                counters[lang] += 1

    def get_dimensions_of_basis(self):
        return self._nao

    dimensions_of_basis = get_dimensions_of_basis

    def max_angular_momentum(self):
        return max((lang for shells in self._atom_shells for lang, _, _ in shells), default=0)

    def number_of_basis_functions(self, lang):
        return self._n_l[lang]

    def angular_block_dims(self):
        """AO offsets of the angular blocks: [0, end_s, end_p, ...] up to max L."""
        mlang = self.max_angular_momentum()
        return [self._offset_l[lang] for lang in range(mlang + 2)]

    def shell_arrays(self, coordinates=None):
        """Structure-of-arrays shell table for the native library.

        Returns a dict of numpy arrays: ``l``, ``nprim``, ``prim_offset``,
        ``exps``, ``norms``, ``centers`` (n_shells x 3, bohr), ``ao_base``
        (offset_L + k), ``ao_stride`` (n_L) and ``atom``.
        """
        coords = self._coords if coordinates is None else np.asarray(coordinates)
        ls, nprims, poff, exps, norms, centers, base, stride, atoms = [], [], [], [], [], [], [], [], []
        ishell = 0
        for iatom, shells in enumerate(self._atom_shells):
            for lang, e, n in shells:
                ls.append(lang)
                nprims.append(len(e))
                poff.append(len(exps))
                exps.extend(e)
                norms.extend(n)
                centers.append(coords[iatom])
                base.append(self._offset_l[lang] + self._shell_k[ishell])
                stride.append(self._n_l[lang])
                atoms.append(iatom)
                ishell += 1
        return {
            'l': np.array(ls, dtype=np.int32),
            'nprim': np.array(nprims, dtype=np.int32),
            'prim_offset': np.array(poff, dtype=np.int32),
            'exps': np.array(exps, dtype=np.float64),
            'norms': np.array(norms, dtype=np.float64),
            'centers': np.ascontiguousarray(np.array(centers, dtype=np.float64).reshape(-1, 3)),
            'ao_base': np.array(base, dtype=np.int32),
            'ao_stride': np.array(stride, dtype=np.int32),
            'atom': np.array(atoms, dtype=np.int32),
        }

    def number_of_shells(self):
        return sum(len(s) for s in self._atom_shells)
