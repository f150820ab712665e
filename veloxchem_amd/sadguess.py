"""Superposition-of-atomic-densities (SAD) initial guess.

Follows SadGuessDriver.compute (src/pymodule/sadguessdriver.py:841-971):
per atom, the minimal-basis (AO-START-GUESS) functions are projected onto the
atom's AO-basis functions with the overlap matrices S12 (minimal x AO) and
S22 (AO x AO), orthonormalised (C2 = S22^-1 A (A^T S22^-1 A)^-1/2 with
A = S12^T), weighted by sqrt of the atomic occupation numbers, and the
closed-shell alpha density is D = C C^T.

Occupation numbers per element restate
SadGuessDriver.get_occupation_numbers_for_element (:402-470) for H..Ne.  The
reference shifts them by EEQ partial charges (src/moldata/PartialCharges.cpp);
this guess distributes the net charge over the atoms in proportion to the
nuclear charge instead (zero shift for neutral molecules is not identical to
the reference, whose EEQ charges are non-zero there).  The guess only sets
the SCF starting point; the converged energy does not depend on it.
"""

import numpy as np

from .oneeints import _one_electron


def _overlap_joint(molecule, basis, ket_basis):
    return _one_electron(molecule, basis, ket_basis=ket_basis, need_v=False)[0]


def occupation_numbers(elem_id, nelec):
    """Alpha occupations of the minimal-basis functions of one atom (H..Ne)."""
    if elem_id in (1, 2):
        return [max(0.0, 0.5 * elem_id + nelec)]
    if elem_id in (3, 4):
        return [1.0, max(0.0, 0.5 * (elem_id - 2) + nelec)]
    if 5 <= elem_id <= 10:
        occ = max(0.0, 0.125 * (elem_id - 2) + nelec / 4.0)
        return [1.0, occ, occ, occ, occ]
    raise ValueError(f'SAD guess: element {elem_id} is not supported')


def sad_density(molecule, min_basis, ao_basis):
    """Closed-shell alpha density (nao x nao) of the SAD guess."""
    # one pass over the joint (minimal + AO) shell table gives S12 and S22
    S = _overlap_joint(molecule, min_basis, ao_basis)
    n1 = min_basis.get_dimensions_of_basis()
    S12, S22 = S[:n1, n1:], S[n1:, n1:]
    nao_1, nao_2 = S12.shape
    ao1 = min_basis.atom_ao_indices()
    ao2 = ao_basis.atom_ao_indices()
    ids = molecule.get_element_ids()
    zsum = float(sum(ids))
    charge = molecule.get_charge()
    csad = np.zeros((nao_2, nao_1))
    for a, z in enumerate(ids):
        i1, i2 = ao1[a], ao2[a]
        occ = np.array(occupation_numbers(z, -0.5 * charge * z / zsum))
        if len(occ) != len(i1):
            raise ValueError('SAD guess: minimal basis does not match the occupation table')
        mat_a = S12[np.ix_(i1, i2)].T
        w, v = np.linalg.eigh(S22[np.ix_(i2, i2)])
        b22_inv = (v / w) @ v.T
        m = mat_a.T @ b22_inv @ mat_a
        w, v = np.linalg.eigh(m)
        m_invsqrt = (v / np.sqrt(w)) @ v.T
        c2 = b22_inv @ mat_a @ m_invsqrt
        csad[np.ix_(i2, i1)] = c2 * np.sqrt(occ)[None, :]
    return csad @ csad.T
