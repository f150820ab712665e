"""One-electron integral matrices S, T and V (native host code in libvlxfock.so).

Mirrors src/pymodule/oneeints.py:33-95 of the reference
(compute_overlap_integrals, compute_kinetic_energy_integrals,
compute_nuclear_potential_integrals, which includes the electron charge) and
the driver classes bound at src/python/ExportT2CIntegrals.cpp:55-105
(OverlapDriver, KineticEnergyDriver, NuclearPotentialDriver; their
``compute(molecule, basis)`` returns a symmetric Matrix, the nuclear
potential without the electron charge).  Shells up to g.
"""

import numpy as np

from . import _native
from .matrix import SubMatrix, make_matrix, mat_t


def _shell_table(molecule, basis, ket_basis=None):
    """Shell table of basis; with ket_basis, the ket shells follow with AO indices shifted by nao(basis)."""
    st = basis.shell_arrays(molecule.get_coordinates_in_bohr())
    nao = basis.get_dimensions_of_basis()
    if ket_basis is None:
        return st, nao
    kt = ket_basis.shell_arrays(molecule.get_coordinates_in_bohr())
    st = dict(st)
    for key in ('l', 'nprim', 'exps', 'norms', 'ao_stride', 'atom'):
        st[key] = np.concatenate([st[key], kt[key]])
    st['prim_offset'] = np.concatenate([st['prim_offset'], kt['prim_offset'] + len(basis.shell_arrays()['exps'])])
    st['ao_base'] = np.concatenate([st['ao_base'], kt['ao_base'] + nao])
    st['centers'] = np.ascontiguousarray(np.vstack([st['centers'], kt['centers']]))
    for key in ('l', 'nprim', 'prim_offset', 'ao_base', 'ao_stride'):
        st[key] = np.ascontiguousarray(st[key], dtype=np.int32)
    return st, nao + ket_basis.get_dimensions_of_basis()


def _one_electron(molecule, basis, charges=None, coordinates=None, ket_basis=None, need_v=True):
    st, nao = _shell_table(molecule, basis, ket_basis)
    if not need_v:
        charges, coordinates = np.zeros(0), np.zeros((0, 3))
    elif charges is None and coordinates is None:
        charges = np.array(molecule.get_element_ids(), dtype=np.float64)
        coordinates = molecule.get_coordinates_in_bohr()
    elif charges is None or coordinates is None:
        raise ValueError('compute_nuclear_potential_integrals: give both charges and coordinates')
    charges = np.ascontiguousarray(charges, dtype=np.float64).ravel()
    coordinates = np.ascontiguousarray(coordinates, dtype=np.float64).reshape(-1, 3)
    if charges.shape[0] != coordinates.shape[0]:
        raise ValueError('compute_nuclear_potential_integrals: charges and coordinates differ in length')
    S = np.zeros((nao, nao))
    T = np.zeros((nao, nao))
    V = np.zeros((nao, nao))
    ip = lambda a: a.ctypes.data_as(_native._c_int_p)  # noqa: E731
    dp = lambda a: a.ctypes.data_as(_native._c_double_p)  # noqa: E731
    _native.check(_native.lib().vlx_one_electron(
        len(st['l']), ip(st['l']), ip(st['nprim']), ip(st['prim_offset']), dp(st['exps']), dp(st['norms']),
        dp(st['centers']), ip(st['ao_base']), ip(st['ao_stride']), nao, len(charges), dp(charges),
        dp(coordinates), dp(S), dp(T), dp(V)))
    return S, T, V


def compute_one_electron_integrals(molecule, basis):
    """(S, T, V) in one pass over the shell pairs; V includes the electron charge."""
    return _one_electron(molecule, basis)


def compute_overlap_integrals(molecule, basis, ket_basis=None):
    """Overlap matrix; with ket_basis the nao(basis) x nao(ket_basis) bra/ket block."""
    S = _one_electron(molecule, basis, ket_basis=ket_basis, need_v=False)[0]
    if ket_basis is None:
        return S
    n = basis.get_dimensions_of_basis()
    return np.ascontiguousarray(S[:n, n:])


def compute_kinetic_energy_integrals(molecule, basis):
    return _one_electron(molecule, basis, need_v=False)[1]


def compute_nuclear_potential_integrals(molecule, basis, charges=None, coordinates=None):
    """Nuclear attraction of the electrons, -sum_C Z_C <a|1/|r - C||b>.

    ``charges``/``coordinates`` (bohr) replace the nuclei when given, as in
    the reference.
    """
    return _one_electron(molecule, basis, charges, coordinates)[2]


def _as_matrix(basis, values):
    m = make_matrix(basis, mat_t.symmetric)
    m.set_values(values)
    return m


class OverlapDriver:

    def compute(self, molecule, basis, ket_basis=None):
        """Symmetric Matrix; with ket_basis the rectangular bra/ket SubMatrix."""
        if ket_basis is None:
            return _as_matrix(basis, compute_overlap_integrals(molecule, basis))
        S = compute_overlap_integrals(molecule, basis, ket_basis)
        return SubMatrix([0, 0, S.shape[0], S.shape[1]], S)


class KineticEnergyDriver:

    def compute(self, molecule, basis):
        return _as_matrix(basis, compute_kinetic_energy_integrals(molecule, basis))


class NuclearPotentialDriver:

    def compute(self, molecule, basis, charges=None, coordinates=None):
        """Positive potential matrix (electron charge not included), as the reference driver."""
        return _as_matrix(basis, -compute_nuclear_potential_integrals(molecule, basis, charges, coordinates))

