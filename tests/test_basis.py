"""Basis normalisation, AO order and the block-matrix surface."""
import math

import numpy as np

from veloxchem_amd import Matrix, MolecularBasis, Molecule, make_matrix, mat_t
from veloxchem_amd.molecularbasis import normalize_contraction
from veloxchem_amd.workloads import water, water_cluster
from tests.common import dimer


def test_dimensions():
    # reference tests/test_fock_driver.py: basdims [0,16,58,78] (SVPD) and [0,22,52,72,86] (TZVP)
    _, svpd = dimer('def2-svpd')
    assert svpd.angular_block_dims() == [0, 16, 58, 78]
    _, tzvp = dimer('def2-tzvp')
    assert tzvp.angular_block_dims() == [0, 22, 52, 72, 86]
    assert MolecularBasis.read(water(), 'def2-svp').get_dimensions_of_basis() == 24


def test_s_shell_self_overlap_is_one():
    exps = [13.010701, 1.9622572, 0.44453796]
    coefs = [0.019682158, 0.13796524, 0.47831935]
    n = normalize_contraction(0, exps, coefs)
    s = sum(n[i] * n[j] * (math.pi / (exps[i] + exps[j])) ** 1.5 for i in range(3) for j in range(3))
    assert abs(s - 1.0) < 1e-14


def test_p_shell_self_overlap_is_one():
    # <p_x|p_x> of a contraction = sum c_i c_j (pi/p)^{3/2} / (2p)
    exps = [3.6316917145, 0.71695694366, 0.19541932860]
    n = normalize_contraction(1, exps, [-0.029033998305, -0.16877854032, -0.51403419628])
    s = sum(n[i] * n[j] * (math.pi / (exps[i] + exps[j])) ** 1.5 / (2 * (exps[i] + exps[j]))
            for i in range(3) for j in range(3))
    assert abs(s - 1.0) < 1e-14


def test_ao_order_l_major_component_then_function():
    mol = water()
    bas = MolecularBasis.read(mol, 'def2-svp')
    sh = bas.shell_arrays(mol.get_coordinates_in_bohr())
    # 7 s functions, then 4 p functions x 3 components, then 1 d x 5
    assert list(sh['ao_stride']) == [7, 7, 7, 4, 4, 1, 7, 7, 4, 7, 7, 4]
    aos = sorted(int(b) + m * int(s) for b, s, l in zip(sh['ao_base'], sh['ao_stride'], sh['l'])
                 for m in range(2 * l + 1))
    assert aos == list(range(24))


def test_matrix_symmetric_set_values_mirrors_upper_blocks():
    _, bas = dimer('def2-svpd')
    m = make_matrix(bas, mat_t.symmetric)
    rng = np.random.default_rng(1)
    a = rng.normal(size=(78, 78))
    m.set_values(a)
    full = m.to_numpy()
    assert np.allclose(full[:16, 16:], a[:16, 16:])
    assert np.allclose(full[16:, :16], a[:16, 16:].T)
    assert m.submatrix((1, 2)).get_dimensions() == [16, 58, 42, 20]


def test_cluster_generator_is_seeded_and_respects_min_distance():
    a = water_cluster(32, 10.0, 32)
    b = water_cluster(32, 10.0, 32)
    assert np.array_equal(a.get_coordinates_in_bohr(), b.get_coordinates_in_bohr())
    o = a.get_coordinates_in_angstrom()[::3]
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(32) * 100
    assert d.min() >= 2.6
    h = a.get_coordinates_in_angstrom()
    assert abs(np.linalg.norm(h[1] - h[0]) - 0.9572) < 1e-12


from veloxchem_amd import T4CScreener as _T4CScreener  # noqa: E402


class _FakeScreener(_T4CScreener):
    """The screener's host bookkeeping with pair maxima from a closed form (no library)."""

    def _partition(self, basis, molecule, label='eri'):
        if label != 'eri':
            raise ValueError(label)
        self._inputs = (basis, molecule, label)
        sh = basis.shell_arrays(molecule.get_coordinates_in_bohr())
        self._shell_l, self._shell_atom = np.asarray(sh['l']), np.asarray(sh['atom'])
        self._nshells = len(sh['l'])
        xyz = np.asarray(sh['centers'])
        self._pm = np.exp(-np.sum((xyz[:, None] - xyz[None]) ** 2, axis=-1))
        self._key = ('fake', xyz.tobytes())

    def pair_maxima(self):
        return self._pm


def test_screener_surface_copy_pickle_pair_blocks():
    """T4CScreener as exported by the reference (src/python/ExportT4CIntegrals.cpp:127-140): copy,
    pickle, gto_pair_blocks and the atom-restricted partitions, on the host side of the screener
    (no build)."""
    import copy
    import pickle

    import numpy as np
    import pytest

    from tests.common import dimer
    mol, bas = dimer('def2-svp')
    s = _FakeScreener().partition(bas, mol, 'eri')
    blocks = s.gto_pair_blocks()
    npairs = sum(len(b['pairs']) for b in blocks)
    assert npairs == s._nshells * (s._nshells + 1) // 2  # exp(-r^2) >= 1e-30 for every dimer pair
    for b in blocks:
        la, lb = b['angular_momentums']
        assert la <= lb
        for (p, q), v in zip(b['pairs'], b['max_values']):
            assert (s._shell_l[p], s._shell_l[q]) == (la, lb)
            assert b['bin'] == (0 if v >= 1.0 else int(-np.floor(np.log10(np.sqrt(v)))))
    a0 = _FakeScreener().partition_atom(bas, mol, 'eri', 0)
    assert a0.is_restricted and not s.is_restricted
    touched = [(a0._shell_atom[p], a0._shell_atom[q]) for b in a0.gto_pair_blocks() for p, q in b['pairs']]
    assert touched and all(0 in t for t in touched)
    ap = _FakeScreener().partition_atom_pair(bas, mol, 'eri', 0, 3)
    pairs = [tuple(sorted((ap._shell_atom[p], ap._shell_atom[q]))) for b in ap.gto_pair_blocks() for p, q in b['pairs']]
    assert pairs and set(pairs) == {(0, 3)}
    assert s != a0
    with pytest.raises(ValueError):
        _FakeScreener().partition(bas, mol, 'ovl')
    # pickle and copy keep the inputs and the restriction (and re-partition)
    for c in (copy.copy(a0), copy.deepcopy(a0), pickle.loads(pickle.dumps(a0))):
        assert isinstance(c, _FakeScreener) and c == a0 and c._atoms == (0,)
