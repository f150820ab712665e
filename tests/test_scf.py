"""Closed-shell SCF around the Fock build.

Reference energies: tests/test_input_energy.py:124-138 (water, def2-SVP, au
geometry, RHF -75.9612569185, 8 iterations, |dE| < 1e-6) and
tests/test_scf_rest.py:12-39 (water dimer, def2-SVP, RHF -151.9220860932,
|dE| < 1e-8).  The CPU tests drive the SCF with the oracle's Fock build, which
pins the SCF logic (guess, DIIS, thresholds, energy); the GPU tests run the
product path end to end.
"""
import numpy as np
import pytest

from veloxchem_amd import MolecularBasis, Molecule
from veloxchem_amd.oneeints import compute_overlap_integrals
from veloxchem_amd.sadguess import sad_density
from veloxchem_amd.scfdriver import ScfRestrictedDriver, c2diis_weights

WATER_AU = """
    O   0.0   0.0   0.0
    H   0.0   1.4   1.1
    H   0.0  -1.4   1.1
"""

DIMER_XYZ = """6
xyz
O   -0.1858140  -1.1749469   0.7662596
H   -0.1285513  -0.8984365   1.6808606
H   -0.0582782  -0.3702550   0.2638279
O    0.1747051  11.1050002  -0.7244430
H   -0.5650842  11.3134964  -1.2949455
H    0.9282185  11.0652990  -1.3134026
"""


def water():
    return Molecule.read_str(WATER_AU, 'au')


def dimer():
    return Molecule.read_xyz_string(DIMER_XYZ)


def _oracle_driver():
    """SCF driven by the oracle's Fock build on the CPU (test double for the GPU build)."""
    import torch
    from oracle import oracle as orc

    def fock_2e(screener, D, thresh_int):
        g, _ = orc.fock(screener, D.numpy(), '2jk', 1.0, thresh_int, 0)
        return torch.from_numpy(g)

    drv = ScfRestrictedDriver()
    drv.fock_2e = fock_2e
    drv.screener_factory = lambda basis, mol: orc.basis_from(basis, mol)
    return drv


def test_sad_density_electron_count():
    for mol in (water(), dimer()):
        bas = MolecularBasis.read(mol, 'def2-svp')
        D = sad_density(mol, MolecularBasis.read(mol, 'ao-start-guess'), bas)
        S = compute_overlap_integrals(mol, bas)
        assert np.allclose(D, D.T)
        assert abs(np.sum(D * S) - mol.number_of_electrons() / 2) < 1e-10


def test_valence_basis():
    mol = water()
    bas = MolecularBasis.read(mol, 'def2-svp')
    val = bas.reduce_to_valence_basis()
    # O keeps s,s,s,p,p (drops d); H keeps s,s (drops p): 3 + 6 + 2 * 2 = 13 AOs
    assert val.get_dimensions_of_basis() == 13
    assert val.max_angular_momentum() == 1


def test_c2diis_weights():
    rng = np.random.default_rng(5)
    errs = [rng.standard_normal((6, 6)) for _ in range(4)]
    w = c2diis_weights(errs)
    assert abs(np.sum(w) - 1.0) < 1e-12
    # an exact affine zero (e, -e) is found: weights (1/2, 1/2, 0)
    a, b = errs[0], errs[1]
    w = c2diis_weights([a, -a, b])
    assert np.allclose(w, [0.5, 0.5, 0.0], atol=1e-10)


def test_dyn_threshold():
    drv = ScfRestrictedDriver()
    assert drv._get_dyn_threshold(1e-7) == 1e-12
    assert drv._get_dyn_threshold(5.0) == 1e-10
    assert drv._get_dyn_threshold(0.5) == pytest.approx(1e-11)
    assert drv._get_dyn_threshold(1e-4) == 1e-12


def test_scf_water_oracle():
    mol = water()
    drv = _oracle_driver()
    res = drv.compute(mol, MolecularBasis.read(mol, 'def2-svp'))
    assert drv.is_converged
    assert abs(res['scf_energy'] - (-75.9612569185)) < 1.0e-6
    assert drv.num_iter == 8


def test_scf_dimer_oracle():
    mol = dimer()
    drv = _oracle_driver()
    res = drv.compute(mol, MolecularBasis.read(mol, 'def2-svp'), MolecularBasis.read(mol, 'ao-start-guess'))
    assert abs(res['scf_energy'] - (-151.9220860932)) < 1.0e-8


def test_scf_incremental_fock_oracle():
    # difference-density builds reach the full-build energy; the last energy is from a full build
    mol = dimer()
    bas, min_bas = MolecularBasis.read(mol, 'def2-svp'), MolecularBasis.read(mol, 'ao-start-guess')
    ref = _oracle_driver()
    e_full = ref.compute(mol, bas, min_bas)['scf_energy']
    drv = _oracle_driver()
    drv.incremental_fock = True
    res = drv.compute(mol, bas, min_bas)
    assert drv.is_converged
    inc = [h for h in drv.history if h['incremental']]
    assert len(inc) >= 3
    assert all(h['build_threshold'] < h['eri_threshold'] for h in inc)
    assert not drv.history[-1]['incremental']
    assert abs(res['scf_energy'] - e_full) < 1.0e-9
    assert abs(res['scf_energy'] - (-151.9220860932)) < 1.0e-8


def test_scf_rejects_open_shell():
    mol = water()
    mol.set_charge(1)
    with pytest.raises(ValueError):
        ScfRestrictedDriver().compute(mol, MolecularBasis.read(mol, 'def2-svp'))


@pytest.mark.gpu
def test_scf_water_gpu():
    mol = water()
    drv = ScfRestrictedDriver()
    res = drv.compute(mol, MolecularBasis.read(mol, 'def2-svp'))
    assert drv.is_converged
    assert abs(res['scf_energy'] - (-75.9612569185)) < 1.0e-6
    assert drv.num_iter == 8


@pytest.mark.gpu
def test_scf_dimer_gpu():
    mol = dimer()
    drv = ScfRestrictedDriver()
    res = drv.compute(mol, MolecularBasis.read(mol, 'def2-svp'))
    assert abs(res['scf_energy'] - (-151.9220860932)) < 1.0e-8


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_scf_dimer_gpu_incremental(precision):
    mol = dimer()
    drv = ScfRestrictedDriver()
    drv.incremental_fock = True
    drv.precision = precision
    drv.mixed_threshold = 1.0e-8
    res = drv.compute(mol, MolecularBasis.read(mol, 'def2-svp'))
    assert drv.is_converged
    assert sum(h['incremental'] for h in drv.history) >= 3
    assert abs(res['scf_energy'] - (-151.9220860932)) < 1.0e-8


def test_canonical_purification_matches_eigh_projector():
    import torch
    from veloxchem_amd.scfdriver import canonical_purification
    rng = np.random.default_rng(3)
    m, nocc = 40, 11
    q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    evals = np.concatenate([np.sort(rng.uniform(-20.0, -0.5, nocc)), np.sort(rng.uniform(0.2, 15.0, m - nocc))])
    fo = (q * evals) @ q.T
    d = canonical_purification(torch.from_numpy(fo), nocc).numpy()
    ref = q[:, :nocc] @ q[:, :nocc].T
    assert np.max(np.abs(d - ref)) < 1e-10
    assert abs(np.trace(d) - nocc) < 1e-10
    # degenerate HOMO/LUMO: either no projector (the caller diagonalises) or, once rounding has
    # split the pair, a valid idempotent projector with the right trace
    evals[nocc] = evals[nocc - 1]
    d = canonical_purification(torch.from_numpy((q * evals) @ q.T), nocc)
    if d is not None:
        d = d.numpy()
        assert np.max(np.abs(d @ d - d)) < 1e-8 and abs(np.trace(d) - nocc) < 1e-8
