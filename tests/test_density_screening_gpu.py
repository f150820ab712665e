"""Density-weighted screening (the fast mode, no reference counterpart): quartets whose Schwarz bound
times the largest |D| of the six D blocks they read (taken per pair of atom groups) is below tau are
skipped.  Pinned by: a block-diagonal density (SAD: zero between atoms of different groups) loses
nothing, since every skipped quartet there contributes exactly zero; the error grows with tau and stays
within a bound of the dropped quartets' size; tau = 0 restores the reference build; the SCF with it
reaches the reference energy."""
import ctypes

import numpy as np
import pytest

from veloxchem_amd import FockDriver, MolecularBasis, T4CScreener, _native, make_matrix, mat_t
from veloxchem_amd.sadguess import sad_density
from veloxchem_amd.workloads import pseudo_density, water_cluster

pytestmark = pytest.mark.gpu


def _cluster(nmol, box, seed, basis='def2-svp'):
    mol = water_cluster(nmol, box, seed)
    bas = MolecularBasis.read(mol, basis)
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    return mol, bas, scr


def _build(scr, D, tau, label='2jk'):
    _native.set_density_screening(tau)
    try:
        return np.asarray(FockDriver().compute(scr, D, label, 1.0, 0.0, 12))
    finally:
        _native.set_density_screening(0.0)


def test_block_diagonal_density_loses_nothing():
    """SAD density of a 16-water cluster (groups of 12 atoms): every quartet whose six D blocks are
    all zero is skipped, and they contribute nothing, so G equals the unscreened build."""
    mol, bas, scr = _cluster(16, 8.5, 5)
    D = sad_density(mol, MolecularBasis.read(mol, 'ao-start-guess'), bas)
    g0 = _build(scr, D, 0.0)
    g1 = _build(scr, D, 1e-300)
    print(f'SAD density: max|dG| {np.max(np.abs(g1 - g0)):.2e}', flush=True)
    assert np.max(np.abs(g1 - g0)) < 1e-12


@pytest.mark.parametrize('basis', ['def2-svp', 'def2-tzvp'])
def test_error_grows_with_threshold_and_stays_bounded(basis):
    """A delocalised density: the screened G stays within a bound of the dropped quartets' size summed
    over the elements they touch (loose: 1e4 tau), and tau = 1e-6 does drop quartets that matter."""
    mol, bas, scr = _cluster(5, 7.0, 23, basis)
    n = bas.get_dimensions_of_basis()
    D = pseudo_density(n, 25, seed=3)
    g0 = _build(scr, D, 0.0)
    prev = 0.0
    for tau in (1e-12, 1e-10, 1e-8, 1e-6):
        g = _build(scr, D, tau)
        err = float(np.max(np.abs(g - g0)))
        de = abs(float(np.sum(D * (g - g0))))
        print(f'{basis} tau {tau:g}: max|dG| {err:.2e}, |dE2| {de:.2e}', flush=True)
        assert np.max(np.abs(g - g.T)) < 1e-12
        assert err <= 1e4 * tau
        prev = err
    assert prev > 1e-12  # tau = 1e-6 skips quartets that matter
    # off again: the reference build
    assert np.max(np.abs(_build(scr, D, 0.0) - g0)) < 1e-12


def test_refusals():
    with pytest.raises(ValueError):
        _native.set_density_screening(-1.0)
    mol, bas, scr = _cluster(2, 4.0, 29)
    n = bas.get_dimensions_of_basis()
    lib, ctx = _native.lib(), _native.context(0)
    d_dm, d_g = _native.DeviceArray((n, n)), _native.DeviceArray((n, n))
    try:
        d_dm.upload(pseudo_density(n, 10, seed=1))
        _native.check(lib.vlx_fock_compute_device(ctx, scr.handle, d_dm.ptr, b'2jk', 1.0, 0.0, 12, d_g.ptr))
        _native.set_density_screening(1e-10)
        with pytest.raises(ValueError, match='no-allocation'):
            _native.check(lib.vlx_fock_compute_device_noalloc(ctx, scr.handle, d_dm.ptr, b'2jk', 1.0, 0.0, 12,
                                                              d_g.ptr))
    finally:
        _native.set_density_screening(0.0)
        d_dm.free()
        d_g.free()


@pytest.mark.parametrize('incremental', [False, True])
def test_scf_dimer_with_density_screening(incremental):
    """The water dimer def2-SVP SCF (tests/test_scf_rest.py:38-40, -151.9220860932 Eh) with screened
    builds, confirmed by one unscreened build: the reference energy to 1e-8 Eh."""
    from tests.test_scf import dimer
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    mol = dimer()
    bas = MolecularBasis.read(mol, 'def2-svp')
    drv = ScfRestrictedDriver()
    drv.density_screening = 1e-10
    drv.incremental_fock = incremental
    res = drv.compute(mol, bas)
    print(f'incremental {incremental}: E {res["scf_energy"]:.10f}, {res["num_iter"]} iterations', flush=True)
    assert res['is_converged']
    assert abs(res['scf_energy'] - (-151.9220860932)) < 1e-8


def test_fast_scf_cluster_converges_to_fp64_energy():
    """A 12-water def2-SVP SCF in the fast mode (mixed precision, incremental difference-density
    builds, density-weighted screening that follows the gradient down) converges to the FP64 SCF's
    energy within 1e-8 Eh, in not many more iterations."""
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    mol = water_cluster(12, 10.0 * (12 / 32.0) ** (1 / 3), 91)
    bas = MolecularBasis.read(mol, 'def2-svp')
    ref = ScfRestrictedDriver()
    ref.conv_thresh = 1e-8
    r0 = ref.compute(mol, bas)
    fast = ScfRestrictedDriver()
    fast.conv_thresh = 1e-8
    fast.precision, fast.mixed_threshold = 'mixed', 1e-6
    fast.incremental_fock = True
    fast.density_screening = 1e-10
    r1 = fast.compute(mol, bas)
    print(f'FP64 {r0["num_iter"]} iterations, fast {r1["num_iter"]}; dE {r1["scf_energy"] - r0["scf_energy"]:.2e}',
          flush=True)
    assert r0['is_converged'] and r1['is_converged']
    assert abs(r1['scf_energy'] - r0['scf_energy']) < 1e-8
    assert r1['num_iter'] <= r0['num_iter'] + 6
