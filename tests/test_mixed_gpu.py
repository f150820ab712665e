"""Mixed precision: FP32 recursions for quartets with small Schwarz bound x max|D|.

Every mixed result is checked against the FP64 build on the same inputs (and
that against the oracle); the FP32 part must stay within a bound that scales
with the threshold, and FP64 mode must be unchanged after a mixed build.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from veloxchem_amd import FockDriver, MolecularBasis, T4CScreener, _native, make_matrix, mat_t
from veloxchem_amd.workloads import pseudo_density, water_cluster

pytestmark = pytest.mark.gpu


def _setup():
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    n = bas.get_dimensions_of_basis()
    D = pseudo_density(n, 20, seed=11)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(D)
    return mol, bas, scr, D, dm


def _build(scr, dm, mode, tau=0.0):
    _native.set_precision(mode, tau)
    try:
        return FockDriver().compute(scr, dm, '2jk', 1.0, 0.0, 12).to_numpy()
    finally:
        _native.set_precision('fp64')


def test_mixed_thresholds_bound_the_error():
    mol, bas, scr, D, dm = _setup()
    g64 = _build(scr, dm, 'fp64')
    ref, _ = orc.fock(orc.basis_from(bas, mol), D, '2jk', 1.0, 12, 0)
    assert np.max(np.abs(g64 - ref)) < 1e-10
    dmax = np.max(np.abs(D))
    prev_frac = -1.0
    for tau in (1e-10, 1e-8, 1e-6, 1e-4):
        g = _build(scr, dm, 'mixed', tau)
        k32 = _native.last_fp32_bin_sum()
        assert k32 == int(np.ceil(2 + np.log10(dmax / tau)))
        p64, p32 = scr.precision_split(12, k32)
        frac = p32 / (p64 + p32)
        assert frac >= prev_frac
        prev_frac = frac
        err = np.max(np.abs(g - g64))
        print(f'tau {tau:g}: k32 {k32}, FP32 share {frac:.3f}, max|dG| {err:.3e}', flush=True)
        # FP32 quartets are bounded by tau in size; their FP32 error is a small fraction of that
        assert err < max(1e-12, 1e-3 * tau)
    assert prev_frac > 0.0


def test_all_fp32_tiled_classes_still_close():
    mol, bas, scr, D, dm = _setup()
    g64 = _build(scr, dm, 'fp64')
    g = _build(scr, dm, 'mixed', 1e6)  # k32 = 0: every tiled class in FP32
    assert _native.last_fp32_bin_sum() == 0
    err = np.max(np.abs(g - g64))
    de = abs(np.sum(D * (g - g64)))
    print(f'all-FP32 tiled classes: max|dG| {err:.3e}, |dE2| {de:.3e}', flush=True)
    assert err < 1e-4 and de < 1e-4
    # FP64 mode is restored and exact again
    assert np.max(np.abs(_build(scr, dm, 'fp64') - g64)) < 1e-12


def test_precision_argument_errors():
    with pytest.raises(ValueError):
        _native.set_precision('mixed', 0.0)
    with pytest.raises(ValueError):
        _native.set_precision('fp16', 1e-8)


def _tzvp_cluster(nmol, box, seed):
    from veloxchem_amd.sadguess import sad_density
    mol = water_cluster(nmol, box, seed)
    bas = MolecularBasis.read(mol, 'def2-tzvp')
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    D = sad_density(mol, MolecularBasis.read(mol, 'ao-start-guess'), bas)
    return mol, bas, scr, D


def _f_classes_fp32(scr, ithr, k32):
    """FP32 primitive slots of the classes with an f shell (l = 3)."""
    return {c: n for c, n in scr.precision_classes(ithr, k32).items() if 3 in c and n > 0}


@pytest.mark.parametrize('tau', [1e-6, 1e-8, 1e-10])
def test_mixed_tzvp_cluster_vs_fp64_and_reference(tau):
    """3-water def2-TZVP cluster (f shells on O), SAD density: the FP64 build equals the reference's
    CFockDriver::compute at 1e-10 (where oracle/_ref is built; else the oracle), and the mixed build
    at threshold tau stays within 1e-3 tau of FP64 elementwise and 1e-6 in the two-electron energy,
    with FP32 work in f-containing classes (the tests/test_fock_driver.py def2-TZVP cases use the
    same basis)."""
    from oracle import refharness
    mol, bas, scr, D = _tzvp_cluster(3, 5.5, 3)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(D)
    g64 = _build(scr, dm, 'fp64')
    if refharness.available():
        ref, _ = refharness.fock(mol, 'def2-tzvp', D, '2jk', 1.0, 0.0, 12)
    else:
        ref, _ = orc.fock(orc.basis_from(bas, mol), D, '2jk', 1.0, 12, 0)
    assert np.max(np.abs(g64 - ref)) < 1e-10
    g = _build(scr, dm, 'mixed', tau)
    k32 = _native.last_fp32_bin_sum()
    assert k32 == int(np.ceil(2 + np.log10(np.max(np.abs(D)) / tau)))
    p64, p32 = scr.precision_split(12, k32)
    f32 = _f_classes_fp32(scr, 12, k32)
    err = np.max(np.abs(g - g64))
    de = abs(float(np.sum(D * (g - g64))))
    print(f'tau {tau:g}: k32 {k32}, FP32 share {p32 / (p64 + p32):.3f}, f classes in FP32 {sorted(f32)}, '
          f'max|dG| {err:.3e}, |dE2| {de:.3e}', flush=True)
    assert p32 > 0 and f32, 'no FP32 work in an f-containing class'
    assert err <= 1e-3 * tau
    assert de <= 1e-6
    # every f class with an FP32 kernel runs FP32 when all tiled work is FP32 (k32 = 0)
    g0 = _build(scr, dm, 'mixed', 1e6)
    assert _native.last_fp32_bin_sum() == 0
    f_all = _f_classes_fp32(scr, 12, 0)
    assert {(3, 0, 0, 0), (3, 1, 0, 0), (3, 0, 1, 0)} <= set(f_all), sorted(f_all)
    assert np.max(np.abs(g0 - g64)) < 1e-4


def test_mixed_tzvp_64_water_properties():
    """The 64-water def2-TZVP cluster (seed 64, 12.8 A box, 2752 AOs) at tau = 1e-6: the mixed G is
    symmetric, its two-electron energy is within 1e-6 of FP64, its elementwise distance from FP64 is
    within 1e-3 tau, and f-containing classes run FP32 work."""
    mol, bas, scr, D = _tzvp_cluster(64, 12.8, 64)
    assert D.shape == (2752, 2752)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(D)
    g64 = _build(scr, dm, 'fp64')
    g = _build(scr, dm, 'mixed', 1e-6)
    k32 = _native.last_fp32_bin_sum()
    f32 = _f_classes_fp32(scr, 12, k32)
    err = np.max(np.abs(g - g64))
    de = abs(float(np.sum(D * (g - g64))))
    print(f'64 waters: k32 {k32}, f classes in FP32 {sorted(f32)}, max|dG| {err:.3e}, |dE2| {de:.3e}', flush=True)
    assert np.max(np.abs(g - g.T)) < 1e-12
    assert np.max(np.abs(g64 - g64.T)) < 1e-12
    assert de <= 1e-6
    assert err <= 1e-3 * 1e-6
    assert f32
