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
