"""GPU parity of the multi-density pass: ``FockDriver.compute(screener, Matrices, labels, ...)`` ->
``vlx_fock_compute_multi``, one pass over the integrals digesting every density with its own label,
as CFockDriver::compute with CMatrices does (src/t4c_func/FockDriver.cpp:173-229,
src/t4c_func/T4CMatricesDistributor.cpp).

Every density of a call is checked on its own: against the oracle (4-water def2-SVP cluster, every
kernel variant forced in turn, symmetric and general densities, 1e-10), against the reference's
fixtures (def2-SVPD general density, def2-TZVP with f shells, erf integrals), and against single
builds for a call with more densities than the LDS tiles of one launch can hold.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.common import dimer, load, reference_equal
from veloxchem_amd import FockDriver, Matrices, MolecularBasis, T4CScreener, make_matrix, mat_t
from veloxchem_amd.workloads import pseudo_density, water_cluster

pytestmark = pytest.mark.gpu

TOL = 1e-10
XF = 0.37


def _screened(mol, bas):
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    return scr


def _matrices(bas, dens, general):
    mats = Matrices()
    for i, (d, g) in enumerate(zip(dens, general)):
        m = make_matrix(bas, mat_t.general if g else mat_t.symmetric)
        m.set_values(d)
        mats.add(m, str(i))
    return mats


def _cluster_case():
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    n = bas.get_dimensions_of_basis()
    rng = np.random.default_rng(21)
    sym = [pseudo_density(n, 20, seed=s) for s in (11, 12, 13)]
    gen = pseudo_density(n, 20, seed=14) + 0.05 * rng.standard_normal((n, n))
    # (label, density, general): every label, two general densities (one with exchange, one without)
    cases = [('2jk', sym[0], False), ('j', sym[1], False), ('kx', sym[2], False), ('2jkx', gen, True),
             ('k', sym[0], False), ('j', gen, True)]
    return mol, bas, cases


def test_multi_matches_oracle_every_variant():
    """Six densities in one call (2jk, j, kx, k on symmetric densities; 2jkx and j on general ones)
    against the oracle's single-density builds at 1e-10, under every kernel variant forced for all
    classes and under the default choice."""
    from veloxchem_amd import _native
    mol, bas, cases = _cluster_case()
    scr = _screened(mol, bas)
    ob = orc.basis_from(bas, mol)
    refs = []
    for lab, d, g in cases:
        refs.append((orc.fock_general if g else orc.fock)(ob, d, lab, XF, 12)[0])
    mats = _matrices(bas, [c[1] for c in cases], [c[2] for c in cases])
    labels = [c[0] for c in cases]
    drv = FockDriver()
    ctx = _native.context()
    try:
        for v, name in [(-1, 'default')] + list(enumerate(FockDriver.VARIANTS)):
            _native.check(_native.lib().vlx_context_set_variant(ctx, v))
            out = drv.compute(scr, mats, labels, XF, 0.0, 12)
            for i, ref in enumerate(refs):
                g = out.matrix(str(i)).to_numpy()
                err = np.max(np.abs(g - ref))
                print(f'variant {name}, density {i} ({labels[i]}): max|dG| {err:.2e}', flush=True)
                assert np.isfinite(g).all(), (name, i)
                assert err < TOL, (name, i)
                assert out.matrix(str(i)).get_type() == (mat_t.general if cases[i][2] else mat_t.symmetric)
    finally:
        _native.check(_native.lib().vlx_context_set_variant(ctx, -1))


def test_multi_general_svpd_fixtures():
    """The reference's general-density fixtures (tests/test_fock_driver.py:320-360: J = j.gen,
    K = k.gen^T) beside the symmetric ones (j, k), four densities in one call."""
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    dg, ds = load('h2o.dimer.svpd.density.gen.npy'), load('h2o.dimer.svpd.density.npy')
    mats = _matrices(bas, [dg, dg, ds, ds], [True, True, False, False])
    out = FockDriver().compute(scr, mats, ['j', 'k', 'j', 'k'], 0.0, 0.0, 15)
    assert reference_equal(out.matrix('0').to_numpy(), load('h2o.dimer.svpd.j.gen.npy'))
    assert reference_equal(out.matrix('1').to_numpy(), load('h2o.dimer.svpd.k.gen.npy').T)
    assert reference_equal(out.matrix('2').to_numpy(), load('h2o.dimer.svpd.j.npy'))
    assert reference_equal(out.matrix('3').to_numpy(), load('h2o.dimer.svpd.k.npy'))


def test_multi_tzvp_fixtures_and_range_separation_rule():
    """f shells (def2-TZVP dimer fixtures) and the rule of one set of integrals per call: all labels
    range-separated -> erf(omega r)/r (the erf fixture), one plain label -> 1/r for every density,
    k_rs then digesting like k (CT4CMatricesDistributor::need_omega)."""
    from veloxchem_amd import _native
    if _native.lib().vlx_max_angular_momentum() < 3:
        pytest.skip('f-shell kernels not built')
    mol, bas = dimer('def2-tzvp')
    scr = _screened(mol, bas)
    d = load('h2o.dimer.tzvp.density.npy')
    j, k, kerf = load('h2o.dimer.tzvp.j.npy'), load('h2o.dimer.tzvp.k.npy'), load('h2o.dimer.tzvp.erf.k.npy')
    drv = FockDriver()
    out = drv.compute(scr, _matrices(bas, [d] * 4, [False] * 4), ['2jk', 'J', 'k', '2jkx'], 0.5, 0.0, 15)
    assert reference_equal(out.matrix('0').to_numpy(), 2 * j - k)
    assert reference_equal(out.matrix('1').to_numpy(), j)  # labels are case-insensitive
    assert reference_equal(out.matrix('2').to_numpy(), k)
    assert reference_equal(out.matrix('3').to_numpy(), 2 * j - 0.5 * k)
    rs = drv.compute(scr, _matrices(bas, [d] * 2, [False] * 2), ['k_rs', 'kx_rs'], 0.21, 0.63, 15)
    assert reference_equal(rs.matrix('0').to_numpy(), kerf)
    assert reference_equal(rs.matrix('1').to_numpy(), 0.21 * kerf)
    mixed = drv.compute(scr, _matrices(bas, [d] * 2, [False] * 2), ['k_rs', 'j'], 0.0, 0.63, 15)
    assert reference_equal(mixed.matrix('0').to_numpy(), k)
    assert reference_equal(mixed.matrix('1').to_numpy(), j)
    zero = drv.compute(scr, _matrices(bas, [d] * 2, [False] * 2), ['k_rs', 'j_rs'], 0.0, 0.0, 15)
    assert not np.any(zero.matrix('0').to_numpy()) and not np.any(zero.matrix('1').to_numpy())


def test_multi_many_densities_match_single_builds():
    """Twenty densities (more than the LDS tiles of one tiled launch hold, so the tiled classes run
    in groups of densities) against one single-density build each, and a one-density call."""
    mol = water_cluster(6, 7.0, 3)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    n = bas.get_dimensions_of_basis()
    dens = [pseudo_density(n, 30, seed=100 + i) for i in range(20)]
    labels = ['2jk', 'j', 'k', 'kx', '2jkx'] * 4
    drv = FockDriver()
    out = drv.compute(scr, _matrices(bas, dens, [False] * 20), labels, 0.6, 0.0, 12)
    for i, (d, lab) in enumerate(zip(dens, labels)):
        single = np.asarray(drv.compute(scr, d, lab, 0.6, 0.0, 12))
        err = np.max(np.abs(out.matrix(str(i)).to_numpy() - single))
        assert err < 1e-12 * max(1.0, np.max(np.abs(single))), (i, lab, err)
    one = drv.compute(scr, _matrices(bas, dens[:1], [False]), ['2jk'], 0.0, 0.0, 12)
    single = np.asarray(drv.compute(scr, dens[0], '2jk', 0.0, 0.0, 12))
    assert np.max(np.abs(one.matrix('0').to_numpy() - single)) < 1e-12 * max(1.0, np.max(np.abs(single)))


def test_multi_argument_errors():
    from veloxchem_amd import _native
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    d = load('h2o.dimer.svpd.density.npy')
    drv = FockDriver()
    with pytest.raises(ValueError):
        drv.compute(scr, _matrices(bas, [d, d], [False, False]), ['2jk'], 0.0, 0.0, 15)
    with pytest.raises(ValueError):
        drv.compute(scr, _matrices(bas, [d, d], [False, False]), ['2jk', 'q'], 0.0, 0.0, 15)
    asym = d.copy()
    asym[0, 1] += 1e-3
    with pytest.raises(ValueError):  # a symmetric Matrix that is not symmetric
        drv.compute(scr, Matrices([asym, d]), ['2jk', 'j'], 0.0, 0.0, 15)
