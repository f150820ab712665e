"""GPU parity of the HIP Fock build against the CPU oracle and the reference fixtures.

FP64 bar: elementwise 1e-10 Hartree against the oracle on the same basis,
geometry and density; the reference's own rule (rtol = atol = 1e-12,
src/math/SubMatrix.cpp:105-115) against its fixtures.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests.common import dimer, load, reference_equal, sto3g_water
from veloxchem_amd import FockDriver, MolecularBasis, T4CScreener, make_matrix, mat_t
from veloxchem_amd.workloads import pseudo_density, water, water_cluster

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _screened(mol, bas):
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    return scr


def test_screener_maxima_and_bins_match_oracle():
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    mx_gpu = scr.pair_maxima()
    mx_ref, bins_ref = orc.pair_maxima(orc.basis_from(bas, mol))
    iu = np.triu_indices(mx_ref.shape[0])
    a, b = mx_gpu[iu], mx_ref[iu]
    # values agree to round-off of two different recursions (MD vs HGP) ...
    assert np.all(np.abs(a - b) <= 1e-10 * np.abs(b) + 1e-300)
    # ... and the Schwarz bins that select quartets are identical
    def bins(v):
        f = np.sqrt(v)
        with np.errstate(divide='ignore'):
            return np.where(f > 1e-15, np.where(f < 1.0, -np.floor(np.log10(np.maximum(f, 1e-300))), 0), -1)
    assert np.array_equal(bins(a), bins_ref[iu])
    assert int(np.sum(bins_ref[iu] >= 0)) == scr.number_of_pairs()


def test_eri_blocks_match_oracle_all_classes():
    mol = water()
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    ob = orc.basis_from(bas, mol)
    ns = bas.number_of_shells()
    rng = np.random.default_rng(7)
    quartets = [tuple(int(x) for x in rng.integers(0, ns, 4)) for _ in range(400)]
    blocks = FockDriver().eri_quartets(scr, quartets)
    worst = 0.0
    for q, blk in zip(quartets, blocks):
        ref = orc.eri_sph(ob, *q)
        worst = max(worst, float(np.max(np.abs(blk - ref))))
    assert worst < 1e-12, worst


def test_eri_blocks_tzvp_all_f_classes():
    from veloxchem_amd import _native
    if _native.lib().vlx_max_angular_momentum() < 3:
        pytest.skip('f-shell kernels not built')
    mol, bas = dimer('def2-tzvp')
    scr = _screened(mol, bas)
    ob = orc.basis_from(bas, mol)
    ls = bas.shell_arrays(mol.get_coordinates_in_bohr())['l']
    by_l = {l: [i for i, x in enumerate(ls) if x == l] for l in range(4)}
    rng = np.random.default_rng(11)
    quartets = []
    for la in range(4):
        for lb in range(4):
            for lc in range(4):
                for ld in range(4):
                    if 3 not in (la, lb, lc, ld):
                        continue
                    for _ in range(2):
                        quartets.append(tuple(int(rng.choice(by_l[x])) for x in (la, lb, lc, ld)))
    blocks = FockDriver().eri_quartets(scr, quartets)
    worst = 0.0
    for q, blk in zip(quartets, blocks):
        ref = orc.eri_sph(ob, *q)
        worst = max(worst, float(np.max(np.abs(blk - ref))))
    assert worst < 1e-12, worst


def test_compute_eri_sto3g_fixture():
    # reference tests/test_fock_eri.py:26-36
    mol, bas = sto3g_water()
    eri = FockDriver().compute_eri(mol, bas)
    assert np.max(np.abs(eri - load('h2o.sto3g.eri.npy'))) < 1e-12


@pytest.mark.parametrize('label,xf', [('j', 0.0), ('k', 0.0), ('2jk', 0.0), ('2jkx', 0.68), ('kx', 0.73)])
def test_sto3g_fixtures_all_labels(label, xf):
    mol, bas = sto3g_water()
    scr = _screened(mol, bas)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(load('h2o.sto3g.density.npy'))
    g = FockDriver()._compute_fock_omp(scr, dm, label, xf, 0.0, 15).to_numpy()
    j, k = load('h2o.sto3g.j.npy'), load('h2o.sto3g.k.npy')
    ref = {'j': j, 'k': k, '2jk': 2 * j - k, '2jkx': 2 * j - xf * k, 'kx': xf * k}[label]
    assert reference_equal(g, ref)


def test_sto3g_unscreened_path():
    mol, bas = sto3g_water()
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(load('h2o.sto3g.density.npy'))
    g = FockDriver()._compute_fock_omp(bas, mol, dm, '2jk', 0.0, 0.0).to_numpy()
    assert reference_equal(g, 2 * load('h2o.sto3g.j.npy') - load('h2o.sto3g.k.npy'))


@pytest.mark.parametrize('label', ['j', 'k', '2jk'])
def test_dimer_svpd_fixtures(label):
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(load('h2o.dimer.svpd.density.npy'))
    g = FockDriver().compute(scr, dm, label, 0.0, 0.0, 15).to_numpy()
    j, k = load('h2o.dimer.svpd.j.npy'), load('h2o.dimer.svpd.k.npy')
    ref = {'j': j, 'k': k, '2jk': 2 * j - k}[label]
    assert reference_equal(g, ref)


@pytest.mark.parametrize('label,xf', [('j', 0.0), ('k', 0.0), ('2jk', 0.0), ('2jkx', 0.31), ('kx', 0.27)])
def test_dimer_svpd_general_density_fixtures(label, xf):
    """mat_t.general density: reference tests/test_fock_driver.py:320-360 (J = j.gen, K = k.gen^T)."""
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    dm = make_matrix(bas, mat_t.general)
    dm.set_values(load('h2o.dimer.svpd.density.gen.npy'))
    g = FockDriver().compute(scr, dm, label, xf, 0.0, 15)
    assert g.get_type() == mat_t.general
    j, kt = load('h2o.dimer.svpd.j.gen.npy'), load('h2o.dimer.svpd.k.gen.npy').T
    ref = {'j': j, 'k': kt, '2jk': 2 * j - kt, '2jkx': 2 * j - xf * kt, 'kx': xf * kt}[label]
    assert reference_equal(g.to_numpy(), ref)


def test_general_density_matrices_overload():
    """Matrices + label list with general densities (reference tests/test_fock_driver.py:60-140)."""
    from veloxchem_amd import Matrices
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    dm = make_matrix(bas, mat_t.general)
    dm.set_values(load('h2o.dimer.svpd.density.gen.npy'))
    dms = Matrices()
    dms.add(dm, '0')
    dms.add(dm, '1')
    out = FockDriver().compute(scr, dms, ['j', 'k'], 0.0, 0.0, 15)
    assert reference_equal(out.matrix('0').to_numpy(), load('h2o.dimer.svpd.j.gen.npy'))
    assert reference_equal(out.matrix('1').to_numpy(), load('h2o.dimer.svpd.k.gen.npy').T)


def test_general_density_cluster_vs_oracle_1e10():
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    n = bas.get_dimensions_of_basis()
    rng = np.random.default_rng(5)
    d = pseudo_density(n, 20, seed=3) + 0.05 * rng.standard_normal((n, n))
    dm = make_matrix(bas, mat_t.general)
    dm.set_values(d)
    ob = orc.basis_from(bas, mol)
    for label, xf in (('2jk', 0.0), ('2jkx', 0.4)):
        g = FockDriver().compute(scr, dm, label, xf, 0.0, 12).to_numpy()
        ref, _ = orc.fock_general(ob, d, label, xf, 12)
        assert np.max(np.abs(g - ref)) < 1e-10
        assert np.max(np.abs(g - g.T)) > 1e-3  # the exchange of the antisymmetric part is there


def test_dimer_tzvp_fixture_2jk():
    from veloxchem_amd import _native
    if _native.lib().vlx_max_angular_momentum() < 3:
        pytest.skip('f-shell kernels not built')
    mol, bas = dimer('def2-tzvp')
    scr = _screened(mol, bas)
    dm = load('h2o.dimer.tzvp.density.npy')
    g = FockDriver().compute(scr, dm, '2jk', 0.0, 0.0, 15)
    assert reference_equal(g, 2 * load('h2o.dimer.tzvp.j.npy') - load('h2o.dimer.tzvp.k.npy'))


def test_water_svp_config_vs_oracle_1e10():
    mol = water()
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    dm = pseudo_density(24, 5, seed=1)
    g = FockDriver().compute(scr, dm, '2jk', 1.0, 0.0, 12)
    ref, _ = orc.fock(orc.basis_from(bas, mol), dm, '2jk', 1.0, 12)
    assert np.max(np.abs(g - ref)) < TOL


def test_four_water_cluster_vs_oracle_1e10():
    mol = water_cluster(4, 6.0, seed=4)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    dm = pseudo_density(96, 20, seed=2)
    g = FockDriver().compute(scr, dm, '2jk', 1.0, 0.0, 12)
    ref, _ = orc.fock(orc.basis_from(bas, mol), dm, '2jk', 1.0, 12)
    assert np.max(np.abs(g - ref)) < TOL


def test_cluster32_properties():
    # full-size workload: symmetry, linearity in D and label consistency
    mol = water_cluster(32, 10.0, seed=32)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    n = bas.get_dimensions_of_basis()
    assert n == 768
    d1 = pseudo_density(n, 160, seed=3)
    d2 = pseudo_density(n, 160, seed=4)
    drv = FockDriver()
    g1 = drv.compute(scr, d1, '2jk', 1.0, 0.0, 12)
    g2 = drv.compute(scr, d2, '2jk', 1.0, 0.0, 12)
    g12 = drv.compute(scr, d1 + d2, '2jk', 1.0, 0.0, 12)
    assert np.max(np.abs(g1 - g1.T)) < 1e-12
    assert np.max(np.abs(g12 - g1 - g2)) < 1e-10
    j = drv.compute(scr, d1, 'j', 0.0, 0.0, 12)
    k = drv.compute(scr, d1, 'k', 0.0, 0.0, 12)
    assert np.max(np.abs(2 * j - k - g1)) < 1e-10
    # Coulomb energy is positive for a positive semi-definite density
    assert float(np.sum(d1 * j)) > 0.0


def test_error_paths():
    mol, bas = sto3g_water()
    scr = _screened(mol, bas)
    dm = load('h2o.sto3g.density.npy')
    drv = FockDriver()
    with pytest.raises(ValueError):
        drv.compute(scr, dm, 'xyz', 0.0, 0.0, 12)
    # omega is ignored by the Coulomb labels, as in the reference
    g0 = drv.compute(scr, dm, '2jk', 1.0, 0.0, 12)
    g1 = drv.compute(scr, dm, '2jk', 1.0, 0.4, 12)
    assert np.max(np.abs(np.asarray(g0) - np.asarray(g1))) < 1e-13
    # erf(0 r)/r = 0
    assert np.max(np.abs(np.asarray(drv.compute(scr, dm, 'k_rs', 0.0, 0.0, 12)))) == 0.0
    with pytest.raises(ValueError):
        drv.compute(scr, dm[:5, :5], '2jk', 1.0, 0.0, 12)
    with pytest.raises(ValueError):
        drv.compute(scr, dm, '2jk', 1.0, 0.0, 31)
    bad = dm.copy()
    bad[0, 1] += 1.0
    with pytest.raises(ValueError):
        drv.compute(scr, bad, '2jk', 1.0, 0.0, 12)


@pytest.mark.parametrize('label,xf', [('k_rs', 0.0), ('kx_rs', 0.21)])
def test_dimer_tzvp_erf_fixture(label, xf):
    """Reference tests/test_fock_erf_driver.py: k_rs / kx_rs (0.21), omega 0.63, ithreshold 15."""
    mol, bas = dimer('def2-tzvp')
    scr = _screened(mol, bas)
    dm = make_matrix(bas, mat_t.symmetric)
    dm.set_values(load('h2o.dimer.tzvp.density.npy'))
    g = FockDriver()._compute_fock_omp(scr, dm, label, xf, 0.63, 15).to_numpy()
    ref = load('h2o.dimer.tzvp.erf.k.npy') * (xf if label == 'kx_rs' else 1.0)
    assert reference_equal(g, ref)


def test_erf_j_vs_oracle():
    """j_rs on a 4-water cluster (every class through d) against the oracle at 1e-10."""
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    D = pseudo_density(bas.get_dimensions_of_basis(), 20, seed=5)
    g = FockDriver().compute(scr, D, 'j_rs', 0.0, 0.4, 12)
    ref, _ = orc.fock_erf(orc.basis_from(bas, mol), D, 'j', 0.0, 0.4, 12)
    assert np.max(np.abs(np.asarray(g) - ref)) < TOL


def test_every_variant_matches_oracle():
    """Every kernel variant forced for all classes (untiled, tiled 4-wave and wide blocks over
    primitive runs and over single primitives, matrix-core MD) gives the oracle's 2jk at 1e-10 on a
    4-water cluster and the reference's def2-TZVP dimer fixture (f shells), each variant run after
    the others (the order in which masked lanes once carried non-finite values)."""
    from veloxchem_amd import _native
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    D = pseudo_density(bas.get_dimensions_of_basis(), 20, seed=11)
    ref, _ = orc.fock(orc.basis_from(bas, mol), D, '2jk', 1.0, 12, 0)
    drv = FockDriver()
    ctx = _native.context()
    try:
        tz = None
        if _native.lib().vlx_max_angular_momentum() >= 3:
            tmol, tbas = dimer('def2-tzvp')
            tz = (_screened(tmol, tbas), load('h2o.dimer.tzvp.density.npy'),
                  2 * load('h2o.dimer.tzvp.j.npy') - load('h2o.dimer.tzvp.k.npy'))
        # four side streams, then the whole sequence again on one stream (the configuration in
        # which the def2-TZVP p-d elements once came out non-finite)
        for nstreams in (4, 1):
            _native.check(_native.lib().vlx_context_set_streams(ctx, nstreams))
            for v, name in enumerate(FockDriver.VARIANTS):
                _native.check(_native.lib().vlx_context_set_variant(ctx, v))
                g = np.asarray(drv.compute(scr, D, '2jk', 1.0, 0.0, 12))
                err = np.max(np.abs(g - ref))
                print(f'{nstreams} streams, variant {name}: max|dG| {err:.2e}', flush=True)
                assert np.isfinite(g).all(), name
                assert err < TOL, name
                if tz is not None:  # f shells: the reference's def2-TZVP dimer fixture
                    gt = np.asarray(drv.compute(tz[0], tz[1], '2jk', 0.0, 0.0, 15))
                    assert np.isfinite(gt).all(), name
                    assert reference_equal(gt, tz[2]), name
    finally:
        _native.check(_native.lib().vlx_context_set_variant(ctx, -1))
        _native.check(_native.lib().vlx_context_set_streams(ctx, _native.default_streams()))


def test_batch_matches_single_builds():
    """vlx_fock_compute_batch_device over clusters of 1-6 waters of uneven size (each molecule
    whole on one stream, concurrently) gives each molecule's single-build G, and the oracle's."""
    drv = FockDriver()
    scrs, dens, mols = [], [], []
    for i, nw in enumerate((3, 1, 6, 2, 4, 1)):
        mol = water_cluster(nw, 6.0, 40 + i)
        bas = MolecularBasis.read(mol, 'def2-svp')
        scrs.append(_screened(mol, bas))
        dens.append(pseudo_density(bas.get_dimensions_of_basis(), 5 * nw, seed=60 + i))
        mols.append((mol, bas))
    for label, xf in (('2jk', 1.0), ('j', 0.0), ('kx', 0.3)):
        batch = drv.compute_batch(scrs, dens, label, xf, 0.0, 12)
        for scr, D, g in zip(scrs, dens, batch):
            single = np.asarray(drv.compute(scr, D, label, xf, 0.0, 12))
            assert np.max(np.abs(g - single)) < 1e-12, label
    mol, bas = mols[2]
    ref, _ = orc.fock(orc.basis_from(bas, mol), dens[2], '2jk', 1.0, 12, 0)
    g = drv.compute_batch(scrs, dens, '2jk', 1.0, 0.0, 12)[2]
    assert np.max(np.abs(g - ref)) < TOL
    assert drv.compute_batch([], [], '2jk', 1.0, 0.0, 12) == []
    # one screener twice with different densities: the two builds share its accumulator in turn
    twice = drv.compute_batch([scrs[0], scrs[2], scrs[0]], [dens[0], dens[2], 2.0 * dens[0]], 'j', 0.0, 0.0, 12)
    assert np.max(np.abs(twice[2] - 2.0 * twice[0])) < 1e-12
    with pytest.raises(ValueError):
        drv.compute_batch(scrs[:1], [dens[0] + np.triu(np.ones_like(dens[0]), 1)], '2jk', 1.0, 0.0, 12)


def test_batch_graph_replays_and_mixed_precision():
    """The batched build is one captured graph per (screeners, label, threshold, precision): every
    replay, with single eager builds run in between, gives the single builds' G.  In mixed mode the
    precision split is decided on the device per molecule (max|D_m|) and matches the single mixed
    build's split, so the batched and single mixed G agree, and both stay within 1e-3 tau of FP64."""
    from veloxchem_amd import _native
    drv = FockDriver()
    scrs, dens = [], []
    for i, nw in enumerate((2, 5, 1, 3)):
        mol = water_cluster(nw, 6.0, 70 + i)
        bas = MolecularBasis.read(mol, 'def2-svp')
        scrs.append(_screened(mol, bas))
        dens.append((1.0 + i) * pseudo_density(bas.get_dimensions_of_basis(), 5 * nw, seed=80 + i))
    single = [np.asarray(drv.compute(s_, D, '2jk', 1.0, 0.0, 12)) for s_, D in zip(scrs, dens)]
    for rep in range(3):
        batch = drv.compute_batch(scrs, dens, '2jk', 1.0, 0.0, 12)
        err = max(float(np.max(np.abs(g - s_))) for g, s_ in zip(batch, single))
        print(f'replay {rep}: max|G_batch - G_single| {err:.2e}', flush=True)
        assert err < 1e-12
        np.asarray(drv.compute(scrs[1], dens[1], 'j', 0.0, 0.0, 12))  # eager work between replays
    tau = 1e-7
    _native.set_precision('mixed', tau)
    try:
        single32 = [np.asarray(drv.compute(s_, D, '2jk', 1.0, 0.0, 12)) for s_, D in zip(scrs, dens)]
        for rep in range(2):
            batch32 = drv.compute_batch(scrs, dens, '2jk', 1.0, 0.0, 12)
            for g, g1, g64 in zip(batch32, single32, single):
                assert np.max(np.abs(g - g1)) < 1e-10
                assert np.max(np.abs(g - g64)) <= 1e-3 * tau
    finally:
        _native.set_precision('fp64')
    # FP32 work was present: the all-FP32 split moves G away from FP64 by more than rounding
    _native.set_precision('mixed', 1e6)
    try:
        batch_all32 = drv.compute_batch(scrs, dens, '2jk', 1.0, 0.0, 12)
    finally:
        _native.set_precision('fp64')
    dev = max(float(np.max(np.abs(g - g64))) for g, g64 in zip(batch_all32, single))
    print(f'all-FP32 batched build: max|dG| vs FP64 {dev:.2e}', flush=True)
    assert 1e-12 < dev < 1e-4


def test_autotune_choices_keep_parity():
    """The autotuner times every kernel variant per class; the build it leaves behind still matches the oracle."""
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    D = pseudo_density(bas.get_dimensions_of_basis(), 20, seed=9)
    drv = FockDriver()
    choice = drv._autotune(scr, D, 12, repeats=1)
    assert choice and set(choice.values()) <= set(FockDriver.VARIANTS)
    # every variant's warm-up G agreed with the untiled one, so none was rejected
    from veloxchem_amd import _native
    mask = ctypes.c_int(-1)
    _native.check(_native.lib().vlx_fock_autotune_rejected(_native.context(), ctypes.byref(mask)))
    assert mask.value == 0
    g = drv.compute(scr, D, '2jk', 1.0, 0.0, 12)
    ref, _ = orc.fock(orc.basis_from(bas, mol), D, '2jk', 1.0, 12, 0)
    assert np.max(np.abs(np.asarray(g) - ref)) < TOL
