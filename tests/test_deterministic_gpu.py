"""GPU tests of the deterministic reduction mode (vlx_set_deterministic): every F' contribution added as
a 128-bit fixed-point integer with integer atomics, so that a build gives the same bits on every run
and for any number of streams.  It must agree with the oracle like the default (atomic FP64) path and
with that path to round-off, and the SCF it drives must repeat bit for bit.  The reference merges its
task-local blocks under an OpenMP critical section (src/t4c_func/T4CMatrixDistributor.cpp:142) in an
order that varies between runs."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.common import dimer, load, reference_equal
from veloxchem_amd import FockDriver, Matrices, MolecularBasis, T4CScreener, make_matrix, mat_t
from veloxchem_amd.workloads import pseudo_density, water_cluster

pytestmark = pytest.mark.gpu


def _screened(mol, bas):
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    return scr


class _Deterministic:
    def __enter__(self):
        from veloxchem_amd import _native
        _native.set_deterministic(True)

    def __exit__(self, *exc):
        from veloxchem_amd import _native
        _native.set_deterministic(False)


def test_deterministic_build_is_bitwise_repeatable_and_matches_oracle():
    """4-water def2-SVP: three builds on four streams and one on a single stream give identical bits;
    they match the oracle at 1e-10 and the atomic path at round-off; a general density too."""
    from veloxchem_amd import _native
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    n = bas.get_dimensions_of_basis()
    D = pseudo_density(n, 20, seed=11)
    ob = orc.basis_from(bas, mol)
    ref, _ = orc.fock(ob, D, '2jkx', 0.3, 12)
    drv = FockDriver()
    atomic = np.asarray(drv.compute(scr, D, '2jkx', 0.3, 0.0, 12))
    lib, ctx = _native.lib(), _native.context()
    with _Deterministic():
        runs = [np.asarray(drv.compute(scr, D, '2jkx', 0.3, 0.0, 12)) for _ in range(3)]
        _native.check(lib.vlx_context_set_streams(ctx, 1))
        try:
            runs.append(np.asarray(drv.compute(scr, D, '2jkx', 0.3, 0.0, 12)))
        finally:
            _native.check(lib.vlx_context_set_streams(ctx, _native.default_streams()))
        rng = np.random.default_rng(3)
        Dg = D + 0.05 * rng.standard_normal((n, n))
        dm = make_matrix(bas, mat_t.general)
        dm.set_values(Dg)
        gen = [drv.compute(scr, dm, '2jk', 0.0, 0.0, 12).to_numpy() for _ in range(2)]
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])
    assert np.max(np.abs(runs[0] - ref)) < 1e-10
    assert np.max(np.abs(runs[0] - atomic)) < 1e-12 * max(1.0, np.max(np.abs(atomic)))
    assert np.array_equal(gen[0], gen[1])
    ref_g, _ = orc.fock_general(ob, Dg, '2jk', 0.0, 12)
    assert np.max(np.abs(gen[0] - ref_g)) < 1e-10


def test_deterministic_multi_density_equals_single_builds_bitwise():
    """The multi-density pass in the deterministic mode digests the same contributions as single
    builds, and integer sums do not depend on the grouping: identical bits, per density."""
    mol = water_cluster(4, 6.0, 7)
    bas = MolecularBasis.read(mol, 'def2-svp')
    scr = _screened(mol, bas)
    n = bas.get_dimensions_of_basis()
    dens = [pseudo_density(n, 20, seed=s) for s in (21, 22, 23)]
    labels = ['2jk', 'j', 'kx']
    mats = Matrices()
    for i, d in enumerate(dens):
        m = make_matrix(bas)
        m.set_values(d)
        mats.add(m, str(i))
    drv = FockDriver()
    with _Deterministic():
        out = drv.compute(scr, mats, labels, 0.45, 0.0, 12)
        singles = [np.asarray(drv.compute(scr, d, lab, 0.45, 0.0, 12)) for d, lab in zip(dens, labels)]
    for i, s in enumerate(singles):
        assert np.array_equal(out.matrix(str(i)).to_numpy(), s), i


def test_deterministic_tzvp_fixture_and_erf():
    """f shells and erf integrals in the deterministic mode: the reference's def2-TZVP dimer fixtures."""
    from veloxchem_amd import _native
    if _native.lib().vlx_max_angular_momentum() < 3:
        pytest.skip('f-shell kernels not built')
    mol, bas = dimer('def2-tzvp')
    scr = _screened(mol, bas)
    d = load('h2o.dimer.tzvp.density.npy')
    drv = FockDriver()
    with _Deterministic():
        g = np.asarray(drv.compute(scr, d, '2jk', 0.0, 0.0, 15))
        k_rs = np.asarray(drv.compute(scr, d, 'k_rs', 0.0, 0.63, 15))
    assert reference_equal(g, 2 * load('h2o.dimer.tzvp.j.npy') - load('h2o.dimer.tzvp.k.npy'))
    assert reference_equal(k_rs, load('h2o.dimer.tzvp.erf.k.npy'))


def test_deterministic_scf_repeats_bit_for_bit():
    """Two deterministic SCFs of a water dimer: the same energy bits and iterations, and the energy of
    the default SCF within 1e-10 Eh."""
    from tests.test_scf import dimer as scf_dimer
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    mol = scf_dimer()
    bas = MolecularBasis.read(mol, 'def2-svp')
    energies, iters = [], []
    for det in (True, True, False):
        drv = ScfRestrictedDriver()
        drv.deterministic = det
        r = drv.compute(mol, bas)
        energies.append(r['scf_energy'])
        iters.append(r['num_iter'])
    assert energies[0] == energies[1] and iters[0] == iters[1]
    assert abs(energies[0] - energies[2]) < 1e-10


def test_deterministic_mode_refusals():
    """The batched build, the no-allocation build and the autotuner refuse the deterministic mode."""
    import ctypes
    from veloxchem_amd import _native
    mol, bas = dimer('def2-svpd')
    scr = _screened(mol, bas)
    d = load('h2o.dimer.svpd.density.npy')
    drv = FockDriver()
    with _Deterministic():
        with pytest.raises(ValueError):
            drv.compute_batch([scr], [d], '2jk', 0.0, 0.0, 15)
        with pytest.raises(ValueError):
            drv._autotune(scr, d, 15, 1)
        buf = _native.DeviceArray(d.shape)
        out = _native.DeviceArray(d.shape)
        try:
            buf.upload(d)
            rc = _native.lib().vlx_fock_compute_device_noalloc(_native.context(), scr.handle, buf.ptr, b'2jk', 0.0,
                                                               0.0, 15, out.ptr)
            assert rc != 0
        finally:
            buf.free()
            out.free()
