"""The CPU oracle against the reference's fixtures and closed forms.

Fixtures (tests/golden/, copied from the reference's tests/data):
h2o.sto3g.{density,j,k,eri}.npy, h2o.dimer.svpd.{density,j,k}.npy,
h2o.dimer.tzvp.{density,j,k}.npy; screener maxima from
tests/test_t4c_screener.py:50-68 of the reference.
"""
import math

import numpy as np
import pytest
import scipy.special

from oracle import oracle as orc
from tests.common import dimer, load, reference_equal, sto3g_water


@pytest.mark.parametrize('m', [0, 1, 4, 8, 12])
@pytest.mark.parametrize('t', [0.0, 1e-9, 0.3, 2.5, 11.7, 30.0, 39.9, 40.1, 75.0, 300.0])
def test_boys_closed_form(m, t):
    # F_m(T) = 1F1(m + 1/2; m + 3/2; -T) / (2m + 1)
    ref = scipy.special.hyp1f1(m + 0.5, m + 1.5, -t) / (2 * m + 1)
    got = orc.boys(m, t)[m]
    assert abs(got - ref) <= 2e-14 * abs(ref)


def test_boys_limits():
    f = orc.boys(12, 0.0)
    assert np.allclose(f, [1.0 / (2 * m + 1) for m in range(13)], rtol=1e-15, atol=0)
    t = 500.0
    f = orc.boys(6, t)
    asym = [math.gamma(m + 0.5) / (2 * t ** (m + 0.5)) for m in range(7)]
    assert np.allclose(f, asym, rtol=1e-14, atol=0)


def test_ssss_closed_form():
    # (ss|ss) of one normalised s primitive on one centre: 2 sqrt(2a/pi) ... use the
    # textbook value for four identical unit-normalised s Gaussians: sqrt(2 a / pi) * 2 / sqrt(2)
    from veloxchem_amd import Molecule, MolecularBasis
    mol = Molecule.read_str('H 0 0 0', 'angstrom')
    bas = MolecularBasis.read(mol, 'def2-svp')
    sh = bas.shell_arrays(mol.get_coordinates_in_bohr())
    ob = orc.OracleBasis(sh, bas.get_dimensions_of_basis())
    a = sh['exps'][3]  # the single-primitive s shell
    val = orc.eri_sph(ob, 1, 1, 1, 1)[0, 0, 0, 0]
    # (aa|aa) for normalised s: 2 pi^{5/2}/(p q sqrt(p+q)) N^4, p = q = 2a, N = (2a/pi)^{3/4}
    p = 2 * a
    ref = 2 * math.pi ** 2.5 / (p * p * math.sqrt(2 * p)) * (2 * a / math.pi) ** 3
    assert abs(val - ref) < 1e-14 * ref


def test_eri_tensor_sto3g_fixture():
    # reference tests/test_fock_eri.py:26-36
    mol, bas = sto3g_water()
    eri = orc.eri_tensor(orc.basis_from(bas, mol))
    assert np.max(np.abs(eri - load('h2o.sto3g.eri.npy'))) < 1e-12


def test_eri_permutational_symmetry():
    mol, bas = dimer('def2-svp')
    ob = orc.basis_from(bas, mol)
    blk = orc.eri_sph(ob, 5, 2, 4, 3)  # (d p | p s)-type blocks of O
    assert np.allclose(blk, orc.eri_sph(ob, 2, 5, 4, 3).transpose(1, 0, 2, 3), atol=1e-15, rtol=0)
    assert np.allclose(blk, orc.eri_sph(ob, 4, 3, 5, 2).transpose(2, 3, 0, 1), atol=1e-15, rtol=0)


def test_screener_known_maxima():
    # reference tests/test_t4c_screener.py:50-68 (STO-3G water); shell order here:
    # O 1s(0), O 2s(1), O 2p(2), H 1s(3), H 1s(4)
    mol, bas = sto3g_water()
    mx, bins = orc.pair_maxima(orc.basis_from(bas, mol))
    ss = [4.78506575181571, 0.136873367356134, 0.817206295835749, 0.00785752036126304, 0.168705263230473,
          0.774605944211488, 0.00785752036126304, 0.168705263230473, 0.0395957018967669, 0.774605944211488]
    ss_idx = [(0, 0), (0, 1), (1, 1), (0, 3), (1, 3), (3, 3), (0, 4), (1, 4), (3, 4), (4, 4)]
    for v, (a, b) in zip(ss, ss_idx):
        assert abs(mx[a, b] - v) < 1e-12 * v
    sp = [0.0244774107811672, 0.180518387631544, 0.112067640667896, 0.112067640667896]
    for v, a in zip(sp, [0, 1, 3, 4]):
        assert abs(mx[min(a, 2), max(a, 2)] - v) < 1e-12 * v
    assert abs(mx[2, 2] - 0.880159089647115) < 1e-12
    # bins follow -floor(log10(sqrt(max))), 0 above 1
    assert bins[0, 0] == 0 and bins[0, 3] == int(-math.floor(math.log10(math.sqrt(ss[3]))))


@pytest.mark.parametrize('label', ['j', 'k'])
def test_fock_sto3g_fixture(label):
    mol, bas = sto3g_water()
    ob = orc.basis_from(bas, mol)
    g, _ = orc.fock(ob, load('h2o.sto3g.density.npy'), label, 0.0, 15)
    assert reference_equal(g, load(f'h2o.sto3g.{label}.npy'))


def test_fock_sto3g_unscreened_2jk():
    mol, bas = sto3g_water()
    ob = orc.basis_from(bas, mol)
    g, _ = orc.fock(ob, load('h2o.sto3g.density.npy'), '2jk', 0.0, -1)
    ref = 2.0 * load('h2o.sto3g.j.npy') - load('h2o.sto3g.k.npy')
    assert reference_equal(g, ref)


@pytest.mark.parametrize('basis,tag', [('def2-svpd', 'svpd'), ('def2-tzvp', 'tzvp')])
def test_fock_dimer_fixtures(basis, tag):
    mol, bas = dimer(basis)
    ob = orc.basis_from(bas, mol)
    dm = load(f'h2o.dimer.{tag}.density.npy')
    j = load(f'h2o.dimer.{tag}.j.npy')
    k = load(f'h2o.dimer.{tag}.k.npy')
    g, _ = orc.fock(ob, dm, '2jk', 0.0, 15)
    assert reference_equal(g, 2.0 * j - k)
    g, _ = orc.fock(ob, dm, '2jkx', 0.38, 15)
    assert reference_equal(g, 2.0 * j - 0.38 * k)
    g, _ = orc.fock(ob, dm, 'kx', 0.23, 15)
    assert reference_equal(g, 0.23 * k)


@pytest.mark.parametrize('label,xf', [('j', 0.0), ('k', 0.0), ('2jk', 0.0), ('2jkx', 0.31), ('kx', 0.27)])
def test_fock_general_density_fixtures(label, xf):
    """Non-symmetric density: reference tests/test_fock_driver.py:60-140 and :320-360 (def2-SVPD
    dimer, ithreshold 15) compare J with j.gen and K with the transposed k.gen fixture."""
    mol, bas = dimer('def2-svpd')
    ob = orc.basis_from(bas, mol)
    dm = load('h2o.dimer.svpd.density.gen.npy')
    j = load('h2o.dimer.svpd.j.gen.npy')
    kt = load('h2o.dimer.svpd.k.gen.npy').T
    ref = {'j': j, 'k': kt, '2jk': 2.0 * j - kt, '2jkx': 2.0 * j - xf * kt, 'kx': xf * kt}[label]
    g, _ = orc.fock_general(ob, dm, label, xf, 15)
    assert reference_equal(g, ref)


def test_fock_general_reduces_to_symmetric():
    # a symmetric density through the general path gives the symmetric build exactly
    mol, bas = dimer('def2-svpd')
    ob = orc.basis_from(bas, mol)
    dm = load('h2o.dimer.svpd.density.npy')
    g1, _ = orc.fock(ob, dm, '2jk', 0.0, 15)
    g2, _ = orc.fock_general(ob, dm, '2jk', 0.0, 15)
    assert np.max(np.abs(g1 - g2)) < 1e-13


def test_planted_mistake_is_caught():
    # a dropped exchange factor in the label handling must fail the fixture check
    mol, bas = sto3g_water()
    ob = orc.basis_from(bas, mol)
    g, _ = orc.fock(ob, load('h2o.sto3g.density.npy'), '2jkx', 0.5, 15)
    ref = 2.0 * load('h2o.sto3g.j.npy') - load('h2o.sto3g.k.npy')
    assert not reference_equal(g, ref)


def test_fock_erf_fixture():
    """Range-separated exchange: reference tests/test_fock_erf_driver.py:30-78 (k_rs, omega 0.63,
    ithreshold 15) and :81-127 (kx_rs, factor 0.21) on the def2-TZVP dimer."""
    mol, bas = dimer('def2-tzvp')
    ob = orc.basis_from(bas, mol)
    dm = load('h2o.dimer.tzvp.density.npy')
    ref = load('h2o.dimer.tzvp.erf.k.npy')
    g, _ = orc.fock_erf(ob, dm, 'k', 0.0, 0.63, 15)
    assert reference_equal(g, ref)
    g, _ = orc.fock_erf(ob, dm, 'kx', 0.21, 0.63, 15)
    assert reference_equal(g, 0.21 * ref)


def test_erf_integral_limits():
    """erf(w r)/r -> 1/r for large w, and -> (2 w / sqrt(pi)) <a|b><c|d> for small w."""
    mol, bas = dimer('def2-svp')
    ob = orc.basis_from(bas, mol)
    from veloxchem_amd.oneeints import compute_overlap_integrals
    S = compute_overlap_integrals(mol, bas)
    st = bas.shell_arrays()
    quartets = [(0, 1, 2, 3), (3, 3, 4, 4), (1, 5, 7, 8), (6, 6, 6, 6)]
    for q in quartets:
        coul = orc.eri_sph(ob, *q)
        big = orc.eri_sph_erf(ob, *q, 1.0e5)
        assert np.allclose(big, coul, rtol=1e-8, atol=1e-12)
        w = 1.0e-5
        small = orc.eri_sph_erf(ob, *q, w)
        base = [st['ao_base'][x] for x in q]
        stride = [st['ao_stride'][x] for x in q]
        dims = small.shape
        for i in range(dims[0]):
            for j in range(dims[1]):
                for k in range(dims[2]):
                    for l in range(dims[3]):
                        p_, q_ = base[0] + i * stride[0], base[1] + j * stride[1]
                        r_, s_ = base[2] + k * stride[2], base[3] + l * stride[3]
                        expect = 2.0 * w / np.sqrt(np.pi) * S[p_, q_] * S[r_, s_]
                        assert abs(small[i, j, k, l] - expect) < 1e-9 * w
