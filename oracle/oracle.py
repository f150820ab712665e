"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product package.
"""

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

LABELS = {'2jk': 0, '2jkx': 1, 'j': 2, 'k': 3, 'kx': 4}


class _Basis(ctypes.Structure):
    _fields_ = [
        ('nshells', ctypes.c_int),
        ('l', ctypes.POINTER(ctypes.c_int)),
        ('nprim', ctypes.POINTER(ctypes.c_int)),
        ('prim_offset', ctypes.POINTER(ctypes.c_int)),
        ('exps', ctypes.POINTER(ctypes.c_double)),
        ('norms', ctypes.POINTER(ctypes.c_double)),
        ('centers', ctypes.POINTER(ctypes.c_double)),
        ('ao_base', ctypes.POINTER(ctypes.c_int)),
        ('ao_stride', ctypes.POINTER(ctypes.c_int)),
        ('nao', ctypes.c_int),
    ]


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(['make', '-s', '-C', str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / 'liboracle.so'
        if not path.exists():
            build()
        _LIB = ctypes.CDLL(str(path))
        _LIB.oracle_fock.restype = ctypes.c_long
        _LIB.oracle_boys.restype = None
    return _LIB


class OracleBasis:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, arrays, nao):
        self.arrays = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        a = self.arrays
        ip = lambda k: a[k].astype(np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        for k in ('l', 'nprim', 'prim_offset', 'ao_base', 'ao_stride'):
            a[k] = np.ascontiguousarray(a[k], dtype=np.int32)
        for k in ('exps', 'norms', 'centers'):
            a[k] = np.ascontiguousarray(a[k], dtype=np.float64)
        dp = lambda k: a[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        ipp = lambda k: a[k].ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        del ip
        self.struct = _Basis(len(a['l']), ipp('l'), ipp('nprim'), ipp('prim_offset'), dp('exps'),
                             dp('norms'), dp('centers'), ipp('ao_base'), ipp('ao_stride'), int(nao))


def basis_from(molbasis, molecule):
    return OracleBasis(molbasis.shell_arrays(molecule.get_coordinates_in_bohr()),
                       molbasis.get_dimensions_of_basis())


def boys(mmax, t):
    out = np.zeros(mmax + 1)
    lib().oracle_boys(ctypes.c_int(mmax), ctypes.c_double(t),
                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def eri_sph(ob, sa, sb, sc, sd):
    ls = ob.arrays['l']
    n = [2 * int(ls[s]) + 1 for s in (sa, sb, sc, sd)]
    out = np.zeros(n)
    rc = lib().oracle_eri_sph(ctypes.byref(ob.struct), sa, sb, sc, sd,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def eri_tensor(ob):
    n = ob.struct.nao
    out = np.zeros((n, n, n, n))
    rc = lib().oracle_eri_tensor(ctypes.byref(ob.struct), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return out


def pair_maxima(ob):
    ns = ob.struct.nshells
    mx = np.zeros((ns, ns))
    bins = np.zeros((ns, ns), dtype=np.int32)
    rc = lib().oracle_pair_bins(ctypes.byref(ob.struct), mx.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                bins.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0
    return mx, bins


def fock(ob, density, label='2jk', exchange_factor=1.0, ithreshold=15, nthreads=0):
    """G = F' + F'^T for a symmetric density; returns (G, number of shell quartets)."""
    n = ob.struct.nao
    dm = np.ascontiguousarray(density, dtype=np.float64)
    assert dm.shape == (n, n)
    g = np.zeros((n, n))
    nq = lib().oracle_fock(ctypes.byref(ob.struct), dm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           ctypes.c_int(LABELS[label]), ctypes.c_double(exchange_factor),
                           ctypes.c_int(ithreshold), ctypes.c_int(nthreads),
                           g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if nq < 0:
        raise ValueError('oracle_fock: bad input')
    return g, int(nq)


def fock_general(ob, density, label='2jk', exchange_factor=1.0, ithreshold=15, nthreads=0):
    """Build for a general (non-symmetric) density: J of its symmetric part plus the exchange
    of the whole density (oracle_fock_general); returns (G, number of shell quartets)."""
    n = ob.struct.nao
    dm = np.ascontiguousarray(density, dtype=np.float64)
    assert dm.shape == (n, n)
    g = np.zeros((n, n))
    L = lib()
    L.oracle_fock_general.restype = ctypes.c_long
    nq = L.oracle_fock_general(ctypes.byref(ob.struct), dm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ctypes.c_int(LABELS[label]), ctypes.c_double(exchange_factor),
                               ctypes.c_int(ithreshold), ctypes.c_int(nthreads),
                               g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if nq < 0:
        raise ValueError('oracle_fock_general: bad input')
    return g, int(nq)


def fock_erf(ob, density, label='k', exchange_factor=1.0, omega=0.0, ithreshold=15, nthreads=0):
    """Range-separated build (labels j/k/kx on erf(omega r)/r integrals): (G, quartets)."""
    n = ob.struct.nao
    dm = np.ascontiguousarray(density, dtype=np.float64)
    assert dm.shape == (n, n)
    g = np.zeros((n, n))
    L = lib()
    L.oracle_fock_erf.restype = ctypes.c_long
    nq = L.oracle_fock_erf(ctypes.byref(ob.struct), dm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           ctypes.c_int(LABELS[label]), ctypes.c_double(exchange_factor), ctypes.c_double(omega),
                           ctypes.c_int(ithreshold), ctypes.c_int(nthreads),
                           g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if nq < 0:
        raise ValueError('oracle_fock_erf: bad input')
    return g, int(nq)


def eri_sph_erf(ob, sa, sb, sc, sd, omega):
    """erf-attenuated spherical block (ab|erf(omega r)/r|cd)."""
    ls = [int(ob.arrays['l'][x]) for x in (sa, sb, sc, sd)]
    out = np.zeros([2 * l + 1 for l in ls])
    rc = lib().oracle_eri_sph_erf(ctypes.byref(ob.struct), sa, sb, sc, sd, ctypes.c_double(omega),
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if rc:
        raise ValueError('oracle_eri_sph_erf: bad input')
    return out


def fock_sampled(ob, density, label, exchange_factor, ithreshold, nthreads, stride, bins):
    """Partial build over bra pairs ip % stride == 0 with precomputed pair bins
    (from pair_maxima): (quartets computed, quartets in a full build)."""
    n = ob.struct.nao
    bins = np.ascontiguousarray(bins, dtype=np.int32)
    dm = np.ascontiguousarray(density, dtype=np.float64)
    g = np.zeros((n, n))
    total = ctypes.c_long()
    L = lib()
    L.oracle_fock_sampled.restype = ctypes.c_long
    nq = L.oracle_fock_sampled(ctypes.byref(ob.struct), dm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               ctypes.c_int(LABELS[label]), ctypes.c_double(exchange_factor),
                               ctypes.c_int(ithreshold), ctypes.c_int(nthreads), ctypes.c_int(stride),
                               bins.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(total))
    if nq < 0:
        raise ValueError('oracle_fock_sampled: bad input')
    return int(nq), int(total.value)


if os.environ.get('VLX_ORACLE_BUILD') == '1':
    build()
