"""FockDriver: drop-in for VeloxChem's two-electron Fock build on MI355X.

Mirrors the Python facade src/pymodule/fockdriver.py:37-135 and the pybind11
``_FockDriver`` surface (src/python/ExportT4CIntegrals.cpp:48-124):

* ``compute(screener, D, label, exchange_factor, omega, ithreshold)`` ->
  ``Matrix`` (same angular layout and type as ``D``), the call
  scfdriver.py:1955-1957 makes with ``'2jk', 1.0, 0.0, thresh_int``;
* ``compute(screener, Ds: Matrices, labels: list, ...)`` -> ``Matrices``,
  one pass over the integrals for all densities (vlx_fock_compute_multi);
* ``_compute_fock_omp(screener, D, ...)`` and the unscreened
  ``_compute_fock_omp(basis, molecule, D, label, xf, omega)``;
* ``compute_eri(molecule, basis, eri_thresh)`` -> N^4 ndarray;
* ``_set_block_size_factor(f)``.

Labels: '2jk' -> 2J - K, '2jkx' -> 2J - xf K, 'j' -> J, 'k' -> K,
'kx' -> xf K, with J_pq = sum_rs (pq|rs) D_rs and K_pr = sum_qs (pq|rs) D_qs;
'j_rs', 'k_rs', 'kx_rs' the same on erf(omega r)/r integrals
(src/t4c_func/T4CMatrixDistributor.cpp:51-54).  omega is ignored by the
other labels, as in the reference.

A density Matrix of type mat_t.general (non-symmetric) takes
vlx_fock_compute_general: J of its symmetric part plus the exchange of the
whole density, K_pr = sum_qs (pq|rs) D_qs, as the reference's general path
(its tests: tests/test_fock_driver.py:320-360).

Deviations from the reference, by design: an unknown label raises
ValueError (the reference silently returns zeros,
src/t4c_func/T4CUtils.cpp:814-1062); ithreshold must be < 0 (unscreened)
or 0..30.
"""

import ctypes
import math

import numpy as np

from . import _native
from .matrix import Matrices, Matrix, mat_t
from .screener import T4CScreener

_ALLOWED_FACTORS = (1, 2, 4, 8, 16, 32, 64, 128)


def _dense(density):
    if isinstance(density, Matrix):
        return density.to_numpy(), density
    arr = np.asarray(density, dtype=np.float64)
    return arr, None


class FockDriver:

    def __init__(self, comm=None, ostream=None):
        self.comm = comm
        self.rank = 0 if comm is None else comm.Get_rank()
        self.nodes = 1 if comm is None else comm.Get_size()
        self.ostream = ostream
        self._block_size_factor = 8
        self._device = _native.current_device()

    # -- reference entry points -------------------------------------------------

    def compute(self, screener, *args):
        return self._compute_local_fock(screener, self.rank, self.nodes, *args)

    def compute_on_subcomm(self, subcomm, screener, *args):
        return self._compute_local_fock(screener, subcomm.Get_rank(), subcomm.Get_size(), *args)

    def _compute_fock_omp(self, *args):
        if isinstance(args[0], T4CScreener):
            return self._compute_local_fock(args[0], 0, 1, *args[1:])
        # (basis, molecule, D, label, exchange_factor, omega): every quartet, no screening
        basis, molecule = args[0], args[1]
        screener = T4CScreener()
        screener.partition(basis, molecule, 'eri')
        rest = list(args[2:])
        while len(rest) < 4:
            rest.append(0.0)
        return self._compute_local_fock(screener, 0, 1, rest[0], rest[1], rest[2], rest[3], -1)

    def _set_block_size_factor(self, factor):
        if factor not in _ALLOWED_FACTORS:
            raise ValueError('FockDriver._set_block_size_factor: Invalid factor')
        self._block_size_factor = int(factor)
        _native.check(_native.lib().vlx_set_block_size_factor(_native.context(self._device), int(factor)))

    VARIANTS = ('untiled', 'tiled4', 'tiled-wide', 'mmd', 'tiled4-p1', 'tiled-wide-p1', 'ks-lds', 'tiled-dlds',
                'tiled-2seg')

    def _autotune(self, screener, density, ithreshold=12, repeats=2):
        """Times each angular class under every kernel variant on this density and keeps
        the fastest per class (vlx_fock_autotune); returns {(la, lb, lc, ld): variant}."""
        dm = density.to_numpy() if hasattr(density, 'to_numpy') else np.asarray(density, dtype=np.float64)
        buf = _native.DeviceArray(dm.shape, self._device)
        try:
            buf.upload(dm)
            n = ctypes.c_int()
            cls = np.zeros(4 * 64, dtype=np.int32)
            var = np.zeros(64, dtype=np.int32)
            _native.check(_native.lib().vlx_fock_autotune(
                _native.context(self._device), screener.handle, buf.ptr, int(ithreshold), int(repeats),
                ctypes.byref(n), cls.ctypes.data_as(_native._c_int_p), var.ctypes.data_as(_native._c_int_p)))
        finally:
            buf.free()
        return {tuple(int(x) for x in cls[4 * r:4 * r + 4]): self.VARIANTS[int(var[r])] for r in range(n.value)}

    def compute_eri(self, molecule, basis, eri_thresh=1.0e-12):
        """Full AO ERI tensor (pq|rs) (src/t4c_func/FockDriver.cpp:76-117).

        Quartets whose binned Schwarz estimate fails ``int(-log10(eri_thresh))``
        are left zero, as the reference does.
        """
        screener = T4CScreener()
        screener.partition(basis, molecule, 'eri')
        thresh_int = int(-math.log10(eri_thresh))
        sh = basis.shell_arrays(molecule.get_coordinates_in_bohr())
        ls = sh['l']
        ns = len(ls)
        n = basis.get_dimensions_of_basis()
        maxv = screener.pair_maxima()

        def pair_bin(v):
            f = math.sqrt(v) if v > 0 else 0.0
            if f > 1.0e-15:
                return int(-math.floor(math.log10(f))) if f < 1.0 else 0
            return -1

        bins = np.vectorize(pair_bin)(maxv)
        pairs = [(a, b) for a in range(ns) for b in range(a, ns) if bins[a, b] >= 0]
        quartets = [(a, b, c, d) for i, (a, b) in enumerate(pairs) for (c, d) in pairs[i:]
                    if bins[a, b] + bins[c, d] <= thresh_int]
        blocks = self.eri_quartets(screener, quartets)
        eri = np.zeros((n, n, n, n))
        base, stride = sh['ao_base'], sh['ao_stride']
        for (a, b, c, d), blk in zip(quartets, blocks):
            ia = base[a] + stride[a] * np.arange(2 * ls[a] + 1)
            ib = base[b] + stride[b] * np.arange(2 * ls[b] + 1)
            ic = base[c] + stride[c] * np.arange(2 * ls[c] + 1)
            id_ = base[d] + stride[d] * np.arange(2 * ls[d] + 1)
            # all eight permutations of (ab|cd)
            t = blk
            for perm_blk, idx in ((t, (ia, ib, ic, id_)),
                                  (t.transpose(1, 0, 2, 3), (ib, ia, ic, id_)),
                                  (t.transpose(0, 1, 3, 2), (ia, ib, id_, ic)),
                                  (t.transpose(1, 0, 3, 2), (ib, ia, id_, ic)),
                                  (t.transpose(2, 3, 0, 1), (ic, id_, ia, ib)),
                                  (t.transpose(3, 2, 0, 1), (id_, ic, ia, ib)),
                                  (t.transpose(2, 3, 1, 0), (ic, id_, ib, ia)),
                                  (t.transpose(3, 2, 1, 0), (id_, ic, ib, ia))):
                eri[np.ix_(*idx)] = perm_blk
        return eri

    def compute_batch(self, screeners, densities, label, exchange_factor=0.0, omega=0.0, ithreshold=15):
        """G of many molecules in one call (vlx_fock_compute_batch_device).

        ``screeners[m]`` and ``densities[m]`` (symmetric, nao_m x nao_m, Matrix or array) describe
        molecule m; returns the list of G as numpy arrays.  One kernel launch per angular class
        covers every molecule, and the whole build is a HIP graph captured on the first call with
        these screeners and settings and replayed afterwards; the precision mode is the context's
        (mixed: the FP32 split per molecule from its max|D|).  No reference counterpart: the
        reference builds one molecule per ``compute`` call.
        """
        if len(screeners) != len(densities):
            raise ValueError('compute_batch: one density per screener')
        if ithreshold > 30:
            raise ValueError(f'FockDriver: ithreshold {ithreshold} outside 0..30')
        bufs = []
        try:
            d_ptrs, g_ptrs = [], []
            for scr, dens in zip(screeners, densities):
                dm, _ = _dense(dens)
                n = scr.number_of_ao()
                if dm.shape != (n, n):
                    raise ValueError(f'FockDriver: density has shape {dm.shape}, basis has {n} AOs')
                # the same test as the symmetric host entry point (vlx_fock_compute)
                if np.any(np.abs(dm - dm.T) > 1.0e-12 * np.maximum(1.0, np.maximum(np.abs(dm), np.abs(dm.T)))):
                    raise ValueError('compute_batch: density is not symmetric')
                d = _native.DeviceArray((n, n), self._device)
                bufs.append(d)
                d.upload(dm)
                g = _native.DeviceArray((n, n), self._device)
                bufs.append(g)
                d_ptrs.append(d.ptr.value)
                g_ptrs.append(g.ptr.value)
            nmol = len(screeners)
            arr = ctypes.c_void_p * max(nmol, 1)
            _native.check(_native.lib().vlx_fock_compute_batch_device(
                _native.context(self._device), nmol, arr(*[s.handle.value for s in screeners]), arr(*d_ptrs),
                label.encode(), ctypes.c_double(exchange_factor), ctypes.c_double(omega), int(ithreshold),
                arr(*g_ptrs)))
            return [bufs[2 * m + 1].download() for m in range(nmol)]
        finally:
            for b in bufs:
                b.free()

    # -- helpers ----------------------------------------------------------------

    def eri_quartets(self, screener, quartets):
        """Spherical (ab|cd) blocks of explicit shell quartets, computed on the GPU."""
        q = np.ascontiguousarray(np.array(quartets, dtype=np.int32).reshape(-1, 4))
        ls = screener._keep['l']
        sizes = [(2 * ls[a] + 1) * (2 * ls[b] + 1) * (2 * ls[c] + 1) * (2 * ls[d] + 1) for a, b, c, d in q]
        out = np.zeros(int(sum(sizes)))
        if len(q):
            _native.check(_native.lib().vlx_eri_quartets(
                _native.context(self._device), screener.handle, len(q), q.ctypes.data_as(_native._c_int_p),
                out.ctypes.data_as(_native._c_double_p)))
        blocks, off = [], 0
        for (a, b, c, d), sz in zip(q, sizes):
            blocks.append(out[off:off + sz].reshape(2 * ls[a] + 1, 2 * ls[b] + 1, 2 * ls[c] + 1, 2 * ls[d] + 1))
            off += sz
        return blocks

    def _compute_local_fock(self, screener, rank, nodes, *args):
        if len(args) < 2:
            raise TypeError('compute(screener, density, label, exchange_factor, omega, ithreshold)')
        density, label = args[0], args[1]
        xf = float(args[2]) if len(args) > 2 else 0.0
        omega = float(args[3]) if len(args) > 3 else 0.0
        ithreshold = int(args[4]) if len(args) > 4 else 15
        if isinstance(density, Matrices):
            if not isinstance(label, (list, tuple)):
                raise TypeError('a Matrices density needs a list of labels')
            keys = density.keys()
            if len(keys) != len(label):
                raise ValueError('number of densities and labels differ')
            return self._multi(screener, rank, nodes, density, keys, list(label), xf, omega, ithreshold)
        return self._one(screener, rank, nodes, density, label, xf, omega, ithreshold)

    def _check_call(self, screener, rank, nodes, ithreshold):
        if screener.is_restricted:
            raise ValueError('FockDriver: screeners of partition_atom / partition_atom_pair feed the geometric-'
                             'derivative drivers; the Fock build takes a partition() screener')
        if ithreshold > 30:
            raise ValueError(f'FockDriver: ithreshold {ithreshold} outside 0..30')
        if nodes != 1 or rank != 0:
            raise NotImplementedError('FockDriver: one process per GPU; rank/nodes partitioning is not built')

    def _multi(self, screener, rank, nodes, mats, keys, labels, xf, omega, ithreshold):
        """Matrices overload (src/t4c_func/FockDriver.cpp:173-229): one pass over the integrals
        digests every density with its own label (vlx_fock_compute_multi).  The integrals are
        erf(omega r)/r only when every label is range-separated, else 1/r for all, the *_rs labels
        digesting like their Coulomb forms (CT4CMatricesDistributor::need_omega,
        src/t4c_func/T4CMatricesDistributor.cpp:55-66)."""
        self._check_call(screener, rank, nodes, ithreshold)
        n = screener.number_of_ao()
        dens, general, mats_in = [], [], []
        for key, lab in zip(keys, labels):
            if not isinstance(lab, str):
                raise TypeError('label must be a string')
            dm, mat = _dense(mats.matrix(key))
            if dm.shape != (n, n):
                raise ValueError(f'FockDriver: density has shape {dm.shape}, basis has {n} AOs')
            dens.append(dm)
            mats_in.append(mat)
            general.append(1 if mat is not None and mat.get_type() == mat_t.general else 0)
        k = len(dens)
        d = np.ascontiguousarray(np.stack(dens).astype(np.float64, copy=False))
        g = np.zeros_like(d)
        flags = np.ascontiguousarray(np.array(general, dtype=np.int32))
        labs = (ctypes.c_char_p * k)(*[lab.encode() for lab in labels])
        _native.check(_native.lib().vlx_fock_compute_multi(
            _native.context(self._device), screener.handle, k, d.ctypes.data_as(_native._c_double_p),
            flags.ctypes.data_as(_native._c_int_p), labs, ctypes.c_double(xf), ctypes.c_double(omega),
            int(ithreshold), g.ctypes.data_as(_native._c_double_p)))
        out = Matrices()
        for i, (key, mat) in enumerate(zip(keys, mats_in)):
            if mat is not None:
                res = Matrix(mat._block_dims, mat.get_type())
                res._values = g[i].copy()
            else:
                res = g[i].copy()
            out.add(res, key)
        return out

    def _one(self, screener, rank, nodes, density, label, xf, omega, ithreshold):
        self._check_call(screener, rank, nodes, ithreshold)
        dm, mat = _dense(density)
        n = screener.number_of_ao()
        if dm.shape != (n, n):
            raise ValueError(f'FockDriver: density has shape {dm.shape}, basis has {n} AOs')
        if not isinstance(label, str):
            raise TypeError('label must be a string')
        dm = np.ascontiguousarray(dm)
        g = np.zeros((n, n))
        general = mat is not None and mat.get_type() == mat_t.general
        fn = _native.lib().vlx_fock_compute_general if general else _native.lib().vlx_fock_compute
        _native.check(fn(
            _native.context(self._device), screener.handle, dm.ctypes.data_as(_native._c_double_p),
            label.encode(), ctypes.c_double(xf), ctypes.c_double(omega), int(ithreshold),
            g.ctypes.data_as(_native._c_double_p)))
        if mat is not None:
            res = Matrix(mat._block_dims, mat.get_type())
            res._values = g
            return res
        return g
