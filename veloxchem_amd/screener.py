"""T4CScreener: Schwarz partition of shell pairs, computed on the GPU.

Mirrors ``CT4CScreener`` (src/t4c_func/T4CScreener.hpp, .cpp:63-118;
exported at src/python/ExportT4CIntegrals.cpp:127-140).  ``partition`` builds
every shell pair, evaluates max over components of (ab|ab) with the gfx950
kernels, bins the pairs by -floor(log10 sqrt(max)) and drops those with
sqrt <= 1e-15 (src/orbdata/BlockedGtoPairBlock.cpp:32-42).
"""

import ctypes

import numpy as np

from . import _native


class T4CScreener:

    def __init__(self, other=None):
        self._handle = None
        self._nshells = 0
        self._nao = 0
        self._device = _native.current_device()
        self._key = None
        self._inputs = None       # (basis, molecule, label) of the last partition
        self._atoms = None        # partition_atom / partition_atom_pair: atoms the pairs must touch
        if other is not None:     # T4CScreener(other): a copy (src/python/ExportT4CIntegrals.cpp:130)
            self.__setstate__(other.__getstate__())

    def partition(self, basis, molecule, label='eri'):
        """Every shell pair of the basis (CT4CScreener::partition, src/t4c_func/T4CScreener.cpp:63-118)."""
        self._partition(basis, molecule, label)
        self._atoms = None
        return self

    def partition_atom(self, basis, molecule, label, iatom):
        """The pairs with a shell on atom ``iatom`` (CT4CScreener::partition_atom,
        src/t4c_func/T4CScreener.cpp:120-185).  The reference feeds such screeners to its
        geometric-derivative Fock drivers, not to FockDriver.compute, which refuses them here."""
        self._partition(basis, molecule, label)
        self._atoms = (int(iatom),)
        return self

    def partition_atom_pair(self, basis, molecule, label, iatom, jatom):
        """The pairs of a shell on atom ``iatom`` with a shell on atom ``jatom``
        (CT4CScreener::partition_atom_pair, src/t4c_func/T4CScreener.cpp:187-254); like
        partition_atom, for the derivative drivers only."""
        self._partition(basis, molecule, label)
        self._atoms = (int(iatom), int(jatom))
        return self

    @property
    def is_restricted(self):
        """True after partition_atom / partition_atom_pair."""
        return self._atoms is not None

    def _partition(self, basis, molecule, label='eri'):
        if label != 'eri':
            raise ValueError(f"T4CScreener.partition: label {label!r} is not supported (only 'eri')")
        self._release()
        self._inputs = (basis, molecule, label)
        sh = basis.shell_arrays(molecule.get_coordinates_in_bohr())
        self._keep = {}
        ptrs = {}
        for k in ('l', 'nprim', 'prim_offset', 'ao_base', 'ao_stride'):
            self._keep[k], ptrs[k] = _native.as_int(sh[k])
        for k in ('exps', 'norms', 'centers'):
            self._keep[k], ptrs[k] = _native.as_double(sh[k])
        ctx = _native.context(self._device)
        h = ctypes.c_void_p()
        nao = basis.get_dimensions_of_basis()
        _native.check(_native.lib().vlx_screener_create(
            ctx, len(sh['l']), ptrs['l'], ptrs['nprim'], ptrs['prim_offset'], ptrs['exps'], ptrs['norms'],
            ptrs['centers'], ptrs['ao_base'], ptrs['ao_stride'], nao, ctypes.byref(h)))
        self._handle = h
        self._nshells = len(sh['l'])
        self._nao = nao
        self._key = (basis.get_label(), np.asarray(molecule.get_coordinates_in_bohr()).tobytes())
        self._shell_l = np.asarray(sh['l'])
        self._shell_atom = np.asarray(sh['atom'])
        return self

    def _release(self):
        if self._handle is not None and _native._lib is not None:
            _native.lib().vlx_screener_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    @property
    def handle(self):
        if self._handle is None:
            raise RuntimeError('T4CScreener: call partition(basis, molecule, "eri") first')
        return self._handle

    def number_of_ao(self):
        return self._nao

    def pair_maxima(self):
        """nshells x nshells matrix of max over components of (ab|ab)."""
        out = np.zeros((self._nshells, self._nshells))
        _native.check(_native.lib().vlx_screener_pair_maxima(self.handle, out.ctypes.data_as(_native._c_double_p)))
        return out

    def number_of_pairs(self):
        n = ctypes.c_longlong()
        _native.check(_native.lib().vlx_screener_number_of_pairs(self.handle, ctypes.byref(n)))
        return n.value

    def work(self, ithreshold):
        """Per-class (la, lb, lc, ld) contracted and primitive quartet counts of one build."""
        L = _native.lib()
        n = ctypes.c_int()
        _native.check(L.vlx_fock_work(self.handle, int(ithreshold), 0, ctypes.byref(n), None, None, None, None))
        cls = np.zeros(4 * max(n.value, 1), dtype=np.int32)
        nq = np.zeros(max(n.value, 1), dtype=np.int64)
        npq = np.zeros(max(n.value, 1), dtype=np.int64)
        fl = np.zeros(max(n.value, 1))
        _native.check(L.vlx_fock_work(self.handle, int(ithreshold), n.value, ctypes.byref(n),
                                      cls.ctypes.data_as(_native._c_int_p), nq.ctypes.data_as(_native._c_ll_p),
                                      npq.ctypes.data_as(_native._c_ll_p), fl.ctypes.data_as(_native._c_double_p)))
        return [(tuple(int(x) for x in cls[4 * i:4 * i + 4]), int(nq[i]), int(npq[i]), float(fl[i]))
                for i in range(n.value)]

    def precision_split(self, ithreshold, k32):
        """(FP64, FP32) primitive slots of one build when bin sums >= k32 run in FP32."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _native.check(_native.lib().vlx_fock_precision_split(self.handle, int(ithreshold), int(k32), ctypes.byref(a),
                                                             ctypes.byref(b)))
        return a.value, b.value

    def precision_classes(self, ithreshold, k32):
        """{(la, lb, lc, ld): FP32 primitive slots} of one build when bin sums >= k32 run in FP32."""
        L = _native.lib()
        n = ctypes.c_int()
        _native.check(L.vlx_fock_precision_classes(self.handle, int(ithreshold), int(k32), 0, ctypes.byref(n), None,
                                                   None))
        cls = np.zeros(4 * max(n.value, 1), dtype=np.int32)
        p32 = np.zeros(max(n.value, 1), dtype=np.int64)
        _native.check(L.vlx_fock_precision_classes(self.handle, int(ithreshold), int(k32), n.value, ctypes.byref(n),
                                                   cls.ctypes.data_as(_native._c_int_p),
                                                   p32.ctypes.data_as(_native._c_ll_p)))
        return {tuple(int(x) for x in cls[4 * i:4 * i + 4]): int(p32[i]) for i in range(n.value)}

    def schedule_stats(self, ithreshold):
        """Per class (same order as work()): (lane slots, primitive slots) the kernels execute."""
        L = _native.lib()
        L.vlx_fock_schedule_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int), _native._c_ll_p, _native._c_ll_p]
        n = ctypes.c_int()
        _native.check(L.vlx_fock_schedule_stats(self.handle, int(ithreshold), 0, ctypes.byref(n), None, None))
        ls = np.zeros(max(n.value, 1), dtype=np.int64)
        ps = np.zeros(max(n.value, 1), dtype=np.int64)
        _native.check(L.vlx_fock_schedule_stats(self.handle, int(ithreshold), n.value, ctypes.byref(n),
                                                ls.ctypes.data_as(_native._c_ll_p),
                                                ps.ctypes.data_as(_native._c_ll_p)))
        return [(int(ls[i]), int(ps[i])) for i in range(n.value)]

    def gto_pair_blocks(self):
        """The kept shell pairs by angular class and Schwarz bin, as CT4CScreener::gto_pair_blocks
        lists them (src/t4c_func/T4CScreener.hpp:67; bins of src/orbdata/BlockedGtoPairBlock.cpp:32-42):
        one dict per (l_a <= l_b, bin) with the shell pairs (a, b), their max over components of
        (ab|ab) and the bin; a restricted screener lists only the pairs touching its atoms."""
        pm = self.pair_maxima()
        la, at = self._shell_l, self._shell_atom
        blocks = {}
        n = self._nshells
        for a in range(n):
            for b in range(a, n):
                if la[a] > la[b] or (la[a] == la[b] and a > b):
                    p, q = b, a
                else:
                    p, q = a, b
                if self._atoms is not None:
                    if len(self._atoms) == 1 and self._atoms[0] not in (at[p], at[q]):
                        continue
                    if len(self._atoms) == 2 and {at[p], at[q]} != set(self._atoms) and not (
                            self._atoms[0] == self._atoms[1] == at[p] == at[q]):
                        continue
                fact = float(np.sqrt(pm[p, q]))
                if not fact > 1.0e-15:
                    continue
                b_ = 0 if fact >= 1.0 else int(-np.floor(np.log10(fact)))
                blk = blocks.setdefault((int(la[p]), int(la[q]), b_), {
                    'angular_momentums': (int(la[p]), int(la[q])), 'bin': b_, 'pairs': [], 'max_values': []})
                blk['pairs'].append((p, q))
                blk['max_values'].append(float(pm[p, q]))
        return [blocks[k] for k in sorted(blocks)]

    def __getstate__(self):
        """Pickled as its inputs (src/python/ExportT4CIntegrals.cpp:131-132 pickles the pair
        blocks); unpickling partitions again on the current device."""
        return {'inputs': self._inputs, 'atoms': self._atoms}

    def __setstate__(self, state):
        self._handle = None
        self._nshells = 0
        self._nao = 0
        self._device = _native.current_device()
        self._key = None
        self._inputs = None
        self._atoms = None
        if state.get('inputs') is not None:
            self._partition(*state['inputs'])
            self._atoms = state.get('atoms')

    def __copy__(self):
        return type(self)(self)

    def __deepcopy__(self, memo):
        return type(self)(self)

    def __eq__(self, other):
        return isinstance(other, T4CScreener) and self._key == other._key and self._atoms == other._atoms

    def __ne__(self, other):
        return not self.__eq__(other)
