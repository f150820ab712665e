"""Loader of the native Fock library (libvlxfock.so, HIP for gfx950).

The product path has no fallback: if the library is missing or cannot
create a device context, every call raises.
"""

import ctypes
import os
import re
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / 'libvlxfock.so'
HEADER_PATH = _PKG.parent / 'include' / 'vlxfock.h'

_lib = None
_contexts = {}

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_ll_p = ctypes.POINTER(ctypes.c_longlong)


class NativeError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64 (SONAME
        # libamdhip64.so.7, loaded through the unversioned name).  Loaded after
        # this library's /opt/rocm copy, it would be a second runtime that
        # finds no GPU; loaded first, this library binds to the same one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise NativeError(f'{LIB_PATH} is missing: build it with __graft_entry__.build() '
                              '(the Fock build has no CPU fallback)')
        L = ctypes.CDLL(str(LIB_PATH))
        L.vlx_last_error.restype = ctypes.c_char_p
        L.vlx_version.restype = ctypes.c_char_p
        L.vlx_context_stream.restype = ctypes.c_void_p
        L.vlx_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.vlx_context_destroy.argtypes = [ctypes.c_void_p]
        L.vlx_context_stream.argtypes = [ctypes.c_void_p]
        L.vlx_context_synchronize.argtypes = [ctypes.c_void_p]
        L.vlx_context_set_streams.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vlx_context_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vlx_set_block_size_factor.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vlx_screener_create.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_double_p,
                                          _c_double_p, _c_double_p, _c_int_p, _c_int_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]
        L.vlx_screener_destroy.argtypes = [ctypes.c_void_p]
        L.vlx_screener_pair_maxima.argtypes = [ctypes.c_void_p, _c_double_p]
        L.vlx_screener_number_of_pairs.argtypes = [ctypes.c_void_p, _c_ll_p]
        L.vlx_fock_compute.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _c_double_p, ctypes.c_char_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_double_p]
        L.vlx_fock_compute_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.vlx_fock_compute_device_noalloc.argtypes = L.vlx_fock_compute_device.argtypes
        L.vlx_fock_compute_general.argtypes = L.vlx_fock_compute.argtypes
        L.vlx_fock_compute_batch_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                                    ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.vlx_fock_compute_general_device.argtypes = L.vlx_fock_compute_device.argtypes
        L.vlx_fock_compute_multi.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_double_p, _c_int_p,
                                             ctypes.POINTER(ctypes.c_char_p), ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, _c_double_p]
        L.vlx_fock_compute_multi_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                    _c_int_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        L.vlx_fock_work.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_ll_p, _c_ll_p,
                                    _c_double_p]
        L.vlx_context_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vlx_context_reset_times.argtypes = [ctypes.c_void_p]
        L.vlx_context_class_times.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_double_p, _c_ll_p]
        L.vlx_device_malloc.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]
        L.vlx_device_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.vlx_copy_to_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong]
        L.vlx_copy_to_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong]
        L.vlx_eri_quartets.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_double_p]
        L.vlx_fock_autotune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        _c_int_p, _c_int_p, _c_int_p]
        L.vlx_fock_autotune_rejected.argtypes = [ctypes.c_void_p, _c_int_p]
        L.vlx_fock_get_variants.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_double_p]
        L.vlx_fock_set_variants.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p,
                                            _c_double_p]
        L.vlx_fock_class_kernel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.vlx_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.vlx_set_density_screening.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.vlx_set_deterministic.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vlx_context_last_fp32_bin_sum.argtypes = [ctypes.c_void_p, _c_int_p]
        L.vlx_fock_precision_split.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_ll_p, _c_ll_p]
        L.vlx_fock_precision_classes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p,
                                                 _c_int_p, _c_ll_p]
        L.vlx_one_electron.argtypes = [ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_double_p, _c_double_p,
                                       _c_double_p, _c_int_p, _c_int_p, ctypes.c_int, ctypes.c_int, _c_double_p,
                                       _c_double_p, _c_double_p, _c_double_p, _c_double_p]
        _lib = L
    return _lib


def header_functions():
    """Names of the functions declared in the library's headers (include/*.h)."""
    names = set()
    for path in sorted(HEADER_PATH.parent.glob('*.h')):
        text = re.sub(r'/\*.*?\*/', '', path.read_text(), flags=re.S)
        names.update(re.findall(r'\b(vlx_[a-z0-9_]+)\s*\(', text))
    return sorted(names)


def check(rc):
    if rc != 0:
        msg = lib().vlx_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        if rc == -2:
            raise NotImplementedError(msg)
        raise NativeError(msg)


def default_streams():
    """The side-stream count a new context starts with (VLX_STREAMS, default 3, within 1..4)."""
    return max(1, min(4, int(os.environ.get('VLX_STREAMS', 3))))


def context(device=0):
    """Per-device context (stream + Boys table), created on first use."""
    if device not in _contexts:
        h = ctypes.c_void_p()
        check(lib().vlx_context_create(int(device), ctypes.byref(h)))
        _contexts[device] = h
    return _contexts[device]


PRECISION_MODES = {'fp64': 0, 'mixed': 1}


def set_precision(mode='fp64', threshold=0.0, device=0):
    """'fp64', or 'mixed': FP32 recursions for quartets whose Schwarz bound x max|D| < threshold."""
    if mode not in PRECISION_MODES:
        raise ValueError(f'unknown precision mode {mode!r}')
    check(lib().vlx_set_precision(context(device), PRECISION_MODES[mode], float(threshold)))


def set_density_screening(threshold=0.0, device=0):
    """Density-weighted screening of later builds: quartets whose Schwarz bound times the largest |D|
    of the blocks they read is below threshold are skipped (0: off, the reference's screening)."""
    check(lib().vlx_set_density_screening(context(device), float(threshold)))


def set_deterministic(enable=True, device=0):
    """Deterministic reduction mode of later builds on the device: fixed-point integer accumulation
    of every F' contribution, bitwise the same result on every run (vlx_set_deterministic)."""
    check(lib().vlx_set_deterministic(context(device), 1 if enable else 0))


def last_fp32_bin_sum(device=0):
    k = ctypes.c_int()
    check(lib().vlx_context_last_fp32_bin_sum(context(device), ctypes.byref(k)))
    return k.value


def class_kernel(screener, ithreshold, cls, device=0):
    """Name of the kernel that class cls = (la, lb, lc, ld) runs with, as profilers report it."""
    buf = ctypes.create_string_buffer(128)
    check(lib().vlx_fock_class_kernel(context(device), screener.handle, int(ithreshold), *[int(x) for x in cls],
                                      buf, len(buf)))
    return buf.value.decode()


def class_times(device=0):
    """{(la, lb, lc, ld): (total_ms, launches)} of the timed class kernels."""
    L = lib()
    ctx = context(device)
    n = ctypes.c_int()
    check(L.vlx_context_class_times(ctx, 0, ctypes.byref(n), None, None, None))
    cls = np.zeros(4 * max(n.value, 1), dtype=np.int32)
    ms = np.zeros(max(n.value, 1))
    cnt = np.zeros(max(n.value, 1), dtype=np.int64)
    check(L.vlx_context_class_times(ctx, n.value, ctypes.byref(n), cls.ctypes.data_as(_c_int_p),
                                    ms.ctypes.data_as(_c_double_p), cnt.ctypes.data_as(_c_ll_p)))
    return {tuple(int(x) for x in cls[4 * i:4 * i + 4]): (float(ms[i]), int(cnt[i])) for i in range(n.value)}


class DeviceArray:
    """float64 buffer in HBM of one device (no torch needed)."""

    def __init__(self, shape, device=0):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.device = device
        self.ptr = ctypes.c_void_p()
        check(lib().vlx_device_malloc(context(device), self.nbytes, ctypes.byref(self.ptr)))

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.nbytes == self.nbytes
        check(lib().vlx_copy_to_device(context(self.device), self.ptr, host.ctypes.data_as(ctypes.c_void_p),
                                       self.nbytes))

    def download(self):
        out = np.empty(self.shape)
        check(lib().vlx_copy_to_host(context(self.device), out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                     self.nbytes))
        return out

    def free(self):
        if self.ptr:
            check(lib().vlx_device_free(context(self.device), self.ptr))
            self.ptr = ctypes.c_void_p()


def current_device():
    return int(os.environ.get('VLX_DEVICE', '0'))


def as_int(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_c_int_p)


def as_double(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_c_double_p)
