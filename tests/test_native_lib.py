"""The C-ABI library loads and exports every function include/vlxfock.h declares (no GPU calls)."""
import ctypes

import pytest

from veloxchem_amd import _native


def test_library_built_and_exports_header():
    if not _native.LIB_PATH.exists():
        pytest.fail('libvlxfock.so is not built; run __graft_entry__.build()')
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    names = _native.header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_lmax():
    lib = _native.lib()
    assert b'gfx950' in lib.vlx_version()
    assert lib.vlx_max_angular_momentum() >= 2


def test_last_error_is_a_string():
    assert isinstance(_native.lib().vlx_last_error(), bytes)
