"""libdgfs_hip.so loads and exports every function declared in
include/dgfs_hip.h; the ctypes signature table covers exactly that set.
No compute call is made (no GPU needed)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, 'include', 'dgfs_hip.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return set(re.findall(r'\b(dgfs_[a-z0-9_]+)\s*\(', text))


def test_library_exports_header():
    from dgfs1d_gpu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail('libdgfs_hip.so is not built (run python -m dgfs1d_gpu_amd.build)')
    h = _lib.lib()
    decl = _declared()
    assert decl, 'no declarations parsed'
    for name in decl:
        assert hasattr(h, name), name
    assert decl == set(_lib.SIGNATURES), decl ^ set(_lib.SIGNATURES)


def test_errors_are_reported():
    import ctypes as C
    from dgfs1d_gpu_amd import _lib
    h = C.c_void_p()
    qz = (C.c_double * 1)(1.0)
    st = _lib.lib().dgfs_collision_plan_create(7, 1, 1, 1.0, qz, qz, qz, 1.0, 0.0, 1.0, 0, 0, 1,
                                               C.byref(h))
    assert st == 2
    assert b'unsupported N=7' in _lib.lib().dgfs_last_error()
