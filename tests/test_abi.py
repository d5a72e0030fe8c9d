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


def test_collision_n48_is_fp32_only():
    """N = 48 single-species plans are built for FP32 only; FP64 is refused
    before any device work."""
    import ctypes as C
    from dgfs1d_gpu_amd import _lib
    h = C.c_void_p()
    qz = (C.c_double * 3)(1.0, 0.0, 0.0)
    st = _lib.lib().dgfs_collision_plan_create(48, 1, 1, 1.0, qz, qz, qz, 1.0, 0.0, 1.0,
                                               _lib.F64, _lib.VHS, 1, C.byref(h))
    assert st == 2 and b'DGFS_F32 only' in _lib.lib().dgfs_last_error()


def test_new_entry_points_reject_bad_arguments():
    """Argument checks of the mixture, limiter, projection and ES-BGK entry
    points fail before any device work (status 2, message kept)."""
    import ctypes as C
    from dgfs1d_gpu_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    one = (C.c_double * 3)(1.0, 0.0, 0.0)
    st = L.dgfs_bi_collision_plan_create(10, 1, 1, 1.0, one, one, one, 1.0, 1, one, 1.0, 1.0,
                                         1.0, 0.5, 1.0, 0, 1, C.byref(h))
    assert st == 2 and b'unsupported N=10' in L.dgfs_last_error()
    st = L.dgfs_bi_collision_plan_create(48, 1, 1, 1.0, one, one, one, 1.0, 1, one, 1.0, 1.0,
                                         1.0, 0.5, 1.0, 0, 1, C.byref(h))
    assert st == 2 and b'DGFS_F32 only' in L.dgfs_last_error()
    st = L.dgfs_bi_collision_plan_create(8, 1, 1, 1.0, one, one, one, 1.0, 1, one, -1.0, 1.0,
                                         1.0, 0.5, 1.0, 0, 1, C.byref(h))
    assert st == 2 and b'masses must be positive' in L.dgfs_last_error()
    assert L.dgfs_limit_linear(0, 0, 3, 8, None, None, None, None, None, None, None) == 2
    assert b'bad sizes' in L.dgfs_last_error()
    assert L.dgfs_limit_positivity(5, 3, 3, 8, None, None, None, None) == 2
    assert b'bad dtype' in L.dgfs_last_error()
    assert L.dgfs_project_conservative_mixture(0, 1, 8, None, None, None, 1.0, -1.0, None, None,
                                               None, None, None, 0, None, None) == 2
    assert b'masses must be positive' in L.dgfs_last_error()
    # ES-BGK: Pr must be positive; null operands are refused (no device work)
    assert L.dgfs_esbgk(0, 1, 1, 8, 1, None, None, None, 1, None, None, None, None, None, None,
                        1.0, 1.0, 0.81, 0.0, None, None, None) == 2
    assert b'Pr must be positive' in L.dgfs_last_error()
    assert L.dgfs_esbgk(0, 1, 1, 8, 1, None, None, None, 1, None, None, None, None, None, None,
                        1.0, 1.0, 0.81, 0.5, None, None, None) == 2
    assert b'null pointer' in L.dgfs_last_error()
    assert L.dgfs_esbgk_work_bytes(2, 600) == 2 * (3 * (13 + 6) + 32) * 8
    assert L.dgfs_bgk_conservative(0, 1, 1, 8, 0, 1, None, None, None, 1, None, None, None, None,
                                   None, None, 1.0, 1.0, 0.81, None, None, None, None) == 2
    assert b'null pointer' in L.dgfs_last_error()
    # grid_n must be 0 or the cube root of Nv
    assert L.dgfs_bgk_conservative(0, 1, 1, 8, 3, 1, None, None, None, 1, None, None, None, None,
                                   None, None, 1.0, 1.0, 0.81, None, None, None, None) == 2
    assert b'cube root' in L.dgfs_last_error()
    # per task: 3 blocks x 13 moment partials, 32 constants, then the Newton
    # partials (2 buffers x 64 work-groups x 30 sums) and 64 doubles holding
    # the two sets of 64 work-group flags
    assert L.dgfs_bgk_work_bytes(2, 600) == (2 * (3 * 13 + 32) + 2 * (2 * 64 * 30 + 64)) * 8
    assert L.dgfs_esbgk_conservative(0, 1, 1, 8, 2, 1, None, None, None, 1, None, None, None,
                                     None, None, None, 1.0, 1.0, 0.81, -1.0, None, None, None,
                                     None) == 2
    assert b'Pr must be positive' in L.dgfs_last_error()
    # ES-BGK: the larger of its BGK solve's 36 sums (30 + the fe tensor) and
    # its own 35 monomial sums per work-group
    assert L.dgfs_esbgk_conservative_work_bytes(2, 600) == \
        (2 * (3 * 13 + 48) + 2 * (2 * 64 * 36 + 64)) * 8
