"""ctypes binding of libdgfs_hip.so (the C ABI declared in include/dgfs_hip.h).

The library is the only compute path of this package: there is no CPU
fallback.  ``lib()`` raises if the shared object is missing or a symbol is
absent, and every wrapper raises ``DGFSError`` with the library's message when
a call returns a non-zero status.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, 'libdgfs_hip.so')

F64, F32 = 0, 1
VHS, VHS_GLL = 0, 1
BC_COPY, BC_DIFFUSE, BC_DIFFUSE_EXPR, BC_INLET, BC_VANISH = 0, 1, 2, 3, 4

_vp, _i, _d, _sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_szp = C.POINTER(C.c_size_t)



class Wall(C.Structure):
    """dgfs_wall_t"""
    _fields_ = [('kind', C.c_int), ('nl', C.c_double), ('ushift', C.c_double),
                ('f0', C.c_void_p), ('ux', C.c_double), ('uy', C.c_double),
                ('uz', C.c_double), ('T', C.c_double)]


class Transport(C.Structure):
    """dgfs_transport_t"""
    _fields_ = [('dtype', C.c_int), ('K', C.c_int), ('Ne', C.c_int), ('Nv', C.c_int),
                ('D', C.c_void_p), ('BL', C.c_void_p), ('BR', C.c_void_p),
                ('gLD', C.c_void_p), ('gRD', C.c_void_p), ('invjac', C.c_void_p),
                ('cx', C.c_void_p), ('cy', C.c_void_p), ('cz', C.c_void_p),
                ('cw', C.c_double), ('left', Wall), ('right', Wall),
                ('work', C.c_void_p), ('nw', C.c_void_p), ('Dm', C.c_void_p),
                ('De', C.c_void_p)]


_l = C.c_long
# name -> (restype, argtypes); the authoritative list of exported entry points
IMEX_MAXS = 8


class LM(C.Structure):
    _fields_ = [('nstages', C.c_int), ('dt', C.c_double), ('prefac', C.c_double),
                ('omega', C.c_double), ('Pr', C.c_double),
                ('a', C.c_double * IMEX_MAXS), ('b', C.c_double * IMEX_MAXS),
                ('c', C.c_double * (IMEX_MAXS + 1)),
                ('U', C.c_void_p * IMEX_MAXS), ('LU', C.c_void_p * IMEX_MAXS),
                ('LM', C.c_void_p * IMEX_MAXS), ('Unew', C.c_void_p),
                ('F', C.c_void_p * IMEX_MAXS), ('L', C.c_void_p * IMEX_MAXS),
                ('Pn', C.c_void_p * IMEX_MAXS), ('Mnew', C.c_void_p), ('fnew', C.c_void_p),
                ('srate', C.c_double)]


class ARS(C.Structure):
    _fields_ = [('stage', C.c_int), ('dt', C.c_double), ('prefac', C.c_double),
                ('omega', C.c_double), ('Pr', C.c_double),
                ('ah', C.c_double * IMEX_MAXS), ('ai', C.c_double * (IMEX_MAXS + 1)),
                ('U', C.c_void_p * (IMEX_MAXS + 1)), ('LU', C.c_void_p * IMEX_MAXS),
                ('F', C.c_void_p * (IMEX_MAXS + 1)), ('L', C.c_void_p * IMEX_MAXS),
                ('M', C.c_void_p * (IMEX_MAXS + 1)), ('srate', C.c_double)]


class ClosureNet(C.Structure):
    _fields_ = [('W1', C.c_void_p), ('b1', C.c_void_p), ('W2f', C.c_void_p), ('W2t', C.c_void_p),
                ('b2', C.c_void_p), ('W3f', C.c_void_p), ('W3t', C.c_void_p), ('b3', C.c_void_p),
                ('W4', C.c_void_p), ('b4', C.c_void_p)]


class ClosureGrad(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ('W1', 'b1', 'W2', 'b2', 'W3', 'b3', 'W4', 'b4')]


SIGNATURES = {
    'dgfs_last_error': (C.c_char_p, []),
    'dgfs_collision_plan_create': (_i, [_i, _i, _i, _d, _dp, _dp, _dp, _d, _d, _d, _i, _i, _i,
                                        C.POINTER(_vp)]),
    'dgfs_collision_plan_destroy': (_i, [_vp]),
    'dgfs_collision_plan_query': (_i, [_vp, _ip, _ip, _ip, _szp]),
    'dgfs_collision_plan_set_qgroups': (_i, [_vp, _i]),
    'dgfs_collision_plan_set_schedule': (_i, [_vp, _i, _i]),
    'dgfs_collide': (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _i,
                          _vp]),
    'dgfs_collision_reference_set': (_i, [_vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'dgfs_collide_reference': (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _vp,
                                    _vp, _vp]),
    'dgfs_collision_timing': (_i, [_vp, _i]),
    'dgfs_bi_collision_plan_create': (_i, [_i, _i, _i, _d, _dp, _dp, _dp, _d, _i, _dp, _d, _d,
                                           _d, _d, _d, _i, _i, C.POINTER(_vp)]),
    'dgfs_bi_collision_plan_destroy': (_i, [_vp]),
    'dgfs_bi_collision_plan_query': (_i, [_vp, _ip, _ip, _szp]),
    'dgfs_bi_collide': (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    'dgfs_collision_timing_read': (_i, [_vp, _dp, _ip]),
    'dgfs_limit_linear': (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'dgfs_limit_positivity': (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'dgfs_transport_work_bytes': (_sz, [_i]),
    'dgfs_transport_rhs': (_i, [C.POINTER(Transport), _vp, _vp, _d, _vp]),
    'dgfs_transport_batch_create': (_i, [C.POINTER(Transport), _i, C.POINTER(_l),
                                         C.POINTER(_vp)]),
    'dgfs_transport_batch_destroy': (_i, [_vp]),
    'dgfs_transport_batch_rhs': (_i, [_vp, _vp, _vp, _vp]),
    'dgfs_transport_batch_nw': (_i, [_vp, _dp]),
    'dgfs_matop': (_i, [_i, _i, _i, _l, _vp, _vp, _vp, _i, _vp]),
    'dgfs_axnpby': (_i, [_i, _l, _i, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp]),
    'dgfs_axnpby_compensated': (_i, [_l, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d,
                                     _d, _d, _vp]),
    'dgfs_moments_work_bytes': (_sz, [_i, _i]),
    'dgfs_moments': (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    'dgfs_residual_work_bytes': (_sz, []),
    'dgfs_residual': (_i, [_i, _l, _vp, _vp, _vp, _vp, _vp]),
    'dgfs_projection_work_bytes': (_sz, [_i, _i]),
    'dgfs_project_conservative': (_i, [_i, _i, _i, _vp, _vp, _vp, _dp, _vp, _vp, _i, _vp, _vp]),
    'dgfs_project_conservative_noalloc': (_i, [_i, _i, _i, _vp, _vp, _vp, _dp, _vp, _vp, _i,
                                               _vp]),
    'dgfs_projection_mixture_work_bytes': (_sz, [_i, _i]),
    'dgfs_project_conservative_mixture': (_i, [_i, _i, _i, _vp, _vp, _vp, _d, _d, _dp, _vp, _vp,
                                               _vp, _vp, _i, _vp, _vp]),
    'dgfs_shakov_work_bytes': (_sz, [_i, _i]),
    'dgfs_mom14_work_bytes': (_sz, [_i, _i]),
    'dgfs_mom14': (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    'dgfs_shakov_target': (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp]),
    'dgfs_esbgk_target': (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _d, _vp, _vp]),
    'dgfs_relax_collide': (_i, [_i, _i, _i, _vp, _vp, _vp, _d, _d, _vp, _vp]),
    'dgfs_imex_lm_update': (_i, [_i, _i, _i, C.POINTER(LM), _i, _vp]),
    'dgfs_imex_ars_update': (_i, [_i, _i, _i, C.POINTER(ARS), _i, _vp]),
    'dgfs_closure_work_bytes': (_sz, [_i, _i, _i]),
    'dgfs_closure_forward': (_i, [C.POINTER(ClosureNet), _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp,
                                  _vp, _vp, _vp, _vp]),
    'dgfs_closure_backward': (_i, [C.POINTER(ClosureNet), C.POINTER(ClosureGrad), _i, _i, _i,
                                   _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp]),
    'dgfs_closure_pack': (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'dgfs_closure_loss_work_bytes': (_sz, [_i, _i]),
    'dgfs_closure_loss': (_i, [_i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp]),
    'dgfs_adam': (_i, [_i, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp]),
    'dgfs_shakov': (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                         _d, _d, _d, _d, _vp, _vp, _vp]),
    'dgfs_esbgk_work_bytes': (_sz, [_i, _i]),
    'dgfs_bgk_work_bytes': (_sz, [_i, _i]),
    'dgfs_esbgk_conservative_work_bytes': (_sz, [_i, _i]),
    'dgfs_esbgk_conservative': (_i, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _d, _d, _d, _d, _vp, _vp, _vp, _vp]),
    'dgfs_bgk_conservative': (_i, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, _vp]),
    'dgfs_esbgk': (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp,
                        _d, _d, _d, _d, _vp, _vp, _vp]),
}


class DGFSError(RuntimeError):
    pass


_LIB = None


def lib():
    """Load (once) and return the ctypes handle; raise if unavailable."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise DGFSError(
                'libdgfs_hip.so not found at %s: build it with '
                '`python -m dgfs1d_gpu_amd.build` (no CPU fallback exists)' % LIB_PATH)
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(status):
    if status != 0:
        msg = lib().dgfs_last_error()
        raise DGFSError((msg or b'unknown error').decode())


def ptr(t):
    """Device pointer of a contiguous torch tensor (None -> NULL)."""
    if t is None:
        return C.c_void_p(None)
    if not t.is_contiguous():
        raise DGFSError('tensor must be contiguous')
    return C.c_void_p(t.data_ptr())


def dtype_code(torch_dtype):
    import torch
    if torch_dtype == torch.float64:
        return F64
    if torch_dtype == torch.float32:
        return F32
    raise DGFSError('unsupported dtype %s' % torch_dtype)


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def dbl_array(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)
