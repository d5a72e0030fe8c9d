"""Host side of the relaxation-type collision term and the conservation projection.

``Shakov`` evaluates the reference's 'shakov-direct-gll' model
(dgfs1D/std/scattering.py:1636-1833) for a whole batch of (mode, element)
tasks in two launches; ``ConservativeProjection`` removes the mass, momentum
and energy moments of collision output slices in FP64 (no reference
counterpart).  Both call the C-ABI of libdgfs_hip.so and never fall back to
host arithmetic.
"""
import logging

import numpy as np

from . import _lib
from .collision import task_table

log = logging.getLogger('dgfs1d_gpu_amd')


def _grid_as_device(cv, tdtype):
    """The velocity grid exactly as the kernels see it (rounded to tdtype)."""
    import torch
    np_dt = np.float64 if tdtype == torch.float64 else np.float32
    return np.asarray(cv, dtype=np_dt).astype(np.float64)


def collision_invariants(cv):
    """Phi = {1, c_x, c_y, c_z, |c|^2} as a (5, n) FP64 array."""
    cv = np.asarray(cv, dtype=np.float64)
    return np.vstack([np.ones(cv.shape[1]), cv, (cv * cv).sum(0)])


class ConservativeProjection:
    """out (+)= Q - Phi^T (Phi Phi^T)^{-1} Phi Q per slice (FP64 moments and solve)."""

    def __init__(self, cv, cx, cy, cz, tdtype, device):
        import torch
        self.torch = torch
        self.tdtype = tdtype
        self.device = device
        self.n = int(np.asarray(cv).shape[1])
        phi = collision_invariants(_grid_as_device(cv, tdtype))
        self.gram_inv = np.ascontiguousarray(np.linalg.inv(phi @ phi.T))
        self._gp = self.gram_inv.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double))
        self.cx, self.cy, self.cz = cx, cy, cz
        self._work = None
        self._buf = None

    def work(self, nslice):
        nb = _lib.lib().dgfs_projection_work_bytes(int(nslice), self.n)
        if self._work is None or self._work.numel() * 8 < nb:
            self._work = self.torch.empty(max(1, nb // 8), dtype=self.torch.float64,
                                          device=self.device)
        return self._work

    def buffer(self, like):
        """A zeroed scratch array shaped like ``like`` (kept between calls)."""
        if self._buf is None or self._buf.numel() != like.numel():
            self._buf = self.torch.empty_like(like)
        self._buf.zero_()
        return self._buf

    def __call__(self, Q, out, nslice, accumulate=True, stream=None):
        if Q.dtype != self.tdtype or out.dtype != self.tdtype:
            raise TypeError('projection operands must be %s' % self.tdtype)
        if Q.numel() < nslice * self.n or out.numel() < nslice * self.n:
            raise ValueError('projection operands smaller than nslice*Nv')
        w = self.work(nslice)
        _lib.check(_lib.lib().dgfs_project_conservative(
            _lib.dtype_code(self.tdtype), int(nslice), self.n, _lib.ptr(self.cx),
            _lib.ptr(self.cy), _lib.ptr(self.cz), self._gp, _lib.ptr(Q), _lib.ptr(out),
            int(bool(accumulate)), _lib.ptr(w), _lib.stream_ptr(stream)))


def mixture_constraints(cv, m0, m1):
    """(Phi0, Phi1): the six constraint rows of the binary mixture on each
    species' grid values, (mass 0, mass 1, m_p c, m_p |c|^2), each (6, n)."""
    phi = collision_invariants(cv)
    n = phi.shape[1]
    out = []
    for p, m in enumerate((m0, m1)):
        P = np.zeros((6, n))
        P[p] = 1.0
        P[2:] = m * phi[1:]
        out.append(P)
    return out


class MixtureProjection:
    """Binary-mixture conservation projection (dgfs_project_conservative_mixture):
    the Euclidean-nearest pair (Q0, Q1) with each species' mass and the
    mixture's momentum and energy unchanged by collisions (FP64 moments and
    solve)."""

    def __init__(self, cv, cx, cy, cz, m0, m1, tdtype, device):
        import torch
        self.torch = torch
        self.tdtype = tdtype
        self.device = device
        self.n = int(np.asarray(cv).shape[1])
        self.m0, self.m1 = float(m0), float(m1)
        P0, P1 = mixture_constraints(_grid_as_device(cv, tdtype), m0, m1)
        self.gram_inv = np.ascontiguousarray(np.linalg.inv(P0 @ P0.T + P1 @ P1.T))
        self._gp = self.gram_inv.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double))
        self.cx, self.cy, self.cz = cx, cy, cz
        self._work = None

    def __call__(self, Q0, Q1, out0, out1, nslice, accumulate=False, stream=None):
        for x in (Q0, Q1, out0, out1):
            if x.dtype != self.tdtype:
                raise TypeError('projection operands must be %s' % self.tdtype)
            if x.numel() < nslice * self.n:
                raise ValueError('projection operands smaller than nslice*Nv')
        nb = _lib.lib().dgfs_projection_mixture_work_bytes(int(nslice), self.n)
        if self._work is None or self._work.numel() * 8 < nb:
            self._work = self.torch.empty(max(1, nb // 8), dtype=self.torch.float64,
                                          device=self.device)
        _lib.check(_lib.lib().dgfs_project_conservative_mixture(
            _lib.dtype_code(self.tdtype), int(nslice), self.n, _lib.ptr(self.cx),
            _lib.ptr(self.cy), _lib.ptr(self.cz), self.m0, self.m1, self._gp, _lib.ptr(Q0),
            _lib.ptr(Q1), _lib.ptr(out0), _lib.ptr(out1), int(bool(accumulate)),
            _lib.ptr(self._work), _lib.stream_ptr(stream)))


# per-task status bits of the Newton solves (dgfs_bgk_conservative)
NEWTON_NONFINITE, NEWTON_NOT_CONVERGED, NEWTON_WAIT_TIMEOUT = 1, 2, 4


def tensor_grid_n(cx, cy, cz, n):
    """N if the n points (cx, cy, cz) are the tensor grid c0[i], c0[j], c0[k]
    at p = (i N + j) N + k (the layout of std/velocitymesh.py), compared
    exactly; 0 otherwise."""
    n = int(n)
    N = int(round(n ** (1.0 / 3.0)))
    if N < 1 or N ** 3 != n:
        return 0
    x, y, z = (np.asarray(a.detach().cpu().numpy() if hasattr(a, 'detach') else a).ravel()
               for a in (cx, cy, cz))
    if x.size != n or y.size != n or z.size != n:
        return 0
    c0 = z[:N]
    p = np.arange(n)
    ok = (np.array_equal(x, c0[p // (N * N)]) and np.array_equal(y, c0[(p // N) % N])
          and np.array_equal(z, c0[p % N]))
    return N if ok else 0


class Shakov:
    """Batched Shakov relaxation term Q += prefac rho T^{1-omega} (f_S - f);
    with target='esbgk' the ES-BGK term (dgfs_esbgk), f_S replaced by the
    anisotropic Gaussian of the node; with target='bgk' the conservative BGK
    term (dgfs_bgk_conservative), f_S replaced by the discrete Maxwellian with
    f's discrete mass, momentum and energy; with target='esbgk-conservative'
    the conservative ES-BGK term (dgfs_esbgk_conservative).

    The two conservative targets run Newton solves on the device.  With
    check=True (default) every call reads their per-task status back and
    raises RuntimeError('NaN encountered ...') when a solve produced a NaN or
    an infinity, as the reference does (std/scattering.py:689, :1084); a solve
    that stops at the iteration limit is logged, as the reference stops there
    without an error.  check=False leaves the call free of host
    synchronisation; ``check_status()`` then reads the last call's status.
    grid_n: the Newton solves' velocity layout (dgfs_bgk_conservative's
    grid_n); None tests the points once (tensor_grid_n), 0 sums over the
    points as given."""

    def __init__(self, n, cx, cy, cz, cw, prefac, omega, Pr, tdtype, device, target='shakov',
                 check=True, grid_n=None):
        fns = {'shakov': 'dgfs_shakov', 'esbgk': 'dgfs_esbgk', 'bgk': 'dgfs_bgk_conservative',
               'esbgk-conservative': 'dgfs_esbgk_conservative'}
        if target not in fns:
            raise ValueError('target must be one of %s' % ', '.join(sorted(fns)))
        self._fn = fns[target]
        self._wfn = ('dgfs_bgk' if target == 'bgk' else fns[target]) + '_work_bytes'
        self.newton = target in ('bgk', 'esbgk-conservative')
        self.check = bool(check)
        if not self.newton:
            self.grid_n = 0
        elif grid_n is None:
            self.grid_n = tensor_grid_n(cx, cy, cz, n)
        else:
            self.grid_n = int(grid_n)
        import torch
        self.torch = torch
        self.n = int(n)
        self.cx, self.cy, self.cz = cx, cy, cz
        self.cw, self.prefac, self.omega, self.Pr = float(cw), float(prefac), float(omega), float(Pr)
        self.tdtype = tdtype
        self.device = device
        self.tasks = None
        self._work = None
        self._status = None
        self._ntask_last = 0

    def set_tasks(self, K, Ne, tasks):
        """tasks: (out_slice, f_terms[, g_terms]); only the f terms are used
        (the reference's fs() reads d_arr_in only, scattering.py:1786-1790)."""
        tasks = [(t[0], t[1], t[1]) for t in tasks]
        self.tasks = task_table(K, Ne, tasks, self.tdtype, self.device)
        return self.tasks

    def work(self, ntask):
        nb = getattr(_lib.lib(), self._wfn)(max(1, int(ntask)), self.n)
        if self._work is None or self._work.numel() * 8 < nb:
            self._work = self.torch.empty(max(1, nb // 8), dtype=self.torch.float64,
                                          device=self.device)
        return self._work

    def __call__(self, u, Q, stream=None, tasks=None):
        t = self.tasks if tasks is None else tasks
        if t is None:
            raise RuntimeError('set_tasks() must be called first')
        if u.dtype != self.tdtype or Q.dtype != self.tdtype:
            raise TypeError('Shakov operands must be %s' % self.tdtype)
        nmin = t['K'] * t['Ne'] * self.n
        if u.numel() < nmin or Q.numel() < nmin:
            raise ValueError('Shakov operands smaller than K*Ne*Nv')
        pr = () if self._fn == 'dgfs_bgk_conservative' else (self.Pr,)
        st, gn = (), ()
        if self.newton:
            gn = (self.grid_n,)
            if self._status is None or self._status.numel() < t['ntask']:
                self._status = self.torch.zeros(max(1, t['ntask']), dtype=self.torch.int32,
                                                device=self.device)
            st = (_lib.ptr(self._status),)
            self._ntask_last = t['ntask']
        _lib.check(getattr(_lib.lib(), self._fn)(
            _lib.dtype_code(self.tdtype), t['K'], t['Ne'], self.n, *gn, t['ntask'],
            _lib.ptr(t['nterm']), _lib.ptr(t['slice']), _lib.ptr(t['coef']), t['ngroup'],
            _lib.ptr(t['gstart']), _lib.ptr(t['gslice']), _lib.ptr(u), _lib.ptr(self.cx),
            _lib.ptr(self.cy), _lib.ptr(self.cz), self.cw, self.prefac, self.omega, *pr,
            _lib.ptr(Q), *st, _lib.ptr(self.work(t['ntask'])), _lib.stream_ptr(stream)))
        if self.newton and self.check:
            self.check_status()

    def status(self):
        """Per-task status bits of the last call's Newton solves (host array;
        synchronises with the device)."""
        import numpy as np
        if not self.newton or self._status is None:
            return np.zeros(0, dtype=np.int32)
        return self._status[:self._ntask_last].cpu().numpy()

    def check_status(self):
        """Raise if the last call's Newton solves met a NaN or an infinity (or a
        work-group wait gave up); log the tasks that hit the iteration limit."""
        st = self.status()
        bad = np.flatnonzero(st & (NEWTON_NONFINITE | NEWTON_WAIT_TIMEOUT))
        if bad.size:
            kind = 'NaN encountered' if np.any(st[bad] & NEWTON_NONFINITE) else \
                'work-group wait gave up'
            raise RuntimeError('%s in the %s Newton solve of task(s) %s'
                               % (kind, self._fn[5:], ', '.join(str(int(i)) for i in bad[:8])))
        slow = np.flatnonzero(st & NEWTON_NOT_CONVERGED)
        if slow.size:
            log.warning('%s: Newton solve stopped at the iteration limit for %d task(s)',
                        self._fn[5:], slow.size)
        return st
