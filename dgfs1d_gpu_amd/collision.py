"""Batched fast-spectral collision operator on the GPU (host side).

``CollisionPlan`` owns a ``dgfs_collision_plan_t`` (tables + scratch on the
device) and a device-resident task table.  A *task* evaluates one collision
integral Q(f_t, g_t), where f_t and g_t are linear combinations of slices of
the coefficient array u (K, Ne, Nv) -- the U[m]^T / V[m] contractions of the
SVD-decomposed H tensor (reference dgfs1D/basis.py:163-177, std/std.py:352-356)
-- and accumulates prefac * (gain - f_t * loss) into one output slice.
"""
import numpy as np

from . import _lib


def task_table(K, Ne, tasks, tdtype, device):
    """Device task table shared by the collision and relaxation entry points.

    tasks: (out_slice, f_terms, g_terms) with terms [(slice, coef), ...]; the
    table is sorted by output slice and grouped (one group per distinct output
    slice): nterm [T][2], slice [T][2][K], coef [T][2][K] (tdtype), group_start
    [G+1], group_slice [G], and fg_same = every task has identical f and g terms."""
    import torch
    tasks = sorted(tasks, key=lambda t: t[0])
    T = len(tasks)
    nterm = np.zeros((T, 2), np.int32)
    sl = np.zeros((T, 2, K), np.int32)
    cf = np.zeros((T, 2, K), np.float64)
    gstart, gslice = [0], []
    fg_same = True
    nsl = int(K) * int(Ne)
    for t, (out, fterms, gterms) in enumerate(tasks):
        if not 0 <= out < nsl:
            raise ValueError('output slice %d outside [0, K*Ne)' % out)
        for w, terms in enumerate((fterms, gterms)):
            if len(terms) > K:
                raise ValueError('more than K terms in a task')
            if any(not 0 <= s < nsl for s, _ in terms):
                raise ValueError('input slice outside [0, K*Ne)')
            nterm[t, w] = len(terms)
            for a, (s, c) in enumerate(terms):
                sl[t, w, a] = s
                cf[t, w, a] = c
        if sorted(fterms) != sorted(gterms):
            fg_same = False
        if t > 0 and out != tasks[t - 1][0]:
            gstart.append(t)
            gslice.append(tasks[t - 1][0])
    if T:
        gstart.append(T)
        gslice.append(tasks[-1][0])
    return dict(
        K=int(K), Ne=int(Ne), ntask=T, ngroup=len(gslice), fg_same=int(fg_same),
        nterm=torch.tensor(nterm, dtype=torch.int32, device=device),
        slice=torch.tensor(sl, dtype=torch.int32, device=device),
        coef=torch.tensor(cf, dtype=tdtype, device=device),
        gstart=torch.tensor(np.array(gstart, np.int32), device=device),
        gslice=torch.tensor(np.array(gslice, np.int32), device=device))


class CollisionPlan:
    def __init__(self, N, M, Nrho, L, qz, qw, sz, sw, gamma, prefac, dtype, variant,
                 ntask_max, device='cuda'):
        import torch
        self.torch = torch
        self.N, self.M, self.Nrho = int(N), int(M), int(Nrho)
        self.n = self.N ** 3
        self.dtype = dtype
        self.tdtype = torch.float64 if dtype == _lib.F64 else torch.float32
        self.variant = variant
        self.prefac = float(prefac)
        self.device = device
        L_ = _lib.lib()
        qz_a, qz_p = _lib.dbl_array(qz)
        qw_a, qw_p = _lib.dbl_array(qw)
        sz_a, sz_p = _lib.dbl_array(np.asarray(sz).reshape(-1, 3))
        if sz_a.shape[0] != self.M or len(qz_a) != self.Nrho or len(qw_a) != self.Nrho:
            raise ValueError('inconsistent quadrature table sizes')
        h = _lib.C.c_void_p()
        _lib.check(L_.dgfs_collision_plan_create(
            self.N, self.M, self.Nrho, float(L), qz_p, qw_p, sz_p, float(sw), float(gamma),
            self.prefac, dtype, variant, int(ntask_max), _lib.C.byref(h)))
        self._h = h
        self.ntask_max = int(ntask_max)
        self.tasks = None

    def query(self):
        s, qg, pg = _lib.C.c_int(), _lib.C.c_int(), _lib.C.c_int()
        b = _lib.C.c_size_t()
        _lib.check(_lib.lib().dgfs_collision_plan_query(
            self._h, _lib.C.byref(s), _lib.C.byref(qg), _lib.C.byref(pg), _lib.C.byref(b)))
        return dict(split=s.value, qgroups=qg.value, pgroups=pg.value, scratch_bytes=b.value)

    def set_qgroups(self, qg):
        """Split the M sphere directions of the gain kernel into qg groups."""
        _lib.check(_lib.lib().dgfs_collision_plan_set_qgroups(self._h, int(qg)))

    def set_schedule(self, split_mult=1, xcd_remap=1):
        """The f = g gain kernel's residue split (1: smallest that fits LDS,
        2: twice that) and work-group order (1: XCD-aware, 0: dispatch order)."""
        _lib.check(_lib.lib().dgfs_collision_plan_set_schedule(self._h, int(split_mult),
                                                               int(xcd_remap)))
        self.schedule = (int(split_mult), int(xcd_remap))

    def set_tasks(self, K, Ne, tasks):
        """tasks: list of (out_slice, f_terms, g_terms), terms = [(slice, coef), ...].
        Output slices are u-layout slice indices mode*Ne + elem."""
        if len(tasks) > self.ntask_max:
            raise ValueError('too many collision tasks (%d > %d)' % (len(tasks), self.ntask_max))
        self.tasks = task_table(K, Ne, tasks, self.tdtype, self.device)
        return self.tasks

    def __call__(self, u, Q, stream=None, ug=None, tasks=None):
        """Q += collision(u) for every task (stream-ordered, no host sync).
        g-terms read ``ug`` when given (the separate d_g array of fs()).
        ``tasks`` (a task_table) overrides the table of set_tasks()."""
        t = self.tasks if tasks is None else tasks
        if t is None:
            raise RuntimeError('set_tasks() must be called first')
        if t['ntask'] > self.ntask_max:
            raise ValueError('too many collision tasks (%d > %d)' % (t['ntask'], self.ntask_max))
        for x in (u, Q) + ((ug,) if ug is not None else ()):
            if x.dtype != self.tdtype:
                raise TypeError('collision operands must be %s' % self.tdtype)
        n = t['K'] * t['Ne'] * self.n
        if u.numel() < n or Q.numel() < n or (ug is not None and ug.numel() < n):
            raise ValueError('collision operands smaller than K*Ne*Nv')
        fg_same = t['fg_same'] if ug is None else 0
        _lib.check(_lib.lib().dgfs_collide(
            self._h, _lib.ptr(u), _lib.ptr(ug), _lib.ptr(Q), t['K'], t['Ne'], t['ntask'],
            _lib.ptr(t['nterm']), _lib.ptr(t['slice']), _lib.ptr(t['coef']), t['ngroup'],
            _lib.ptr(t['gstart']), _lib.ptr(t['gslice']), fg_same,
            _lib.stream_ptr(stream)))

    def reference_set(self, u0, stream=None, tasks=None):
        """Make u0 the plan's reference state (dgfs_collision_reference_set):
        its F0^ and inverse transforms for the task table, used by every
        reference() call until the next reference_set (u0 must keep its
        storage and contents until then)."""
        t = self.tasks if tasks is None else tasks
        if t is None:
            raise RuntimeError('set_tasks() must be called first')
        if not t['fg_same']:
            raise ValueError('the reference-state form evaluates Q(f, f): every task needs f = g')
        if u0.dtype != self.tdtype:
            raise TypeError('collision operands must be %s' % self.tdtype)
        if u0.numel() < t['K'] * t['Ne'] * self.n:
            raise ValueError('collision operands smaller than K*Ne*Nv')
        _lib.check(_lib.lib().dgfs_collision_reference_set(
            self._h, _lib.ptr(u0), t['K'], t['Ne'], t['ntask'], _lib.ptr(t['nterm']),
            _lib.ptr(t['slice']), _lib.ptr(t['coef']), _lib.stream_ptr(stream)))

    def reference(self, u, u0, Q0, Q, stream=None, tasks=None):
        """Q += Q(f, f) in the reference-state form (dgfs_collide_reference):
        Q0 (FP64, (K, Ne, Nv)) must hold Q(f0, f0) of the reference state u0,
        which reference_set(u0) must have made the plan's reference state;
        the FP32 rounding error then scales with |f - f0|.  f = g tasks only."""
        import torch
        t = self.tasks if tasks is None else tasks
        if t is None:
            raise RuntimeError('set_tasks() must be called first')
        if not t['fg_same']:
            raise ValueError('the reference-state form evaluates Q(f, f): every task needs f = g')
        for x in (u, u0, Q):
            if x.dtype != self.tdtype:
                raise TypeError('collision operands must be %s' % self.tdtype)
        if Q0.dtype != torch.float64:
            raise TypeError('Q0 must be float64')
        n = t['K'] * t['Ne'] * self.n
        if min(u.numel(), u0.numel(), Q0.numel(), Q.numel()) < n:
            raise ValueError('collision operands smaller than K*Ne*Nv')
        _lib.check(_lib.lib().dgfs_collide_reference(
            self._h, _lib.ptr(u), _lib.ptr(u0), _lib.ptr(Q0), _lib.ptr(Q), t['K'], t['Ne'],
            t['ntask'], _lib.ptr(t['nterm']), _lib.ptr(t['slice']), _lib.ptr(t['coef']),
            t['ngroup'], _lib.ptr(t['gstart']), _lib.ptr(t['gslice']), _lib.stream_ptr(stream)))

    def timing(self, max_calls):
        """Record per-stage HIP events for the next max_calls calls (0 = off)."""
        _lib.check(_lib.lib().dgfs_collision_timing(self._h, int(max_calls)))

    STAGES = ('forward', 'nyquist', 'gain', 'plane', 'line')

    def timing_read(self):
        """(ms per stage [forward, nyquist, gain, plane, line] summed, ncalls)."""
        ms = (_lib.C.c_double * 5)()
        nc = _lib.C.c_int()
        _lib.check(_lib.lib().dgfs_collision_timing_read(
            self._h, _lib.C.cast(ms, _lib.C.POINTER(_lib.C.c_double)), _lib.C.byref(nc)))
        return list(ms), nc.value

    def close(self):
        if getattr(self, '_h', None) is not None and self._h.value:
            _lib.lib().dgfs_collision_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BiCollisionPlan:
    """One species pair (p, q) of the binary-mixture VSS operator
    (reference DGFSBiVSSScatteringModel, bi/scattering.py:47-321) on a
    ``dgfs_bi_collision_plan_t``: the complex kernels b_pq, c_pq are built on
    the device at creation.  ``__call__(up, uq, Q)`` adds Q_pq(f_p, f_q) of
    every task to Q; the task table is that of CollisionPlan (f terms
    contract up, g terms contract uq)."""

    def __init__(self, N, M, Nrho, L, qz, qw, sz, sw, szpre, mp, mq, alpha, omega, prefac,
                 dtype, ntask_max, device='cuda'):
        import torch
        self.N, self.M, self.Nrho = int(N), int(M), int(Nrho)
        self.n = self.N ** 3
        self.dtype = dtype
        self.tdtype = torch.float64 if dtype == _lib.F64 else torch.float32
        self.prefac = float(prefac)
        self.device = device
        qz_a, qz_p = _lib.dbl_array(qz)
        qw_a, qw_p = _lib.dbl_array(qw)
        sz_a, sz_p = _lib.dbl_array(np.asarray(sz).reshape(-1, 3))
        sp_a, sp_p = _lib.dbl_array(np.asarray(szpre).reshape(-1, 3))
        if sz_a.shape[0] != self.M or len(qz_a) != self.Nrho or len(qw_a) != self.Nrho:
            raise ValueError('inconsistent quadrature table sizes')
        h = _lib.C.c_void_p()
        _lib.check(_lib.lib().dgfs_bi_collision_plan_create(
            self.N, self.M, self.Nrho, float(L), qz_p, qw_p, sz_p, float(sw), sp_a.shape[0], sp_p,
            float(mp), float(mq), float(alpha), float(omega), self.prefac, dtype, int(ntask_max),
            _lib.C.byref(h)))
        self._h = h
        self.ntask_max = int(ntask_max)
        self.tasks = None

    def query(self):
        s, ch = _lib.C.c_int(), _lib.C.c_int()
        b = _lib.C.c_size_t()
        _lib.check(_lib.lib().dgfs_bi_collision_plan_query(
            self._h, _lib.C.byref(s), _lib.C.byref(ch), _lib.C.byref(b)))
        return dict(split=s.value, chunk=ch.value, scratch_bytes=b.value)

    def set_tasks(self, K, Ne, tasks):
        if len(tasks) > self.ntask_max:
            raise ValueError('too many collision tasks (%d > %d)' % (len(tasks), self.ntask_max))
        self.tasks = task_table(K, Ne, tasks, self.tdtype, self.device)
        return self.tasks

    def __call__(self, up, uq, Q, stream=None, tasks=None):
        """Q += Q_pq(up, uq) for every task (stream-ordered, no host sync)."""
        t = self.tasks if tasks is None else tasks
        if t is None:
            raise RuntimeError('set_tasks() must be called first')
        if t['ntask'] > self.ntask_max:
            raise ValueError('too many collision tasks (%d > %d)' % (t['ntask'], self.ntask_max))
        for x in (up, uq, Q):
            if x.dtype != self.tdtype:
                raise TypeError('collision operands must be %s' % self.tdtype)
        n = t['K'] * t['Ne'] * self.n
        if min(up.numel(), uq.numel(), Q.numel()) < n:
            raise ValueError('collision operands smaller than K*Ne*Nv')
        _lib.check(_lib.lib().dgfs_bi_collide(
            self._h, _lib.ptr(up), _lib.ptr(uq), _lib.ptr(Q), t['K'], t['Ne'], t['ntask'],
            _lib.ptr(t['nterm']), _lib.ptr(t['slice']), _lib.ptr(t['coef']), t['ngroup'],
            _lib.ptr(t['gstart']), _lib.ptr(t['gslice']), _lib.stream_ptr(stream)))

    def close(self):
        if getattr(self, '_h', None) is not None and self._h.value:
            _lib.lib().dgfs_bi_collision_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
