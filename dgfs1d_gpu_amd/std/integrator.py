"""Explicit time integrators (reference dgfs1D/std/integrator.py:7-157).
Registry key ``intg_kind``; ``integrate(t, dt, u)`` advances u in place with
rhs(t, u_in, u_out) and fused axnpby stage updates on the device.

``compensated=True`` (FP32 only) carries the state and every stage register
that holds a state as an unevaluated FP32 pair hi + lo, updated in
double-float arithmetic (dgfs_axnpby_compensated): the RHS reads hi, and the
dt*L increments are not lost to FP32 rounding of the state.  ``lo_of(u)``
returns the low word of the state (zero until the first step)."""
from ..transport import axnpby, axnpby_compensated


class DGFSIntegratorStd:
    intg_kind = None
    _nscratch = 0
    _state_regs = ()          # scratch registers that hold states (get a low word)

    def __init__(self, rhs, shape, dtype, device=None, compensated=False, **kwargs):
        import torch
        self.torch = torch
        self._rhs = rhs
        K, Ne, Nv = shape
        self.scratch = [torch.zeros(K * Ne * Nv, dtype=dtype, device=device)
                        for _ in range(self._nscratch)]
        self.compensated = bool(compensated) and dtype == torch.float32
        self._lo = {}
        if self.compensated:
            for i in self._state_regs:
                self._lo[id(self.scratch[i])] = torch.zeros_like(self.scratch[i])

    def lo_of(self, u):
        """Low word of a compensated state (None when not compensated)."""
        if not self.compensated:
            return None
        if id(u) not in self._lo:
            self._lo[id(u)] = self.torch.zeros_like(u)
        return self._lo[id(u)]

    def reset(self, u):
        """Forget the low word of u (after u is overwritten from outside)."""
        if id(u) in self._lo:
            self._lo[id(u)].zero_()

    def _ax(self, x0, terms, a0):
        if not self.compensated:
            axnpby(x0, terms, a0)
            return
        axnpby_compensated(x0, self._lo.get(id(x0)),
                           [(a, x, self._lo.get(id(x))) for a, x in terms], a0)

    def integrate(self, t, dt, u):
        raise NotImplementedError


class EulerIntegratorStd(DGFSIntegratorStd):
    intg_kind = 'euler'
    _nscratch = 1

    def integrate(self, t, dt, u):
        r0, = self.scratch
        self.lo_of(u)
        self._rhs(t, u, r0)
        self._ax(u, [(dt, r0)], 1.0)


class SSPRK2IntegratorStd(DGFSIntegratorStd):
    intg_kind = 'ssp-rk2'
    _nscratch = 2
    _state_regs = (0,)

    def integrate(self, t, dt, u):
        r0, r1 = self.scratch
        self.lo_of(u)
        self._rhs(t, u, r0)
        if self.compensated:                # r0 = u + dt L(u) as a state; r0's old low word unused
            axnpby_compensated(r0, self._lo[id(r0)], [(1.0, u, self._lo[id(u)]), (dt, r0, None)],
                               0.0)
        else:
            axnpby(r0, [(1.0, u)], dt)
        self._rhs(t + dt, r0, r1)
        self._ax(u, [(0.5, r0), (0.5 * dt, r1)], 0.5)      # u = u/2 + r0/2 + dt/2 L(r0)


class SSPRK3IntegratorStd(DGFSIntegratorStd):
    intg_kind = 'ssp-rk3'
    _nscratch = 2
    _state_regs = (0,)

    def integrate(self, t, dt, u):
        r0, r1 = self.scratch
        self.lo_of(u)
        self._rhs(t, u, r1)
        self._ax(r0, [(1.0, u), (dt, r1)], 0.0)
        self._rhs(t + dt, r0, r1)
        self._ax(r0, [(0.75, u), (0.25 * dt, r1)], 0.25)
        self._rhs(t + 0.5 * dt, r0, r1)
        self._ax(u, [(2.0 / 3.0, r0), (2.0 / 3.0 * dt, r1)], 1.0 / 3.0)


class LSERK45IntegratorStd(DGFSIntegratorStd):
    """Carpenter-Kennedy 4th order, 5-stage low-storage RK (same tables as
    integrator.py:115-137)."""
    intg_kind = 'lse-rk45'
    _nscratch = 2
    rk45a = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
             -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
    rk45b = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
             1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
             2277821191437.0 / 14882151754819.0]
    rk45c = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
             2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]

    def integrate(self, t, dt, u):
        r0, r1 = self.scratch
        self.lo_of(u)
        for i in range(5):
            self._rhs(t + self.rk45c[i] * dt, u, r0)
            axnpby(r1, [(dt, r0)], self.rk45a[i])          # increment register: no low word
            self._ax(u, [(self.rk45b[i], r1)], 1.0)


class LESRK45IntegratorStd(LSERK45IntegratorStd):
    """README spelling of the same scheme (README.md:81)."""
    intg_kind = 'lesrk-45'
