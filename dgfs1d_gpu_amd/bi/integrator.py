"""Explicit integrators of the binary mixture (reference
dgfs1D/bi/integrator.py:12-206).  ``integrate(t, dt, us)`` advances both
species' coefficient arrays in place; ``rhs(p, t, us_in, out)`` evaluates
species p's RHS from both species' states.  Within a stage every species'
RHS reads the same stage states, as in the reference.

ssp-rk3: the reference's second stage forms r1 = 0.25 dt L(r0) + 0.75 u +
dt r0 (integrator.py:121-124), which is not a Runge-Kutta stage for dt != 0.25;
this implements the scheme its comment and the single-species solver state,
r1 = 0.75 u + 0.25 r0 + 0.25 dt L(r0).
"""
from ..transport import axnpby


class DGFSIntegratorBi:
    intg_kind = None
    _nscratch_per_species = 0

    def __init__(self, rhs, shape, dtype, nspcs, device=None, **kwargs):
        import torch
        self._rhs = rhs
        self._nspcs = nspcs
        K, Ne, Nv = shape
        self.scratch = [[torch.zeros(K * Ne * Nv, dtype=dtype, device=device)
                         for _ in range(nspcs)] for _ in range(self._nscratch_per_species)]

    @staticmethod
    def copy(olds, news):
        for o, n in zip(olds, news):
            n.copy_(o)

    def integrate(self, t, dt, us):
        raise NotImplementedError


class EulerIntegratorBi(DGFSIntegratorBi):
    """integrator.py:48-73"""
    intg_kind = 'euler'
    _nscratch_per_species = 1

    def integrate(self, t, dt, us):
        r0, = self.scratch
        for p in range(self._nspcs):
            self._rhs(p, t, us, r0[p])
            axnpby(r0[p], [(1.0, us[p])], dt)              # r0 = u + dt L
        self.copy(r0, us)


class SSPRK2IntegratorBi(DGFSIntegratorBi):
    """integrator.py:75-102"""
    intg_kind = 'ssp-rk2'
    _nscratch_per_species = 2

    def integrate(self, t, dt, us):
        r0, r1 = self.scratch
        for p in range(self._nspcs):
            self._rhs(p, t, us, r0[p])
            axnpby(r0[p], [(1.0, us[p])], dt)
        for p in range(self._nspcs):
            self._rhs(p, t + dt, r0, r1[p])
            axnpby(us[p], [(0.5, r0[p]), (0.5 * dt, r1[p])], 0.5)


class SSPRK3IntegratorBi(DGFSIntegratorBi):
    """integrator.py:104-138 (second stage as the scheme states; see the
    module docstring)"""
    intg_kind = 'ssp-rk3'
    _nscratch_per_species = 3

    def integrate(self, t, dt, us):
        r0, r1, r2 = self.scratch
        for p in range(self._nspcs):
            self._rhs(p, t, us, r1[p])
            axnpby(r0[p], [(1.0, us[p]), (dt, r1[p])], 0.0)
        for p in range(self._nspcs):
            self._rhs(p, t + dt, r0, r2[p])
            axnpby(r1[p], [(0.75, us[p]), (0.25, r0[p]), (0.25 * dt, r2[p])], 0.0)
        for p in range(self._nspcs):
            self._rhs(p, t + 0.5 * dt, r1, r2[p])
        for p in range(self._nspcs):
            axnpby(us[p], [(2.0 / 3.0, r1[p]), (2.0 / 3.0 * dt, r2[p])], 1.0 / 3.0)


class LSERK45IntegratorBi(DGFSIntegratorBi):
    """integrator.py:140-176: five-stage low-storage RK (Carpenter-Kennedy)."""
    intg_kind = 'lse-rk45'
    _nscratch_per_species = 2
    rk45a = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
             -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
    rk45b = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
             1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
             2277821191437.0 / 14882151754819.0]
    rk45c = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
             2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]

    def integrate(self, t, dt, us):
        r0, r1 = self.scratch
        for s in range(5):
            for p in range(self._nspcs):
                self._rhs(p, t + self.rk45c[s] * dt, us, r0[p])
                axnpby(r1[p], [(dt, r0[p])], self.rk45a[s])           # r1 = a r1 + dt L
                axnpby(r0[p], [(1.0, us[p]), (self.rk45b[s], r1[p])], 0.0)
            self.copy(r0, us)
