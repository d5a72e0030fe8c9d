"""Boundary conditions (reference dgfs1D/std/bc.py:9-324).

A boundary object only carries the parameters of its kind; the update
(diffuse-wall number density) and application are fused into the transport
kernel (dgfs_transport_rhs).  ``wall_params(t)`` returns the dgfs_wall_t
fields; time-dependent expressions are evaluated on the host per stage."""
import numpy as np

from .. import _lib
from ..nputil import npeval
from .initcond import DGFSMaxwellianInitConditionStd, maxwellian


def _tensor(x, cfg, device):
    import torch
    td = torch.float64 if cfg.dtype == np.float64 else torch.float32
    return torch.tensor(np.ascontiguousarray(x), dtype=td, device=device)


class DGFSBCStd:
    type = None

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        self._vm, self._nl, self.cfg, self.sect = vm, float(nl), cfg, cfgsect
        self.device = device

    def wall_params(self, t):
        return dict(kind=_lib.BC_COPY, nl=self._nl)


class DGFSPeriodicBCStd(DGFSBCStd):
    type = 'dgfs-periodic'


class DGFSCyclicBCStd(DGFSBCStd):
    type = 'dgfs-cyclic'


class DGFSWallDiffuseBCStd(DGFSBCStd):
    type = 'dgfs-wall-diffuse'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        ic = DGFSMaxwellianInitConditionStd(cfg, vm, cfgsect, wall=True)
        self.f0_host = ic.get_init_vals()
        self.f0 = _tensor(self.f0_host, cfg, device)
        self.ushift = float(ic.unondim()[0, 0])

    def wall_params(self, t):
        return dict(kind=_lib.BC_DIFFUSE, nl=self._nl, ushift=self.ushift, f0=self.f0)


class DGFSWallExprDiffuseBCStd(DGFSBCStd):
    type = 'dgfs-wall-expr-diffuse'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        self.exprs = {k: cfg.lookupexpr(cfgsect, k) for k in ('ux', 'uy', 'uz', 'T')}
        self.x = float(xsol)

    def wall_params(self, t):
        ev = {k: float(npeval(e, {'t': t, 'x': self.x})) for k, e in self.exprs.items()}
        u0, T0 = self._vm.u0(), self._vm.T0()
        ux = ev['ux'] / u0
        return dict(kind=_lib.BC_DIFFUSE_EXPR, nl=self._nl, ushift=ux, ux=ux,
                    uy=ev['uy'] / u0, uz=ev['uz'] / u0, T=ev['T'] / T0)


class DGFSInletBCStd(DGFSBCStd):
    type = 'dgfs-inlet'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        ic = DGFSMaxwellianInitConditionStd(cfg, vm, cfgsect, wall=False)
        self.f0 = _tensor(ic.get_init_vals(), cfg, device)

    def wall_params(self, t):
        return dict(kind=_lib.BC_INLET, nl=self._nl, ushift=0.0, f0=self.f0)


class DGFSInletNonDimBCStd(DGFSBCStd):
    type = 'dgfs-inlet-nondim'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        rho, ux, T = (float(cfg.lookupfloat(cfgsect, k)) for k in ('rho', 'ux', 'T'))
        f0 = maxwellian(vm.cv(), rho, ux, 0.0, 0.0, T)
        if not np.allclose(rho, np.sum(f0) * vm.cw(), atol=1e-5):
            raise ValueError('Bulk properties not conserved! Check Nv, dev')
        self.f0 = _tensor(f0, cfg, device)
        self.ushift = ux

    def wall_params(self, t):
        return dict(kind=_lib.BC_INLET, nl=self._nl, ushift=self.ushift, f0=self.f0)


class DGFSVanishBCStd(DGFSBCStd):
    type = 'dgfs-vanish'

    def wall_params(self, t):
        return dict(kind=_lib.BC_VANISH, nl=self._nl, ushift=0.0)
