"""Boundary conditions of the binary mixture (reference dgfs1D/bi/bc.py:9-249
with bi/kernels/bcs/*.mako).  As in the single-species solver the update
(diffuse-wall number density) and application are fused into the transport
kernel; ``wall_params(p, t)`` returns species p's dgfs_wall_t fields."""
import numpy as np

from .. import _lib
from ..nputil import npeval
from .initcond import DGFSMaxwellianInitConditionBi


def _tensor(x, cfg, device):
    import torch
    td = torch.float64 if cfg.dtype == np.float64 else torch.float32
    return torch.tensor(np.ascontiguousarray(x), dtype=td, device=device)


class DGFSBCBi:
    type = None

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        self._vm, self._nl, self.cfg, self.sect = vm, float(nl), cfg, cfgsect
        self.device = device

    def wall_params(self, p, t):
        return dict(kind=_lib.BC_COPY, nl=self._nl)

    def species(self, p):
        """The single-species boundary view that the transport operator reads."""
        return _SpeciesView(self, p)


class _SpeciesView:
    def __init__(self, bc, p):
        self.bc, self.p = bc, p

    def wall_params(self, t):
        return self.bc.wall_params(self.p, t)


class DGFSPeriodicBCBi(DGFSBCBi):
    type = 'dgfs-periodic'


class DGFSWallDiffuseBCBi(DGFSBCBi):
    """bc.py:25-90, dgfs-wall-diffuse.mako: ghost = fac < 0 ? n_w f0_p : interior
    with f0_p the unit-density Maxwellian of species p at the wall state."""
    type = 'dgfs-wall-diffuse'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        ic = DGFSMaxwellianInitConditionBi(cfg, vm, cfgsect, wall=True)
        self.f0_host = ic.get_init_vals()
        self.f0 = [_tensor(f, cfg, device) for f in self.f0_host]
        self.ushift = float(ic.unondim()[0, 0])

    def wall_params(self, p, t):
        return dict(kind=_lib.BC_DIFFUSE, nl=self._nl, ushift=self.ushift, f0=self.f0[p])


class DGFSWallExprDiffuseBCBi(DGFSBCBi):
    """bc.py:122-206, dgfs-wall-expr-diffuse.mako: the wall Maxwellian
    exp(-m_p |c - u|^2 / T) evaluated in the kernel; passed to the transport
    kernel's unnormalised exp(-|c - u|^2 / T') as T' = T / m_p."""
    type = 'dgfs-wall-expr-diffuse'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        self.exprs = {k: cfg.lookupexpr(cfgsect, k) for k in ('ux', 'uy', 'uz', 'T')}
        self.x = float(xsol)

    def wall_params(self, p, t):
        ev = {k: float(npeval(e, {'t': t, 'x': self.x})) for k, e in self.exprs.items()}
        u0, T0 = self._vm.u0(), self._vm.T0()
        ux = ev['ux'] / u0
        return dict(kind=_lib.BC_DIFFUSE_EXPR, nl=self._nl, ushift=ux, ux=ux,
                    uy=ev['uy'] / u0, uz=ev['uz'] / u0,
                    T=ev['T'] / T0 / self._vm.masses()[p])


class DGFSInletBCBi(DGFSBCBi):
    """bc.py:208-249, dgfs-inlet.mako: ghost = fac < 0 ? f0_p : interior with
    the inlet Maxwellian of section nden1, nden2, ux, uy, uz, T."""
    type = 'dgfs-inlet'

    def __init__(self, xsol, nl, vm, cfg, cfgsect, device=None, **kwargs):
        super().__init__(xsol, nl, vm, cfg, cfgsect, device)
        ic = DGFSMaxwellianInitConditionBi(cfg, vm, cfgsect, wall=False)
        self.f0 = [_tensor(f, cfg, device) for f in ic.get_init_vals()]

    def wall_params(self, p, t):
        # the inlet kernel takes fac = c_x nl with no velocity shift (dgfs-inlet.mako:14)
        return dict(kind=_lib.BC_INLET, nl=self._nl, ushift=0.0, f0=self.f0[p])
