"""dgfsStd1D: single-species DGFS in one physical dimension on one MI355X.

Same INI schema, plugin registry keys, RHS composition and time loop as the
reference driver dgfs1D/std/std.py:37-423:

    rhs(t, u_in, u_out):  u_out = invMass( transport(u_in) + sum Q(U u_in, V u_in) )

with the transport (face reconstruction, boundaries, upwind flux, -c D u,
lift, 1/J) fused in one kernel and every collision of the RHS evaluated by one
batched launch sequence.  The RHS never synchronises the host (the reference
copies the boundary faces to the host every RHS, std.py:295).

Usage:  python -m dgfs1d_gpu_amd.std.std run dgfs1D.ini [-v sect::key value]...
"""
import logging
import sys
import time as _time

import numpy as np

from ..basis import Basis
from ..initialize import initialize
from ..mesh import Mesh
from ..nputil import subclass_where
from ..transport import MatOp, Transport
from .bc import DGFSBCStd
from .distribution import DGFSDistributionStd, load_distribution
from .initcond import DGFSInitConditionStd
from .integrator import DGFSIntegratorStd
from .moments import DGFSMomWriterStd
from .residual import DGFSResidualStd
from .scattering import DGFSScatteringModelStd
from .velocitymesh import DGFSVelocityMeshStd

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSStd1D:
    """Solver state and operators; ``rhs``, ``step`` and ``run`` drive it."""

    def __init__(self, cfg, device='cuda', restart=None, writers=True):
        import torch
        self.torch = torch
        self.cfg = cfg
        self.device = torch.device(device)
        self.tdtype = torch.float64 if cfg.dtype == np.float64 else torch.float32

        self.mesh = mesh = Mesh(cfg)
        self.Ne = Ne = mesh.Ne
        bkind = cfg.lookup('basis', 'kind')
        self.basis = basis = subclass_where(Basis, basis_kind=bkind)(cfg)
        self.K = K = basis.K
        self.Nq = Nq = basis.Nq
        self.vm = vm = DGFSVelocityMeshStd(cfg, device=self.device)
        self.Nv = Nv = vm.vsize()

        tasks = basis.collision_tasks(Ne)
        smn = cfg.lookup('scattering-model', 'type')
        smcls = subclass_where(DGFSScatteringModelStd, scattering_model=smn)
        self.sm = smcls(cfg, vm, Ne=Ne, device=self.device, ntask_max=max(1, len(tasks)))
        self.sm.set_tasks(K, Ne, tasks)
        self.ntasks = len(tasks)

        ti, dt, tf = (float(x) for x in cfg.lookupfloats('time-integrator',
                                                         ('tstart', 'dt', 'tend')))
        nsteps = np.ceil((tf - ti) / dt)
        self.dt = (tf - ti) / nsteps                       # std.py:85-87
        self.tstart, self.tend = ti, tf

        z, jac = basis.z, mesh.jac
        xm = mesh.xmesh
        self.xsol = np.array([0.5 * (xm[j] + xm[j + 1]) + jac[j] * z for j in range(Ne)]).T
        self.xcoeff = np.einsum('kq,qe->ke', basis.fwdTransMat, self.xsol)

        self.bcl = subclass_where(DGFSBCStd, type=cfg.lookup('soln-bcs-xlo', 'type'))(
            xm[0], -1.0, vm, cfg, 'soln-bcs-xlo', device=self.device)
        self.bcr = subclass_where(DGFSBCStd, type=cfg.lookup('soln-bcs-xhi', 'type'))(
            xm[-1], 1.0, vm, cfg, 'soln-bcs-xhi', device=self.device)
        self.transport = Transport(basis, mesh.invjac[:, 0], vm, self.bcl, self.bcr,
                                   self.tdtype, self.device)
        self.invmass = MatOp(basis.mat_invmass, self.device)
        self.fwd = MatOp(basis.mat_fwd, self.device)
        self.bwd = MatOp(basis.mat_bwd, self.device)

        n = K * Ne * Nv
        self.u = torch.zeros(n, dtype=self.tdtype, device=self.device)
        self.ux = torch.zeros(n, dtype=self.tdtype, device=self.device)
        if restart is None:
            usol = np.empty((Nq, Ne, Nv))
            icn = cfg.lookup('soln-ics', 'type')
            ic = subclass_where(DGFSInitConditionStd, model=icn)(cfg, vm, 'soln-ics')
            ic.apply_init_vals(usol, Nq, Ne, self.xsol)
            d_usol = torch.tensor(usol.ravel(), dtype=self.tdtype, device=self.device)
            self.fwd(d_usol, self.u)
        else:
            ti, coeff = load_distribution(restart, K, Ne, Nv)
            self.u.copy_(torch.tensor(coeff, dtype=self.tdtype))
        self.time = ti
        self.uprev = self.u.clone()
        self.nacptsteps = 0

        self.odes = self._make_integrator(cfg, (K, Ne, Nv))

        self.residual = None
        self.moments = None
        self.distribution = None
        if writers:
            if cfg.has_section('dgfsresidual'):
                self.residual = DGFSResidualStd(cfg, 'dgfsresidual', device=self.device)
            if cfg.has_section('dgfsmomwriter'):
                self.moments = DGFSMomWriterStd(ti, basis.interpMat, self.xcoeff, self.u, vm,
                                                cfg, 'dgfsmomwriter', device=self.device)
            if cfg.has_section('dgfsdistwriter'):
                self.distribution = DGFSDistributionStd(ti, (K, Ne, Nv), cfg, 'dgfsdistwriter')

    # ------------------------------------------------------------------ RHS
    def rhs(self, t, u_in, u_out):
        """u_out = invMass(transport(u_in) + collisions(u_in))  (std.py:281-366)."""
        if self.invmass.is_identity:
            self.transport(t, u_in, u_out)
            self.sm.collide(u_in, u_out)
        else:
            self.transport(t, u_in, self.ux)
            self.sm.collide(u_in, self.ux)
            self.invmass(self.ux, u_out)

    def _make_integrator(self, cfg, shape):
        odescls = subclass_where(DGFSIntegratorStd,
                                 intg_kind=cfg.lookup('time-integrator', 'scheme'))
        comp = cfg.lookupordefault('time-integrator', 'compensated', 'false').lower()
        if comp not in ('true', 'false', '1', '0', 'yes', 'no'):
            raise ValueError('time-integrator::compensated must be true or false')
        return odescls(self.rhs, shape, self.tdtype, device=self.device,
                       compensated=comp in ('true', '1', 'yes'))

    def _advance(self):
        self.odes.integrate(self.time, self.dt, self.u)

    def state_fp64(self):
        """The state in FP64: u, plus the low word of a compensated FP32 state."""
        lo = self.odes.lo_of(self.u) if hasattr(self.odes, 'lo_of') else None
        return self.u.double() if lo is None else self.u.double() + lo.double()

    # ----------------------------------------------------------- time loop
    def step(self):
        self._advance()
        self.time += self.dt
        self.nacptsteps += 1
        if self.residual is not None:
            self.residual(self.time, self.nacptsteps, self.uprev, self.u)
        if self.moments is not None:
            self.moments(self.dt, self.time, self.u)
        if self.distribution is not None:
            self.distribution(self.dt, self.time, self.u)
        if self.residual is not None:       # the previous state serves the residual only
            self.uprev.copy_(self.u)

    def run(self, nsteps=None):
        """nsteps=None: the reference loop ``while time < tend`` (std.py:396-404,
        which may take one step more than (tend-tstart)/dt); else nsteps steps."""
        start = _time.perf_counter()
        if nsteps is None:
            while self.time < self.tend:
                self.step()
        else:
            for _ in range(int(nsteps)):
                self.step()
        self.torch.cuda.synchronize()
        elapsed = _time.perf_counter() - start
        log.info('Nsteps %d elapsed time %s s', self.nacptsteps, elapsed)
        return elapsed


def main(argv=None):
    logging.basicConfig(level=logging.INFO, format='%(asctime)s %(message)s')
    cfg, args = initialize(argv)
    restart = args.dist[0][0].name if hasattr(args, 'process_restart') else None
    solver = DGFSStd1D(cfg, restart=restart)
    solver.run()


if __name__ == '__main__':
    sys.exit(main())
