"""dgfsBi1D: binary-mixture DGFS in one physical dimension on one MI355X.

Same INI schema, plugin registry keys, RHS composition and time loop as the
reference driver dgfs1D/bi/bi.py:31-421:

    rhs(p, t, us_in, u_out):
        u_out = invMass( transport_p(us_in[p]) + sum_q Q_pq(U us_in[p], V us_in[q]) )

with species p's transport (its own wall Maxwellians) fused in one kernel and
every collision of a species pair evaluated by one batched launch sequence.

Usage:  python -m dgfs1d_gpu_amd.bi.bi run dgfs1D.ini [-v sect::key value]...
"""
import logging
import sys
import time as _time

import numpy as np

from ..basis import Basis
from ..initialize import initialize
from ..mesh import Mesh
from ..nputil import subclass_where
from ..relax import MixtureProjection
from ..transport import MatOp, Transport, axnpby
from .bc import DGFSBCBi
from .distribution import DGFSDistributionBi, load_distribution
from .initcond import DGFSInitConditionBi
from .integrator import DGFSIntegratorBi
from .moments import DGFSMomWriterBi
from .residual import DGFSResidualBi
from .scattering import DGFSScatteringModelBi
from .velocitymesh import DGFSVelocityMeshBi

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSBi1D:
    def __init__(self, cfg, device='cuda', restart=None, writers=True):
        import torch
        self.torch = torch
        self.cfg = cfg
        self.device = torch.device(device)
        self.tdtype = torch.float64 if cfg.dtype == np.float64 else torch.float32

        self.mesh = mesh = Mesh(cfg)
        self.Ne = Ne = mesh.Ne
        self.basis = basis = subclass_where(Basis, basis_kind=cfg.lookup('basis', 'kind'))(cfg)
        self.K = K = basis.K
        self.Nq = Nq = basis.Nq
        self.vm = vm = DGFSVelocityMeshBi(cfg, device=self.device)
        self.Nv = Nv = vm.vsize()
        self.nspcs = ns = vm.nspcs()

        tasks = basis.collision_tasks(Ne)
        smcls = subclass_where(DGFSScatteringModelBi,
                               scattering_model=cfg.lookup('scattering-model', 'type'))
        self.sm = smcls(cfg, vm, Ne=Ne, device=self.device, ntask_max=max(1, len(tasks)))
        self.sm.set_tasks(K, Ne, tasks)

        ti, dt, tf = (float(x) for x in cfg.lookupfloats('time-integrator',
                                                         ('tstart', 'dt', 'tend')))
        nsteps = np.ceil((tf - ti) / dt)
        self.dt = (tf - ti) / nsteps                       # bi.py:71-73
        self.tstart, self.tend = ti, tf

        z, jac = basis.z, mesh.jac
        xm = mesh.xmesh
        self.xsol = np.array([0.5 * (xm[j] + xm[j + 1]) + jac[j] * z for j in range(Ne)]).T
        self.xcoeff = np.einsum('kq,qe->ke', basis.fwdTransMat, self.xsol)

        self.bcl = subclass_where(DGFSBCBi, type=cfg.lookup('soln-bcs-xlo', 'type'))(
            xm[0], -1.0, vm, cfg, 'soln-bcs-xlo', device=self.device)
        self.bcr = subclass_where(DGFSBCBi, type=cfg.lookup('soln-bcs-xhi', 'type'))(
            xm[-1], 1.0, vm, cfg, 'soln-bcs-xhi', device=self.device)
        self.transport = [Transport(basis, mesh.invjac[:, 0], vm, self.bcl.species(p),
                                    self.bcr.species(p), self.tdtype, self.device)
                          for p in range(ns)]
        self.invmass = MatOp(basis.mat_invmass, self.device)
        self.fwd = MatOp(basis.mat_fwd, self.device)

        n = K * Ne * Nv
        # [scattering-model] projection = conservative: the two species'
        # collision terms of every node are projected together (each species'
        # mass, the mixture's momentum and energy), FP64 moments and solve
        mode = cfg.lookupordefault('scattering-model', 'projection', 'none')
        if mode not in ('none', 'conservative'):
            raise ValueError("scattering-model::projection must be 'none' or 'conservative'")
        self.projection = None
        if mode == 'conservative':
            self.projection = MixtureProjection(vm.cv(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(),
                                                vm.masses()[0], vm.masses()[1], self.tdtype,
                                                self.device)
            self._qc = [torch.zeros(n, dtype=self.tdtype, device=self.device)
                        for _ in range(ns)]
            self._pending = None
        # both species in one buffer (species-major); us[p] are views of it
        self.u = torch.zeros(ns * n, dtype=self.tdtype, device=self.device)
        self.us = list(self.u.view(ns, n))
        self.ux = torch.zeros(n, dtype=self.tdtype, device=self.device)
        if restart is None:
            usol = np.empty((Nq, Ne, Nv))
            ic = subclass_where(DGFSInitConditionBi, model=cfg.lookup('soln-ics', 'type'))(
                cfg, vm, 'soln-ics')
            for p in range(ns):
                ic.apply_init_vals(p, usol, Nq, Ne, self.xsol)
                d_usol = torch.tensor(usol.ravel(), dtype=self.tdtype, device=self.device)
                self.fwd(d_usol, self.us[p])
        else:
            ti, coeffs = load_distribution(restart, K, Ne, Nv, ns)
            for p in range(ns):
                self.us[p].copy_(torch.tensor(coeffs[p], dtype=self.tdtype))
        self.time = ti
        self.uprevs = [u.clone() for u in self.us]
        self.nacptsteps = 0

        self.odes = self._make_integrator(cfg, (K, Ne, Nv))

        self.residual = self.moments = self.distribution = None
        if writers:
            if cfg.has_section('dgfsresidual'):
                self.residual = DGFSResidualBi(cfg, 'dgfsresidual', device=self.device)
            if cfg.has_section('dgfsmomwriter'):
                self.moments = DGFSMomWriterBi(ti, basis.interpMat, self.xcoeff, self.us, vm,
                                               cfg, 'dgfsmomwriter')
            if cfg.has_section('dgfsdistwriter'):
                self.distribution = DGFSDistributionBi(ti, (K, Ne, Nv), cfg, 'dgfsdistwriter')

    def _make_integrator(self, cfg, shape):
        odescls = subclass_where(DGFSIntegratorBi,
                                 intg_kind=cfg.lookup('time-integrator', 'scheme'))
        return odescls(self.rhs, shape, self.tdtype, self.nspcs, device=self.device)

    def _collide(self, p, us_in, Q):
        if self.projection is None:
            self.sm.collide(p, us_in, Q)
            return
        # both species' terms of the same state are projected together; the
        # integrators evaluate species 0 then species 1 on one state, so the
        # pair computed for p = 0 serves the following p = 1 call
        if p == 0 or self._pending is not us_in:
            for q, buf in enumerate(self._qc):
                buf.zero_()
                self.sm.collide(q, us_in, buf)
            self.projection(self._qc[0], self._qc[1], self._qc[0], self._qc[1],
                            self.K * self.Ne, accumulate=False)
            self._pending = us_in if p == 0 else None
        else:
            self._pending = None
        axnpby(Q, [(1.0, self._qc[p])], 1.0)

    def rhs(self, p, t, us_in, u_out):
        """u_out = invMass(transport_p(u_p) + sum_q Q_pq)  (bi.py:265-338)."""
        if self.invmass.is_identity:
            self.transport[p](t, us_in[p], u_out)
            self._collide(p, us_in, u_out)
        else:
            self.transport[p](t, us_in[p], self.ux)
            self._collide(p, us_in, self.ux)
            self.invmass(self.ux, u_out)

    def _advance(self):
        self.odes.integrate(self.time, self.dt, self.us)

    def step(self):
        self._advance()
        self.time += self.dt
        self.nacptsteps += 1
        if self.residual is not None:
            self.residual(self.time, self.nacptsteps, self.uprevs, self.us)
        if self.moments is not None:
            self.moments(self.dt, self.time, self.us)
        if self.distribution is not None:
            self.distribution(self.dt, self.time, self.us)
        for a, b in zip(self.uprevs, self.us):
            a.copy_(b)

    def run(self, nsteps=None):
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
    DGFSBi1D(cfg, restart=restart).run()


if __name__ == '__main__':
    sys.exit(main())
