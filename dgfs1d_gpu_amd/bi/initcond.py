"""Initial conditions of the binary mixture (host, FP64; reference
dgfs1D/bi/initcond.py:8-203).  Species p is the Maxwellian
n_p (m_p/(pi T))^1.5 exp(-m_p |c - u|^2 / T) in the solver's units (number
density / n0, velocity / u0, temperature / T0)."""
import logging

import numpy as np

from ..nputil import ndrange, npeval

log = logging.getLogger('dgfs1d_gpu_amd')


def maxwellian_m(cv, m, nden, ux, uy, uz, T):
    u = np.array([ux, uy, uz]).reshape(3, 1)
    return nden * (m / np.pi / T) ** 1.5 * np.exp(-m * np.sum((cv - u) ** 2, axis=0) / T)


class DGFSInitConditionBi:
    model = None

    def __init__(self, cfg, vm, name, **kwargs):
        self.cfg, self.vm, self.name = cfg, vm, name
        self._nspcs = vm.nspcs()
        self._init_vals = None
        self.load_parameters(name, **kwargs)

    def load_parameters(self, name, **kwargs):
        pass

    def get_init_vals(self):
        return self._init_vals

    def apply_init_vals(self, p, usol, Nq, Ne, xsol):
        if not 0 <= p < self._nspcs:
            raise ValueError('species index out of range')
        usol[:, :, :] = self._init_vals[p][None, None, :]


class DGFSMaxwellianInitConditionBi(DGFSInitConditionBi):
    """initcond.py:38-128; wall=True sets every number density to 1 (the
    diffuse wall's f0).  The moment check of :68-126 is reported, not raised,
    as in the reference."""
    model = 'maxwellian'

    def load_parameters(self, name, wall=False):
        vm = self.vm
        u0 = vm.u0()
        self._ndenini = [1.0] * self._nspcs
        if not wall:
            self._ndenini = [float(self.cfg.lookupfloat(name, 'nden%d' % (p + 1))) / vm.n0()
                             for p in range(self._nspcs)]
        self._Tini = float(self.cfg.lookupfloat(name, 'T')) / vm.T0()
        self._uini = np.array([float(self.cfg.lookupfloat(name, k)) / u0
                               for k in ('ux', 'uy', 'uz')]).reshape(3, 1)
        m = vm.masses()
        self._init_vals = [maxwellian_m(vm.cv(), m[p], self._ndenini[p], *self._uini[:, 0],
                                        self._Tini) for p in range(self._nspcs)]
        cv, cw = vm.cv(), vm.cw()
        for p in range(self._nspcs):
            f, mr = self._init_vals[p], m[p]
            rho = f.sum() * mr * cw
            u = (cv * f).sum(1) * mr * cw / rho
            T = (f * ((cv - u[:, None]) ** 2).sum(0)).sum() * (2.0 / 3.0 * mr * cw * mr) / rho
            ok = (np.allclose(self._ndenini[p], rho / mr) and np.allclose(self._uini[:, 0], u)
                  and np.allclose(self._Tini, T))
            if not ok:
                log.warning('Bulk properties not conserved! See Nv, dev (species %d)', p)

    def unondim(self):
        return self._uini


class DGFSMaxwellianExprInitConditionBi(DGFSInitConditionBi):
    """initcond.py:131-203: nden1, nden2, ux, uy, uz, T as expressions of x."""
    model = 'maxwellian-expr'

    def apply_init_vals(self, p, usol, Nq, Ne, xsol):
        if not 0 <= p < self._nspcs:
            raise ValueError('species index out of range')
        vm, cfg = self.vm, self.cfg
        m = vm.masses()[p]
        env = {'x': xsol}
        nden = npeval(cfg.lookupexpr(self.name, 'nden%d' % (p + 1)), env) / vm.n0()
        ux = npeval(cfg.lookupexpr(self.name, 'ux'), env) / vm.u0()
        uy = npeval(cfg.lookupexpr(self.name, 'uy'), env) / vm.u0()
        uz = npeval(cfg.lookupexpr(self.name, 'uz'), env) / vm.u0()
        T = npeval(cfg.lookupexpr(self.name, 'T'), env) / vm.T0()
        vals = [np.broadcast_to(np.asarray(v, dtype=np.float64), (Nq, Ne))
                for v in (nden, ux, uy, uz, T)]
        for q, e in ndrange(Nq, Ne):
            f = maxwellian_m(vm.cv(), m, *(v[q, e] for v in vals))
            if not np.allclose(vals[0][q, e], f.sum() * vm.cw(), atol=1e-5):
                raise ValueError('Bulk properties not conserved! Check Nv, dev')
            usol[q, e, :] = f
