"""Initial conditions (host, FP64; reference dgfs1D/std/initcond.py:36-454).
Values are evaluated on the solution points of every element and uploaded
once; the registry keys are the reference's ``model`` strings."""
import numpy as np

from ..nputil import ndrange, npeval


def maxwellian(cv, rho, ux, uy, uz, T):
    u = np.array([ux, uy, uz]).reshape(3, 1)
    return (rho / (np.pi * T) ** 1.5) * np.exp(-np.sum((cv - u) ** 2, axis=0) / T)


class DGFSInitConditionStd:
    model = None

    def __init__(self, cfg, vm, name, **kwargs):
        self.cfg, self.vm, self.name = cfg, vm, name
        self._init_vals = None
        self.load_parameters(name, **kwargs)

    def load_parameters(self, name, **kwargs):
        pass

    def get_init_vals(self):
        return self._init_vals

    def apply_init_vals(self, usol, Nq, Ne, xsol, **kwargs):
        usol[:, :, :] = self._init_vals[None, None, :]


class DGFSMaxwellianInitConditionStd(DGFSInitConditionStd):
    """initcond.py:36-117; wall=True forces rho = 1 (used by the diffuse wall)."""
    model = 'maxwellian'

    def load_parameters(self, name, wall=False):
        u0 = self.vm.u0()
        self._rhoini = 1.0 if wall else float(self.cfg.lookupfloat(name, 'rho')) / self.vm.rho0()
        self._Tini = float(self.cfg.lookupfloat(name, 'T')) / self.vm.T0()
        self._uini = np.array([float(self.cfg.lookupfloat(name, k)) / u0
                               for k in ('ux', 'uy', 'uz')]).reshape(3, 1)
        self._init_vals = maxwellian(self.vm.cv(), self._rhoini, *self._uini[:, 0], self._Tini)

    def unondim(self):
        return self._uini


class DGFSBKWInitConditionStd(DGFSInitConditionStd):
    """Bobylev-Krook-Wu closed-form state (initcond.py:121-141)."""
    model = 'bkw'

    def load_parameters(self, name, **kwargs):
        t0 = float(self.cfg.lookupfloat(name, 't0'))
        K = 1.0 - np.exp(-t0 / 6.0)
        c2 = np.sum(self.vm.cv() ** 2, axis=0)
        self._init_vals = (1.0 / (2 * (2 * np.pi * K) ** 1.5) * np.exp(-c2 / (2.0 * K))
                           * ((5 * K - 3) / K + (1 - K) / (K * K) * c2))


class DGFSMaxwellianExprInitConditionStd(DGFSInitConditionStd):
    """Maxwellian with rho, u, T given as expressions of x (initcond.py:144-238)."""
    model = 'maxwellian-expr'
    nondim = False

    def read_params(self, xsol):
        vars = dict(self.cfg.section_values('non-dim', np.float64))
        vars.update(self.cfg.section_values('mesh', np.float64))
        vars['x'] = xsol
        sc = (1.0, 1.0, 1.0) if self.nondim else (self.vm.rho0(), self.vm.u0(), self.vm.T0())
        ev = lambda k: npeval(self.cfg.lookupexpr(self.name, k), vars)
        return (ev('rho') / sc[0], ev('ux') / sc[1], ev('uy') / sc[1], ev('uz') / sc[1],
                ev('T') / sc[2])

    def apply_init_vals(self, usol, Nq, Ne, xsol, **kwargs):
        vals = [np.broadcast_to(np.asarray(v, dtype=np.float64), (Nq, Ne))
                for v in self.read_params(xsol)]
        for q, e in ndrange(Nq, Ne):
            usol[q, e, :] = maxwellian(self.vm.cv(), *(v[q, e] for v in vals))

    def get_init_vals(self):
        raise ValueError('this model depends on spatial coordinates!')


class DGFSMaxwellianExprNonDimInitConditionStd(DGFSMaxwellianExprInitConditionStd):
    model = 'maxwellian-expr-nondim'
    nondim = True


class DGFSSodShockTubeInitConditionStd(DGFSInitConditionStd):
    """Left/right Maxwellians split at the middle element (initcond.py:312-368)."""
    model = 'sod-shock-tube'

    def apply_init_vals(self, usol, Nq, Ne, xsol, **kwargs):
        c, vm = self.cfg, self.vm
        Ml = maxwellian(vm.cv(), float(c.lookupfloat(self.name, 'rho-left')) / vm.rho0(),
                        c.lookupordefault(self.name, 'ux-left', 0.0) / vm.u0(), 0.0, 0.0,
                        float(c.lookupfloat(self.name, 'T-left')) / vm.T0())
        Mr = maxwellian(vm.cv(), float(c.lookupfloat(self.name, 'rho-right')) / vm.rho0(),
                        c.lookupordefault(self.name, 'ux-right', 0.0) / vm.u0(), 0.0, 0.0,
                        float(c.lookupfloat(self.name, 'T-right')) / vm.T0())
        if Ne % 2:
            raise ValueError('Not implemented')
        for q, e in ndrange(Nq, Ne):
            usol[q, e, :] = Ml if e < Ne // 2 else Mr
