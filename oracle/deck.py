"""Oracle: build the FP64 reference operator from an INI deck.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Reads the deck with configparser directly (independent of the product's
Dictionary) and assembles mesh (dgfs1D/mesh.py:85-121), basis, velocity mesh,
boundaries (std/bc.py), prefactor (std/scattering.py:70-83) and the initial
state (std/initcond.py:36-117; std/std.py:89-91, :219-232).
"""
from configparser import ConfigParser

import numpy as np

from . import collision as OC
from . import relax as OR
from . import std as OS
from .basis import OracleBasis
from .velocitymesh import VelocityMesh


class Deck:
    def __init__(self, ini_text, overrides=None, sz_override=None):
        cp = ConfigParser(inline_comment_prefixes=[';'])
        cp.optionxform = str
        cp.read_string(ini_text)
        for (sect, key), val in (overrides or {}).items():
            cp.set(sect, key, str(val))
        g = lambda s, k: float(cp.get(s, k))
        self.cp = cp
        # mesh (classic)
        H0 = g('mesh', 'H0')
        Ne = int(cp.get('mesh', 'Ne'))
        xm = np.linspace(g('mesh', 'xlo'), g('mesh', 'xhi'), Ne + 1) / H0
        h = np.diff(xm)
        self.Ne, self.xmesh, self.jac, self.invjac = Ne, xm, h / 2.0, 2.0 / h
        # basis
        K = int(cp.get('basis', 'K'))
        kind = cp.get('basis', 'kind')
        Nq = int(cp.get('basis', 'Nq')) if cp.has_option('basis', 'Nq') and \
            kind.startswith('modal') else None
        self.basis = OracleBasis(kind, K, Nq=Nq, Nqr=int(cp.get('basis', 'Nqr')))
        self.K = K
        # velocity mesh
        Nv = int(cp.get('velocity-mesh', 'Nv'))
        Nrho = int(cp.get('velocity-mesh', 'Nrho')) if cp.has_option('velocity-mesh', 'Nrho') \
            else Nv // 2
        M = int(cp.get('spherical-design-rule', 'M'))
        rule = cp.get('spherical-design-rule', 'ssrule')
        if rule == 'octahedral' and sz_override is None:
            sz_override = octahedral24_oracle()
        self.vm = vm = VelocityMesh(H0, g('non-dim', 'T0'), g('non-dim', 'rho0'),
                                    g('non-dim', 'molarMass0'), Nv, Nrho,
                                    g('velocity-mesh', 'cmax'), g('velocity-mesh', 'Tmax'),
                                    g('velocity-mesh', 'dev'), M, ssrule=rule,
                                    sz_override=sz_override)
        # scattering
        omega = g('scattering-model', 'omega')
        self.model = cp.get('scattering-model', 'type')
        shakov = None
        if self.model in ('shakov-direct-gll', 'bgk-direct-gll', 'esbgk-direct-gll', 'bgk-gll',
                          'esbgk-gll'):
            Pr = g('scattering-model', 'Pr') if cp.has_option('scattering-model', 'Pr') \
                else 2.0 / 3.0
            self.prefac = OR.shakov_prefactor(H0, vm.u0, vm.T0, vm.n0,
                                              g('scattering-model', 'muRef'),
                                              g('scattering-model', 'Tref'), omega)
            if self.model in ('esbgk-direct-gll', 'esbgk-gll'):
                self.prefac *= Pr                 # t0 Pr p0 / mu(T0)
            shakov = (omega, Pr)
            self.pre = None
        else:
            self.prefac, self.invKn = OC.vhs_prefactor(H0, vm.n0, vm.T0, omega,
                                                       g('scattering-model', 'dRef'),
                                                       g('scattering-model', 'Tref'))
            self.pre = OC.Precomp(Nv, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, 2.0 * (1.0 - omega))
        project = cp.has_option('scattering-model', 'projection') and \
            cp.get('scattering-model', 'projection') == 'conservative'
        # boundaries
        self.walls = [self._wall('soln-bcs-xlo', -1.0), self._wall('soln-bcs-xhi', 1.0)]
        self.op = OS.StdOracle(self.basis, self.invjac, vm, self.walls[0], self.walls[1],
                               self.pre, self.prefac, model=self.model, shakov=shakov,
                               project=project)
        # time
        ti, dt, tf = g('time-integrator', 'tstart'), g('time-integrator', 'dt'), \
            g('time-integrator', 'tend')
        nsteps = np.ceil((tf - ti) / dt)
        self.dt, self.t0, self.tend = (tf - ti) / nsteps, ti, tf
        self.scheme = cp.get('time-integrator', 'scheme')
        # solution points and initial state (Maxwellian IC only)
        b = self.basis
        self.xsol = np.array([0.5 * (xm[j] + xm[j + 1]) + self.jac[j] * b.z
                              for j in range(Ne)]).T
        sect = 'soln-ics'
        if cp.get(sect, 'type') != 'maxwellian':
            raise ValueError('oracle deck supports the maxwellian IC')
        f0 = OS.maxwellian(vm.cv, g(sect, 'rho') / vm.rho0, g(sect, 'ux') / vm.u0,
                           g(sect, 'uy') / vm.u0, g(sect, 'uz') / vm.u0, g(sect, 'T') / vm.T0)
        usol = np.broadcast_to(f0, (b.Nq, Ne, vm.vsize))
        self.u0 = np.einsum('kq,qen->ken', b.fwd, usol)

    def _wall(self, sect, nl):
        cp, vm = self.cp, self.vm
        kind = cp.get(sect, 'type')
        g = lambda k: float(cp.get(sect, k))
        if kind in ('dgfs-periodic', 'dgfs-cyclic'):
            return OS.Wall('copy', nl, vm.cv, vm.cw)
        if kind == 'dgfs-wall-diffuse':
            u = (g('ux') / vm.u0, g('uy') / vm.u0, g('uz') / vm.u0)
            f0 = OS.maxwellian(vm.cv, 1.0, *u, g('T') / vm.T0)
            return OS.Wall('diffuse', nl, vm.cv, vm.cw, f0=f0, ushift=u[0])
        if kind == 'dgfs-inlet':
            f0 = OS.maxwellian(vm.cv, g('rho') / vm.rho0, g('ux') / vm.u0, g('uy') / vm.u0,
                               g('uz') / vm.u0, g('T') / vm.T0)
            return OS.Wall('inlet', nl, vm.cv, vm.cw, f0=f0)
        if kind == 'dgfs-vanish':
            return OS.Wall('vanish', nl, vm.cv, vm.cw)
        if kind == 'dgfs-inlet-nondim':                    # std/bc.py:269-324
            ux = g('ux')
            f0 = OS.maxwellian(vm.cv, g('rho'), ux, 0.0, 0.0, g('T'))
            return OS.Wall('inlet', nl, vm.cv, vm.cw, f0=f0, ushift=ux)
        if kind == 'dgfs-wall-expr-diffuse':               # std/bc.py:136-198
            x = self.xmesh[0] if nl < 0 else self.xmesh[-1]
            ex = [cp.get(sect, k) for k in ('ux', 'uy', 'uz', 'T')]
            scale = (vm.u0, vm.u0, vm.u0, vm.T0)

            def state(t, ex=ex, x=x, scale=scale):
                env = {'__builtins__': {}, 't': t, 'x': x, 'pi': np.pi, 'sin': np.sin,
                       'cos': np.cos, 'exp': np.exp, 'tanh': np.tanh, 'sqrt': np.sqrt}
                return tuple(float(eval(e, env)) / sc for e, sc in zip(ex, scale))
            return OS.Wall('diffuse-expr', nl, vm.cv, vm.cw, state=state)
        raise ValueError('oracle deck: unsupported boundary ' + kind)

    def step(self, t, u):
        return OS.integrate(self.op, self.scheme, t, self.dt, u)


def octahedral24_oracle():
    """Same 24-point O_h-orbit 5-design, derived independently: a^2 is the
    admissible root of 2a^4 + (1-2a^2)^2 = 3/5."""
    roots = np.roots([6.0, -4.0, 0.4])
    s = float(min(r.real for r in roots))
    a, b = np.sqrt(s), np.sqrt(1 - 2 * s)
    half = []
    for base in ((a, a, b), (a, b, a), (b, a, a)):
        for sx in (1, -1):
            for sy in (1, -1):
                half.append((sx * base[0], sy * base[1], base[2]))
    half = np.array(half)
    return np.vstack([half, -half])
