"""Oracle: build the FP64 binary-mixture operator from an INI deck.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Reads the deck with configparser directly and assembles mesh
(dgfs1D/mesh.py:85-121), basis, mixture velocity mesh (bi/velocitymesh.py),
walls per species (bi/bc.py with bi/kernels/bcs/*.mako), pair tables and
prefactors (bi/scattering.py:54-206) and the Maxwellian initial state of
each species (bi/initcond.py:38-66; bi/bi.py:177-186).
"""
from configparser import ConfigParser

import numpy as np

from . import bivss as OB
from . import std as OS
from .basis import OracleBasis
from .velocitymesh import load_design


class BiDeck:
    def __init__(self, ini_text):
        cp = ConfigParser(inline_comment_prefixes=[';'])
        cp.optionxform = str
        cp.read_string(ini_text)
        g = lambda s, k: float(cp.get(s, k))
        self.cp = cp
        H0 = g('mesh', 'H0')
        Ne = int(cp.get('mesh', 'Ne'))
        xm = np.linspace(g('mesh', 'xlo'), g('mesh', 'xhi'), Ne + 1) / H0
        h = np.diff(xm)
        self.Ne, self.jac, self.invjac = Ne, h / 2.0, 2.0 / h
        K = int(cp.get('basis', 'K'))
        kind = cp.get('basis', 'kind')
        Nq = int(cp.get('basis', 'Nq')) if cp.has_option('basis', 'Nq') and \
            kind.startswith('modal') else None
        self.basis = b = OracleBasis(kind, K, Nq=Nq, Nqr=int(cp.get('basis', 'Nqr')))
        self.K = K
        ns = int(cp.get('velocity-mesh', 'nspcs'))
        Nv = int(cp.get('velocity-mesh', 'Nv'))
        Nrho = int(cp.get('velocity-mesh', 'Nrho'))
        mm = [g('velocity-mesh', 'molarMass%d' % (p + 1)) for p in range(ns)]
        self.vm = vm = OB.BiVelocityMesh(
            H0, g('non-dim', 'T0'), g('non-dim', 'rho0'), g('non-dim', 'molarMass0'), Nv, Nrho,
            g('velocity-mesh', 'cmax'), g('velocity-mesh', 'Tmax'), g('velocity-mesh', 'dev'),
            mm, int(cp.get('spherical-design-rule', 'M')),
            ssrule=cp.get('spherical-design-rule', 'ssrule'))
        m = vm.masses
        szpre = load_design(cp.get('scattering-model', 'ssrulepre'),
                            int(cp.get('scattering-model', 'Mpre')))
        self.pres, self.prefac = {}, {}
        for p in range(ns):
            for q in range(ns):
                pq = '%d%d' % (p, q)
                al, om = g('scattering-model', 'alpha' + pq), g('scattering-model', 'omega' + pq)
                self.prefac[p, q], _ = OB.vss_prefactor(
                    m, p, q, vm.n0, H0, vm.T0, al, om, g('scattering-model', 'dRef' + pq),
                    g('scattering-model', 'Tref' + pq))
                self.pres[p, q] = OB.BiPrecomp(Nv, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, szpre,
                                               m[p], m[q], al, om)
        trs = []
        for p in range(ns):
            wl, wr = self._wall('soln-bcs-xlo', -1.0, p), self._wall('soln-bcs-xhi', 1.0, p)
            trs.append(OS.StdOracle(b, self.invjac, vm, wl, wr, None, 0.0))
        project = None
        if cp.has_option('scattering-model', 'projection') and \
                cp.get('scattering-model', 'projection') == 'conservative':
            project = (vm.cv, m[0], m[1])
        self.op = OB.BiOracle(trs, self.pres, self.prefac, project=project)
        ti, dt, tf = g('time-integrator', 'tstart'), g('time-integrator', 'dt'), \
            g('time-integrator', 'tend')
        nsteps = np.ceil((tf - ti) / dt)
        self.dt, self.t0 = (tf - ti) / nsteps, ti
        self.scheme = cp.get('time-integrator', 'scheme')
        sect = 'soln-ics'
        if cp.get(sect, 'type') != 'maxwellian':
            raise ValueError('oracle deck supports the maxwellian IC')
        u = [g(sect, k) / vm.u0 for k in ('ux', 'uy', 'uz')]
        self.u0 = []
        for p in range(ns):
            f0 = OB.maxwellian(vm.cv, m[p], g(sect, 'nden%d' % (p + 1)) / vm.n0, u,
                               g(sect, 'T') / vm.T0)
            usol = np.broadcast_to(f0, (b.Nq, Ne, vm.vsize))
            self.u0.append(np.einsum('kq,qen->ken', b.fwd, usol))

    def _wall(self, sect, nl, p):
        cp, vm = self.cp, self.vm
        kind = cp.get(sect, 'type')
        g = lambda k: float(cp.get(sect, k))
        m = vm.masses[p]
        if kind == 'dgfs-periodic':
            return OS.Wall('copy', nl, vm.cv, vm.cw)
        if kind == 'dgfs-wall-diffuse':                       # bc.py:25-90
            u = (g('ux') / vm.u0, g('uy') / vm.u0, g('uz') / vm.u0)
            f0 = OB.maxwellian(vm.cv, m, 1.0, u, g('T') / vm.T0)
            return OS.Wall('diffuse', nl, vm.cv, vm.cw, f0=f0, ushift=u[0])
        if kind == 'dgfs-inlet':                              # bc.py:208-249 (no shift)
            u = (g('ux') / vm.u0, g('uy') / vm.u0, g('uz') / vm.u0)
            f0 = OB.maxwellian(vm.cv, m, g('nden%d' % (p + 1)) / vm.n0, u, g('T') / vm.T0)
            return OS.Wall('inlet', nl, vm.cv, vm.cw, f0=f0)
        raise ValueError('oracle deck: unsupported boundary ' + kind)

    def step(self, t, us):
        return OB.integrate_bi(self.op, self.scheme, t, self.dt, us)

    # --------------------------------------------------------- IMEX path
    def imex_parts(self):
        """(explicit, boltz, MixPenal, to_points, to_coeffs) of the IMEX mixture
        on the species-major stack F of shape (2 K Ne, n): explicit =
        invMass(transport_p(f_p)), boltz = invMass(sum_q Q_pq) (with the deck's
        projection), and the species' Shakov penalisations
        ([penalized-scattering-model] Pr).  On a modal basis the stack holds
        the values at the K Gauss-Legendre points of each element (the modes
        evaluated there with numpy.polynomial.legendre); to_points/to_coeffs
        map coefficient stacks to that state and back (identity on a nodal
        basis)."""
        from . import imex as OI
        from .velocitymesh import NA
        op, vm, cp = self.op, self.vm, self.cp
        b, K, Ne, n = op.b, op.K, op.Ne, vm.vsize
        kind = cp.get('basis', 'kind')
        if kind.startswith('nodal'):
            V = np.eye(K)
        elif kind == 'modal-orthonormal-legendre':
            from numpy.polynomial import legendre
            z, _ = legendre.leggauss(K)
            V = np.stack([np.sqrt(i + 0.5) * legendre.legval(z, np.eye(K)[i])
                          for i in range(K)]).T                         # V[q, m]
        else:
            raise ValueError('oracle IMEX mixture: unsupported basis ' + kind)
        Vi = np.linalg.inv(V)
        amap = lambda A, F: np.concatenate(
            [(A @ F[p * K * Ne:(p + 1) * K * Ne].reshape(K, -1)).reshape(K * Ne, n)
             for p in range(2)])
        to_points = lambda F: amap(V, F)
        to_coeffs = lambda F: amap(Vi, F)
        split = lambda F: [F[p * K * Ne:(p + 1) * K * Ne].reshape(K, Ne, n)
                           for p in range(2)]
        im = lambda x: (b.invmass @ x.reshape(K, -1)).reshape(K * Ne, n)

        def explicit(t, W):
            us = split(to_coeffs(W))
            return to_points(np.concatenate([im(op.tr[p].transport(t, us[p]))
                                             for p in range(2)]))

        def boltz(t, W):
            us = split(to_coeffs(W))
            c = [op.collision(p, us) for p in range(2)]
            if op.project is not None:
                from .relax import project_mixture
                cv, m0, m1 = op.project
                for m in range(K):
                    for e in range(Ne):
                        c[0][m, e], c[1][m, e] = project_mixture(c[0][m, e], c[1][m, e], cv,
                                                                 m0, m1)
            return to_points(np.concatenate([im(c[p]) for p in range(2)]))

        Pr = float(cp.get('penalized-scattering-model', 'Pr')) \
            if cp.has_option('penalized-scattering-model', 'Pr') else 2.0 / 3.0
        g = lambda k: float(cp.get('scattering-model', k))
        sps = []
        for p in range(2):
            pp = '%d%d' % (p, p)
            mm = float(cp.get('velocity-mesh', 'molarMass%d' % (p + 1)))
            pre = OI.species_prefactor(float(cp.get('mesh', 'H0')), vm.u0, vm.n0, vm.T0, mm,
                                       NA, g('alpha' + pp), g('omega' + pp), g('dRef' + pp),
                                       g('Tref' + pp))
            sps.append(OI.SpeciesPenal(vm.cv, vm.cw, vm.masses[p], pre, g('omega' + pp), Pr))
        return explicit, boltz, OI.MixPenal(sps, K * Ne), to_points, to_coeffs
