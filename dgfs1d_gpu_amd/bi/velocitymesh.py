"""Velocity mesh of the binary-mixture solver (same construction and accessor
names as the reference dgfs1D/bi/velocitymesh.py:8-146): masses = molar mass /
molarMass0 (largest ratio at most 4), L = cmax/u0 + dev sqrt(Tmax/T0),
S = min over the two mass-ratio bounds (:60-63), R = 2S, full spherical design
of M points with sw = 4 pi / M."""
import numpy as np

from ..quadratures import zwgj, zwglj, zwgrjm, zwgrjp
from ..sphericaldesign import get_sphquadrule


class DGFSVelocityMeshBi:
    R0 = 8.3144598
    NA = 6.0221409e+23

    def __init__(self, cfg, device=None):
        self.cfg = cfg
        self._device = device
        self._nspcs = cfg.lookupint('velocity-mesh', 'nspcs')
        if self._nspcs != 2:
            raise ValueError('the binary-mixture solver needs nspcs = 2')
        self._construct_velocity_mesh()
        self._load_quadrature()
        self._construct_spherical_mesh()
        self._dev = {}

    def _construct_velocity_mesh(self):
        cfg = self.cfg
        self._H0 = float(cfg.lookupfloat('mesh', 'H0'))
        self._T0 = float(cfg.lookupfloat('non-dim', 'T0'))
        self._rho0 = float(cfg.lookupfloat('non-dim', 'rho0'))
        self._molarMass0 = float(cfg.lookupfloat('non-dim', 'molarMass0'))
        self._n0 = self._rho0 / self._molarMass0 * self.NA
        self._u0 = np.sqrt(2 * self.R0 / self._molarMass0 * self._T0)
        self._Nv = cfg.lookupint('velocity-mesh', 'Nv')
        self._vsize = self._Nv ** 3
        self._Nrho = cfg.lookupordefault('velocity-mesh', 'Nrho', int(self._Nv / 2))
        cmax = float(cfg.lookupfloat('velocity-mesh', 'cmax')) / self._u0
        Tmax = float(cfg.lookupfloat('velocity-mesh', 'Tmax')) / self._T0
        dev = float(cfg.lookupfloat('velocity-mesh', 'dev'))
        self._masses = [float(cfg.lookupfloat('velocity-mesh', 'molarMass%d' % (p + 1)))
                        / self._molarMass0 for p in range(self._nspcs)]
        mr = max(self._masses)
        if mr > 4:
            raise ValueError('Mass ratio is too high!')
        self._mr = mr
        self._L = cmax + dev * np.sqrt(Tmax)
        self._S = min(
            self._L * 2.0 / (max(4.0 * mr / (1.0 + mr), 2.0) + np.sqrt(1.0 + mr) + 1.0),
            self._L * 2.0 / (max(4.0 / (1.0 + mr), 2.0) + np.sqrt(1.0 + 1.0 / mr) + 1.0))
        self._R = 2 * self._S
        self._cw = (2.0 * self._L / self._Nv) ** 3
        N = self._Nv
        c0 = np.linspace(-self._L + self._L / N, self._L - self._L / N, N)
        idx = np.arange(self._vsize)
        self._cv = np.stack([c0[idx // (N * N)], c0[(idx % (N * N)) // N], c0[idx % N]])

    def _load_quadrature(self):
        rule = self.cfg.lookupordefault('velocity-mesh', 'quad-rule', 'jacobi')
        rules = {'jacobi': zwgj, 'lobatto': zwglj, 'radau-m': zwgrjm, 'radau-p': zwgrjp}
        if rule not in rules:
            raise ValueError('Valid quads: %s' % list(rules))
        a = self.cfg.lookupordefault('velocity-mesh', 'quad-rule-alpha', 0.0)
        b = self.cfg.lookupordefault('velocity-mesh', 'quad-rule-beta', 0.0)
        z, w = rules[rule](self._Nrho, a, b)
        self._qz = (self._R / 2.0) * (1.0 + z)
        self._qw = (self._R / 2.0) * w

    def _construct_spherical_mesh(self):
        self._ssrule = self.cfg.lookup('spherical-design-rule', 'ssrule')
        self._M = self.cfg.lookupint('spherical-design-rule', 'M')
        self._sz = get_sphquadrule('symmetric', rule=self._ssrule, npts=self._M)
        self._sw = 4 * np.pi / self._M

    def _device_comp(self, i):
        import torch
        if i not in self._dev:
            td = torch.float64 if self.cfg.dtype == np.float64 else torch.float32
            self._dev[i] = torch.tensor(self._cv[i], dtype=td, device=self._device)
        return self._dev[i]

    def nspcs(self): return self._nspcs
    def masses(self): return self._masses
    def Nv(self): return self._Nv
    def Nrho(self): return self._Nrho
    def M(self): return self._M
    def L(self): return self._L
    def vsize(self): return self._vsize
    def qz(self): return self._qz
    def qw(self): return self._qw
    def cv(self): return self._cv
    def cw(self): return self._cw
    def sz(self): return self._sz
    def sw(self): return self._sw
    def H0(self): return self._H0
    def T0(self): return self._T0
    def rho0(self): return self._rho0
    def n0(self): return self._n0
    def u0(self): return self._u0
    def molarMass0(self): return self._molarMass0
    def d_cvx(self): return self._device_comp(0)
    def d_cvy(self): return self._device_comp(1)
    def d_cvz(self): return self._device_comp(2)
