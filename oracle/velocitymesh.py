"""Oracle: velocity mesh, radial quadrature, half-sphere rule (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates ``dgfs1D/std/velocitymesh.py:24-105`` and the stored/tabulated rule
loader ``dgfs1D/sphericaldesign/__init__.py:7-105``.  The Womersley point sets
are data; they are read from the product package's data directory (plain text
tables, no code is imported from the package).
"""
import os
import re

import numpy as np

from . import quadratures as Q

R0 = 8.3144598          # std/velocitymesh.py:9
NA = 6.0221409e+23      # std/velocitymesh.py:10

_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     'dgfs1d_gpu_amd', 'sphericaldesign', 'symmetric')


def load_design(name, npts):
    """Stored symmetric design ``<name>-ssXXX-m<npts>.txt``
    (sphericaldesign/__init__.py:174-210)."""
    best = None
    for fn in os.listdir(_DATA):
        m = re.match(r'([a-zA-Z0-9\-~+]+)-ss(\d+)(?:-m(\d+))?\.txt$', fn)
        if m and m.group(1) == name and int(m.group(3) or -1) == npts:
            if best is None or int(m.group(2)) > best[1]:
                best = (fn, int(m.group(2)))
    if best is None:
        raise ValueError('No suitable spherical quadrature rule found')
    return np.loadtxt(os.path.join(_DATA, best[0]))


class VelocityMesh:
    """Velocity mesh of std/velocitymesh.py:24-105 (plus the rule loader).

    ``sz_override`` supplies a tabulated half-sphere (M, 3) when no stored
    rule with 2M points exists (reference accepts an inline ``ssrule``).
    """

    def __init__(self, H0, T0, rho0, molarMass0, Nv, Nrho, cmax, Tmax, dev,
                 M, ssrule='womersley', quad='jacobi', qa=0.0, qb=0.0,
                 sz_override=None):
        self.H0, self.T0, self.rho0, self.molarMass0 = H0, T0, rho0, molarMass0
        self.n0 = rho0 / molarMass0 * NA                     # :30
        self.u0 = np.sqrt(2.0 * R0 / molarMass0 * T0)        # :31
        self.Nv = Nv
        self.vsize = Nv ** 3
        self.Nrho = Nrho
        cmax = cmax / self.u0                                # :45
        Tmax = Tmax / T0                                     # :48
        self.L = cmax + dev * np.sqrt(Tmax)                  # :51
        self.S = self.L * 2.0 / (3.0 + np.sqrt(2.0))         # :52
        self.R = 2.0 * self.S                                # :53
        self.dev = dev
        self.cw = (2.0 * self.L / Nv) ** 3                   # :60
        c0 = np.linspace(-self.L + self.L / Nv, self.L - self.L / Nv, Nv)
        I, J, K = np.meshgrid(np.arange(Nv), np.arange(Nv), np.arange(Nv),
                              indexing='ij')
        self.cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[K.ravel()]])  # :65-72
        rules = {'jacobi': Q.zwgj, 'lobatto': Q.zwglj,
                 'radau-m': Q.zwgrjm, 'radau-p': Q.zwgrjp}
        z, w = rules[quad](Nrho, qa, qb)                     # :93
        self.qz = (self.R / 2.0) * (1.0 + z)                 # :95
        self.qw = (self.R / 2.0) * w                         # :96
        self.M = M
        if sz_override is not None:
            self.sz = np.asarray(sz_override, dtype=np.float64)[:M]
        else:
            self.sz = load_design(ssrule, 2 * M)[:M]         # :101-104
        self.sw = 2.0 * np.pi / M                            # :105
