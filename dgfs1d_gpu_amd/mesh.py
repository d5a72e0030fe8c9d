"""1D element mesh (reference dgfs1D/mesh.py:8-142): point distributions
'classic' | 'expr' | 'glj' | 'bump' | 'loadtxt', non-dimensionalised by
mesh::H0, jac = h/2, invjac = 2/h.  One GPU owns the whole mesh; the slab
split across ranks of the reference (mesh.py:101-111) is a design note only."""
import numpy as np

from .nputil import npeval
from .quadratures import zwglj

msect = 'mesh'


class Mesh:
    def classic(self):
        xlo, xhi = self.cfg.lookupfloat(msect, 'xlo'), self.cfg.lookupfloat(msect, 'xhi')
        if not xlo < xhi:
            raise ValueError('xlo must be less than xhi')
        Ne = self.cfg.lookupint(msect, 'Ne')
        if Ne <= 0:
            raise ValueError('Need at least one element to define a mesh')
        return np.linspace(xlo, xhi, Ne + 1, dtype=np.float64)

    def expr(self):
        xm = np.sort(np.array(npeval(self.cfg.lookupexpr(msect, 'expr'), {}),
                              dtype=np.float64).ravel())
        if len(xm) < 2:
            raise ValueError('Need at least two points to define the mesh')
        return xm

    def glj(self):
        xlo, xhi = self.cfg.lookupfloat(msect, 'xlo'), self.cfg.lookupfloat(msect, 'xhi')
        a = self.cfg.lookupordefault(msect, 'alpha', 0.0)
        b = self.cfg.lookupordefault(msect, 'beta', 0.0)
        Ne = self.cfg.lookupint(msect, 'Ne')
        z, _ = zwglj(Ne + 1, a, b)
        return 0.5 * (xhi + xlo) + 0.5 * (xhi - xlo) * z

    def bump(self):
        xlo, xhi = self.cfg.lookupfloat(msect, 'xlo'), self.cfg.lookupfloat(msect, 'xhi')
        mu = self.cfg.lookupordefault(msect, 'mu', 0.0)
        Ne = self.cfg.lookupint(msect, 'Ne')
        z = np.tanh(2 * mu * np.arange(Ne + 1) / Ne - mu) / np.tanh(mu)
        return 0.5 * (xhi + xlo) + 0.5 * (xhi - xlo) * z

    def loadtxt(self):
        xm = np.sort(np.loadtxt(self.cfg.lookupordefault(msect, 'file', 'mesh.txt'),
                                dtype=np.float64, comments='#').ravel())
        if len(xm) < 2:
            raise ValueError('Need at least two points to define the mesh')
        return xm

    _types = {'classic': classic, 'expr': expr, 'glj': glj, 'bump': bump, 'loadtxt': loadtxt}

    def __init__(self, cfg):
        self.cfg = cfg
        mt = cfg.lookupordefault(msect, 'type', 'classic')
        if mt not in self._types:
            raise ValueError('Valid mesh: %s' % list(self._types))
        xmesh = np.asarray(self._types[mt](self), dtype=np.float64)
        self._H0 = float(cfg.lookupfloat(msect, 'H0'))
        xmesh = xmesh / self._H0
        self._Ne = len(xmesh) - 1
        h = np.diff(xmesh).reshape(-1, 1)
        if np.min(h) <= 1e-8:
            raise ValueError('Too small element size')
        self._jac, self._invjac = h / 2.0, 2.0 / h
        self._xmesh = xmesh
        self._xlo, self._xhi = xmesh.min(), xmesh.max()

    Ne = property(lambda s: s._Ne)
    jac = property(lambda s: s._jac)
    invjac = property(lambda s: s._invjac)
    xmesh = property(lambda s: s._xmesh)
    xlo = property(lambda s: s._xlo)
    xhi = property(lambda s: s._xhi)
    H0 = property(lambda s: s._H0)
