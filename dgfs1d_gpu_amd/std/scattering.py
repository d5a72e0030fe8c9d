"""Scattering-model plugins (reference dgfs1D/std/scattering.py).

``vhs-gll`` (:51-313) and ``vhs`` (:317-429) keep the reference constructor
``(cfg, vm, Ne=Ne)``, the ``prefac``/``omega`` properties and the single-node
``fs(d_f, d_g, d_Q, elem, modein, modeout)``; the solver uses the batched
``collide(u, Q)``, which evaluates every (m, r in sigModes[m], e) collision of
one RHS in one set of launches with the U/V contraction fused (std.py:352-356).
"""
from math import gamma

import numpy as np

from .. import _lib
from ..collision import CollisionPlan


class DGFSScatteringModelStd:
    scattering_model = None

    def __init__(self, cfg, velocitymesh, **kwargs):
        self.cfg = cfg
        self.vm = velocitymesh
        self.load_parameters()
        self.perform_precomputation()

    def load_parameters(self):
        raise NotImplementedError

    def perform_precomputation(self):
        raise NotImplementedError

    @property
    def prefac(self):
        return self._prefactor

    @property
    def omega(self):
        return self._omega


class DGFSVHSGLLScatteringModelStd(DGFSScatteringModelStd):
    scattering_model = 'vhs-gll'
    _variant = _lib.VHS_GLL

    def __init__(self, cfg, velocitymesh, **kwargs):
        self._Ne = kwargs.get('Ne')
        self._device = kwargs.get('device')
        self._ntask_max = kwargs.get('ntask_max')
        kind = cfg.lookupordefault('basis', 'kind', 'nodal-sem-gll')
        if self.scattering_model == 'vhs-gll' and kind not in (
                'nodal-sem-gll', 'nodal-gll', 'nodal-sem-gll-2'):
            raise RuntimeError('Only tested for nodal basis')
        super().__init__(cfg, velocitymesh, **kwargs)

    prefacBGK = property(lambda s: s._prefactorBGK)
    prefacESBGK = property(lambda s: s._prefactorESBGK)
    prefacSK = property(lambda s: s._prefactorSK)

    def load_parameters(self):
        """std/scattering.py:70-104"""
        cfg, vm = self.cfg, self.vm
        alpha = 1.0
        omega = float(cfg.lookupfloat('scattering-model', 'omega'))
        self._gamma = 2.0 * (1 - omega)
        dRef = float(cfg.lookupfloat('scattering-model', 'dRef'))
        Tref = float(cfg.lookupfloat('scattering-model', 'Tref'))
        invKn = vm.H0() * np.sqrt(2.0) * np.pi * vm.n0() * dRef * dRef * pow(
            Tref / vm.T0(), omega - 0.5)
        self._invKn = invKn
        self._prefactor = invKn * alpha / (pow(2.0, 2 - omega + alpha) * gamma(2.5 - omega) * np.pi)
        self._omega = omega
        # penalisation prefactors; muRef is optional here (the reference requires it)
        Pr = cfg.lookupordefault('scattering-model', 'Pr', 2. / 3.)
        if cfg.has_option('scattering-model', 'muRef'):
            muRef = float(cfg.lookupfloat('scattering-model', 'muRef'))
            t0 = vm.H0() / vm.u0()
            visc = muRef * ((vm.T0() / Tref) ** omega)
            p0 = vm.n0() * vm.R0 / vm.NA * vm.T0()
            self._prefactorBGK = t0 * 1.0 * p0 / visc
            self._prefactorESBGK = t0 * Pr * p0 / visc
            self._prefactorSK = self._prefactorBGK

    def perform_precomputation(self):
        vm = self.vm
        dt = _lib.F64 if self.cfg.dtype == np.float64 else _lib.F32
        ntask = self._ntask_max or max(1, self._Ne or 1)
        self.plan = CollisionPlan(vm.Nv(), vm.M(), vm.Nrho(), vm.L(), vm.qz(), vm.qw(),
                                  vm.sz(), vm.sw(), self._gamma, self._prefactor, dt,
                                  self._variant, ntask_max=ntask, device=self._device)
        self._single = {}

    # ------------------------------------------------------------ batched
    def set_tasks(self, K, Ne, tasks):
        if self._variant == _lib.VHS_GLL:
            tasks = [(o, f, f) for o, f, _ in tasks]       # fs() uses d_arr_in only
        self.plan.set_tasks(K, Ne, _fold_same_slice(tasks))
        self._single = {}

    def collide(self, u, Q, stream=None):
        self.plan(u, Q, stream)

    # ---------------------------------------------------- single node fs()
    def fs(self, d_f, d_g, d_Q, elem, modein, modeout):
        """d_Q[modeout, elem] += Q(d_f[modein, elem], d_g[modein, elem]) on
        (K, Ne, Nv) arrays -- the reference's per-node call
        (std/scattering.py:251-313 'vhs-gll' uses d_f only, :346-429 'vhs').
        Slow path (one task per call); the solver uses collide()."""
        K = d_Q.numel() // (self._Ne * self.vm.vsize())
        key = (K, elem, modein, modeout)
        if self._single.get('key') != key:
            s = modein * self._Ne + elem
            self.plan.set_tasks(K, self._Ne, [(modeout * self._Ne + elem, [(s, 1.0)],
                                               [(s, 1.0)])])
            self._single['key'] = key
        gl = self._variant == _lib.VHS_GLL or d_g is d_f
        self.plan(d_f, d_Q, ug=None if gl else d_g)


class DGFSVHSScatteringModelStd(DGFSVHSGLLScatteringModelStd):
    scattering_model = 'vhs'
    _variant = _lib.VHS


def _fold_same_slice(tasks):
    """Q(c1 x, c2 x) = c1 c2 Q(x, x): single-term tasks on one slice become
    f = g = sqrt(c1 c2) x (positive product), enabling the f == g kernel."""
    out = []
    for o, f, g in tasks:
        if len(f) == 1 and len(g) == 1 and f[0][0] == g[0][0] and f[0][1] * g[0][1] > 0:
            c = float(np.sqrt(f[0][1] * g[0][1]))
            out.append((o, [(f[0][0], c)], [(g[0][0], c)]))
        else:
            out.append((o, f, g))
    return out
