"""Scattering-model plugins (reference dgfs1D/std/scattering.py).

``vhs-gll`` (:51-313), ``vhs`` (:317-429), ``maxwell-gll`` (:707-743),
``bgk-gll`` (:445-706), ``esbgk-gll`` (:752-1185), ``bgk-direct-gll`` (:1186-1369), ``esbgk-direct-gll``
(:1372-1632) and ``shakov-direct-gll`` (:1636-1833) keep the reference constructor ``(cfg, vm, Ne=Ne)``, the
``prefac``/``omega`` properties and the single-node
``fs(d_f, d_g, d_Q, elem, modein, modeout)``; the solver uses the batched
``collide(u, Q)``, which evaluates every (m, r in sigModes[m], e) collision of
one RHS in one set of launches with the U/V contraction fused (std.py:352-356).

``[scattering-model] projection = conservative`` (default ``none``, the
reference behaviour) routes the collision output through the FP64
mass/momentum/energy projection before it is added to the RHS.
"""
import os
from math import gamma

import numpy as np

from .. import _lib
from ..collision import CollisionPlan, task_table
from ..relax import ConservativeProjection, Shakov


def _torch_dtype(cfg):
    import torch
    return torch.float64 if cfg.dtype == np.float64 else torch.float32


class DGFSScatteringModelStd:
    scattering_model = None

    def __init__(self, cfg, velocitymesh, **kwargs):
        self.cfg = cfg
        self.vm = velocitymesh
        self._Ne = kwargs.get('Ne')
        self._device = kwargs.get('device')
        self._ntask_max = kwargs.get('ntask_max')
        self.load_parameters()
        self.perform_precomputation()
        self._projection = self._make_projection()

    def _make_projection(self):
        mode = self.cfg.lookupordefault('scattering-model', 'projection', 'none')
        if mode == 'none':
            return None
        if mode != 'conservative':
            raise ValueError("scattering-model::projection must be 'none' or 'conservative'")
        vm = self.vm
        return ConservativeProjection(vm.cv(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(),
                                      _torch_dtype(self.cfg), self._device)

    def collide(self, u, Q, stream=None):
        """Q += the collision term of every task (projected when enabled)."""
        if self._projection is None:
            self._collide(u, Q, stream)
            return
        buf = self._projection.buffer(Q)
        self._collide(u, buf, stream)
        self._projection(buf, Q, Q.numel() // self.vm.vsize(), accumulate=True, stream=stream)

    def _collide(self, u, Q, stream=None):
        raise NotImplementedError

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
        folded = _fold_same_slice(tasks)
        tab = self.plan.set_tasks(K, Ne, folded)
        self._single = {}
        self.closure = self._make_closure(K, Ne, tasks)
        self._make_reference_state(K, Ne, tab, folded)

    def _make_reference_state(self, K, Ne, tab, tasks):
        """[scattering-model] reference-refresh = R (R > 0, FP32 runs with f = g
        tasks): every collide() evaluates Q(f, f) in the reference-state form
        Q(f0, f0) + Gain(h, f + f0) - h Loss(f) - f0 Loss(h), h = f - f0
        (dgfs_collide_reference), where f0 is the state of the most recent
        refresh and Q(f0, f0) comes from an FP64 plan.  The form is exact; it
        makes the FP32 rounding error scale with |f - f0| instead of |f|.  The
        reference state is refreshed on the first call and every R-th call
        after it.  R = 0 (default): the direct form."""
        self._ref = None
        R = int(self.cfg.lookupordefault('scattering-model', 'reference-refresh', 0))
        if R < 0:
            raise ValueError('scattering-model::reference-refresh must be >= 0')
        if R == 0 or self.plan.dtype == _lib.F64:
            return
        if not tab['fg_same']:
            raise ValueError('scattering-model::reference-refresh needs f = g collision tasks '
                             '(a nodal basis)')
        if self.closure is not None:
            raise ValueError('scattering-model::reference-refresh and closure = learned are '
                             'exclusive')
        import torch
        vm = self.vm
        plan64 = CollisionPlan(vm.Nv(), vm.M(), vm.Nrho(), vm.L(), vm.qz(), vm.qw(), vm.sz(),
                               vm.sw(), self._gamma, self._prefactor, _lib.F64, self._variant,
                               ntask_max=self.plan.ntask_max, device=self._device)
        plan64.set_tasks(K, Ne, tasks)
        n = K * Ne * vm.vsize()
        self._ref = dict(R=R, calls=0, plan64=plan64,
                         u0=torch.zeros(n, dtype=self.plan.tdtype, device=self._device),
                         Q0=torch.zeros(n, dtype=torch.float64, device=self._device))

    def refresh_reference(self, u, stream=None):
        """Make u the reference state: u0 = u, Q0 = Q(u, u) in FP64, and the
        FP32 plan's transforms of u0."""
        r = self._ref
        n = r['u0'].numel()
        r['u0'].copy_(u[:n])
        r['Q0'].zero_()
        r['plan64'](r['u0'].double(), r['Q0'], stream)
        self.plan.reference_set(r['u0'], stream)

    def _make_closure(self, K, Ne, tasks):
        """[scattering-model] closure = learned: Q = Q_Shakov + R_net per node,
        spectral fallback (see learned.LearnedClosure).  Needs a nodal basis
        (one task per slice, f = g = that slice) and muRef for Q_Shakov."""
        cfg = self.cfg
        mode = cfg.lookupordefault('scattering-model', 'closure', 'none')
        if mode == 'none':
            return None
        if mode != 'learned':
            raise ValueError("scattering-model::closure must be 'none' or 'learned'")
        for o, f, g in tasks:
            if len(f) != 1 or len(g) != 1 or f[0][0] != o or g[0][0] != o or \
                    abs(f[0][1] * g[0][1] - 1.0) > 1e-12:
                raise ValueError('the learned closure needs a nodal basis (slice -> slice tasks)')
        from ..learned import ClosureNet, LearnedClosure
        from ..relax import Shakov
        vm = self.vm
        Pr = float(cfg.lookupordefault('scattering-model', 'Pr', 2. / 3.))
        muRef = float(cfg.lookupfloat('scattering-model', 'muRef'))
        Tref = float(cfg.lookupfloat('scattering-model', 'Tref'))
        t0 = vm.H0() / vm.u0()
        visc = muRef * ((vm.T0() / Tref) ** self._omega)
        p0 = vm.n0() * vm.R0 / vm.NA * vm.T0()
        td = _torch_dtype(cfg)
        sk = Shakov(vm.vsize(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(), t0 * p0 / visc,
                    self._omega, Pr, td, self._device)
        net = ClosureNet.load(os.path.expanduser(cfg.lookup('scattering-model', 'closure-weights')),
                              self._device)
        return LearnedClosure(
            net, vm.Nv(), K * Ne, sk, self.plan, self._prefactor, vm.cw(), self._device,
            check_every=int(cfg.lookupordefault('scattering-model', 'closure-check-every', 10)),
            sample=float(cfg.lookupordefault('scattering-model', 'closure-sample', 0.25)),
            tol=float(cfg.lookupordefault('scattering-model', 'closure-tol', 0.05)),
            use_net=cfg.lookupordefault('scattering-model', 'closure-use-net', 'true').lower()
            in ('1', 'true', 'yes'),
            start=int(cfg.lookupordefault('scattering-model', 'closure-start', 0)),
            readmit=float(cfg.lookupordefault('scattering-model', 'closure-readmit', 0.5)),
            net_every=int(cfg.lookupordefault('scattering-model', 'closure-net-every', 1)))

    def _collide(self, u, Q, stream=None):
        if self.closure is not None:
            self.closure(u, Q)
        elif self._ref is not None:
            r = self._ref
            if r['calls'] % r['R'] == 0:
                self.refresh_reference(u, stream)
            r['calls'] += 1
            self.plan.reference(u, r['u0'], r['Q0'], Q, stream)
        else:
            self.plan(u, Q, stream)

    # ---------------------------------------------------- single node fs()
    def fs(self, d_f, d_g, d_Q, elem, modein, modeout):
        """d_Q[modeout, elem] += Q(d_f[modein, elem], d_g[modein, elem]) on
        (K, Ne, Nv) arrays -- the reference's per-node call
        (std/scattering.py:251-313 'vhs-gll' uses d_f only, :346-429 'vhs').
        Slow path (one task per call, never projected); the solver uses collide()."""
        tab = _single_task(self._single, self._Ne, self.vm.vsize(), d_Q, elem, modein, modeout,
                           self.plan.tdtype, self._device)
        gl = self._variant == _lib.VHS_GLL or d_g is d_f
        self.plan(d_f, d_Q, ug=None if gl else d_g, tasks=tab)


class DGFSVHSScatteringModelStd(DGFSVHSGLLScatteringModelStd):
    scattering_model = 'vhs'
    _variant = _lib.VHS


class DGFSMaxwellGLLScatteringModelStd(DGFSVHSGLLScatteringModelStd):
    """Maxwell molecules on the vhs-gll operator (std/scattering.py:707-743):
    gamma = 0, omega = 1 and prefactor H0 u0^2 / n0 / eps, eps read from
    [scattering-model]; the BGK and ES-BGK prefactors equal it.  With
    prefactor 1/(4 pi) the operator is the Boltzmann operator of the kernel
    B = 1/(4 pi), whose closed-form Bobylev-Krook-Wu solution is the 'bkw'
    initial condition.  As in the reference, the nodal-basis check applies to
    'vhs-gll' only."""
    scattering_model = 'maxwell-gll'

    def load_parameters(self):
        vm = self.vm
        self._gamma, self._omega = 0.0, 1.0
        nd = vm.H0() * vm.u0() ** 2 / vm.n0()
        eps = float(self.cfg.lookupfloat('scattering-model', 'eps'))
        self._prefactor = nd / eps
        self._prefactorBGK = self._prefactorESBGK = self._prefactor


class DGFSShakovDirectGLLScatteringModelStd(DGFSScatteringModelStd):
    """BGK-Shakov relaxation, iteration-free (std/scattering.py:1636-1833)."""
    scattering_model = 'shakov-direct-gll'

    def __init__(self, cfg, velocitymesh, **kwargs):
        kind = cfg.lookupordefault('basis', 'kind', 'nodal-sem-gll')
        if kind not in ('nodal-sem-gll', 'nodal-gll'):
            raise RuntimeError('Only tested for nodal basis')
        super().__init__(cfg, velocitymesh, **kwargs)

    def load_parameters(self):
        """prefac = t0 p0 / mu(T0), t0 = H0/u0 (std/scattering.py:1645-1661)."""
        cfg, vm = self.cfg, self.vm
        Pr = float(cfg.lookupordefault('scattering-model', 'Pr', 2. / 3.))
        omega = float(cfg.lookupfloat('scattering-model', 'omega'))
        muRef = float(cfg.lookupfloat('scattering-model', 'muRef'))
        Tref = float(cfg.lookupfloat('scattering-model', 'Tref'))
        t0 = vm.H0() / vm.u0()
        visc = muRef * ((vm.T0() / Tref) ** omega)
        p0 = vm.n0() * vm.R0 / vm.NA * vm.T0()
        self._prefactor = t0 * p0 / visc
        self._omega = omega
        self._Pr = Pr

    @property
    def Pr(self):
        return self._Pr

    def perform_precomputation(self):
        vm = self.vm
        self.model = Shakov(vm.vsize(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(),
                            self._prefactor, self._omega, self._Pr, _torch_dtype(self.cfg),
                            self._device)
        self._single = {}

    def set_tasks(self, K, Ne, tasks):
        self.model.set_tasks(K, Ne, tasks)

    def _collide(self, u, Q, stream=None):
        self.model(u, Q, stream)

    def fs(self, d_f, d_g, d_Q, elem, modein, modeout):
        """d_Q[modeout, elem] += Q_S(d_f[modein, elem]); d_g is not read
        (std/scattering.py:1786-1833).  Slow path, never projected."""
        tab = _single_task(self._single, self._Ne, self.vm.vsize(), d_Q, elem, modein, modeout,
                           self.model.tdtype, self._device)
        self.model(d_f, d_Q, tasks=tab)


class DGFSBGKDirectGLLScatteringModelStd(DGFSShakovDirectGLLScatteringModelStd):
    """BGK relaxation, iteration-free (std/scattering.py:1186-1369): the
    Maxwellian of the node's (rho, u, T) as target, frequency
    prefac rho T^(1-omega), prefac = t0 Pr p0 / mu(T0) with Pr = 1.  That is
    the Shakov term with Pr = 1 (its heat-flux factor (1 - Pr)/5 vanishes and
    its prefactor t0 p0 / mu(T0) is the same), so the Shakov kernels run it;
    a Pr in the deck is ignored, as in the reference."""
    scattering_model = 'bgk-direct-gll'

    def load_parameters(self):
        super().load_parameters()
        self._Pr = 1.0


class DGFSBGKGLLScatteringModelStd(DGFSShakovDirectGLLScatteringModelStd):
    """Conservative BGK (std/scattering.py:445-706): the discrete Maxwellian
    whose discrete mass, momentum and energy equal the node's, found by
    Newton's method (dgfs_bgk_conservative), as target; frequency
    prefac rho T^(1-omega), prefac = t0 p0 / mu(T0).  nodal-sem-gll only, as
    the reference."""
    scattering_model = 'bgk-gll'

    def __init__(self, cfg, velocitymesh, **kwargs):
        if cfg.lookupordefault('basis', 'kind', 'nodal-sem-gll') != 'nodal-sem-gll':
            raise RuntimeError('Only tested for nodal basis')
        super().__init__(cfg, velocitymesh, **kwargs)

    def load_parameters(self):
        super().load_parameters()
        self._Pr = 1.0

    def perform_precomputation(self):
        vm = self.vm
        self.model = Shakov(vm.vsize(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(),
                            self._prefactor, self._omega, 1.0, _torch_dtype(self.cfg),
                            self._device, target='bgk')
        self._single = {}


class DGFSESBGKGLLScatteringModelStd(DGFSBGKGLLScatteringModelStd):
    """Conservative ES-BGK (std/scattering.py:752-1185): the discrete
    Gaussian whose mass, momentum and second moments equal those of the ES
    tensor built from f and the conservative BGK target
    (dgfs_esbgk_conservative); prefac = t0 Pr p0 / mu(T0).  nodal-sem-gll
    only."""
    scattering_model = 'esbgk-gll'

    def load_parameters(self):
        DGFSShakovDirectGLLScatteringModelStd.load_parameters(self)
        self._prefactor *= self._Pr

    def perform_precomputation(self):
        vm = self.vm
        self.model = Shakov(vm.vsize(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(),
                            self._prefactor, self._omega, self._Pr, _torch_dtype(self.cfg),
                            self._device, target='esbgk-conservative')
        self._single = {}


class DGFSESBGKDirectGLLScatteringModelStd(DGFSShakovDirectGLLScatteringModelStd):
    """ES-BGK relaxation, iteration-free (std/scattering.py:1372-1632): the
    node's anisotropic Gaussian with tensor
    L = cw/rho ((1 - 1/Pr) sum f C C^T + (1/Pr) sum f_M C C^T) as target
    (dgfs_esbgk), frequency prefac rho T^(1-omega), prefac = t0 Pr p0 / mu(T0).
    Nodal bases only, as the reference."""
    scattering_model = 'esbgk-direct-gll'

    def load_parameters(self):
        super().load_parameters()
        self._prefactor *= self._Pr

    def perform_precomputation(self):
        vm = self.vm
        self.model = Shakov(vm.vsize(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(),
                            self._prefactor, self._omega, self._Pr, _torch_dtype(self.cfg),
                            self._device, target='esbgk')
        self._single = {}


def _single_task(cache, Ne, Nv, d_Q, elem, modein, modeout, tdtype, device):
    """Cached one-task table: slice (modein, elem) -> slice (modeout, elem)."""
    K = d_Q.numel() // (Ne * Nv)
    key = (K, elem, modein, modeout)
    if cache.get('key') != key:
        s = modein * Ne + elem
        cache['tab'] = task_table(K, Ne, [(modeout * Ne + elem, [(s, 1.0)], [(s, 1.0)])],
                                  tdtype, device)
        cache['key'] = key
    return cache['tab']


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
