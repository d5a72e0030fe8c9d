"""Penalisation models of the IMEX path (reference dgfs1D/astd/scattering.py).

``shakov-direct-gll`` (:355-397) keeps the reference interface used by the
integrators: ``moment``, ``constructMaxwellian``, ``collide``,
``updateMomentLM``, ``updateDistLM``, ``updateMomentARS``, ``updateDistARS``,
``nalph`` and ``prefac``.  Moment vectors are FP64 device tensors
(K*Ne, 14); distributions are (K*Ne*Nv) tensors of the solver dtype.  Every
operation is one or two HIP launches (libdgfs_hip.so, imex.hip).
"""
import numpy as np

from .. import _lib


class DGFSScatteringModelAstd:
    scattering_model = None

    def __init__(self, cfg, velocitymesh, **kwargs):
        self.cfg = cfg
        self.vm = velocitymesh
        self._Ne = kwargs.get('Ne')
        self._device = kwargs.get('device')
        self.load_parameters()
        self.perform_precomputation()

    def load_parameters(self):
        raise NotImplementedError

    def perform_precomputation(self):
        raise NotImplementedError

    @property
    def prefac(self):
        return self._prefactor


class DGFSShakovDirectGLLScatteringModelAstd(DGFSScatteringModelAstd):
    """14-moment Shakov penalisation P = nu (M_S - f)."""
    scattering_model = 'shakov-direct-gll'
    nalph = 14
    srate = 1.0          # the stresses relax at nu

    def load_parameters(self):
        """t0 p0 / mu(T0), parameters of [scattering-model] (astd/scattering.py:362-376)."""
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

    def perform_precomputation(self):
        import torch
        self.torch = torch
        vm = self.vm
        self.tdtype = torch.float64 if self.cfg.dtype == np.float64 else torch.float32
        self.dt_code = _lib.dtype_code(self.tdtype)
        self.Nv = vm.vsize()
        self.cx, self.cy, self.cz = vm.d_cvx(), vm.d_cvy(), vm.d_cvz()
        self.cw = float(vm.cw())
        self._work = None

    # ------------------------------------------------------------- helpers
    def _nslice(self, f):
        n = f.numel() // self.Nv
        if n * self.Nv != f.numel():
            raise ValueError('distribution size is not a multiple of Nv')
        return n

    def new_moments(self, nslice):
        return self.torch.zeros(nslice * self.nalph, dtype=self.torch.float64,
                                device=self._device)

    def _check_U(self, U, nslice):
        if U.dtype != self.torch.float64 or U.numel() != nslice * self.nalph:
            raise ValueError('moment array must be FP64 with nslice*14 entries')

    # ----------------------------------------------------- reference API
    def moment(self, t, f, U, stream=None):
        """U = f . mat (astd/scattering.py:155-162)."""
        ns = self._nslice(f)
        self._check_U(U, ns)
        nb = _lib.lib().dgfs_mom14_work_bytes(ns, self.Nv)
        if self._work is None or self._work.numel() * 8 < nb:
            self._work = self.torch.empty(max(1, nb // 8), dtype=self.torch.float64,
                                          device=self._device)
        _lib.check(_lib.lib().dgfs_mom14(
            self.dt_code, ns, self.Nv, _lib.ptr(f), _lib.ptr(self.cx), _lib.ptr(self.cy),
            _lib.ptr(self.cz), self.cw, _lib.ptr(self._work), _lib.ptr(U),
            _lib.stream_ptr(stream)))

    def constructMaxwellian(self, t, U, M, stream=None):
        """M = Shakov target of U (momentNorm + cmaxwellian)."""
        ns = self._nslice(M)
        self._check_U(U, ns)
        _lib.check(_lib.lib().dgfs_shakov_target(
            self.dt_code, ns, self.Nv, _lib.ptr(U), _lib.ptr(self.cx), _lib.ptr(self.cy),
            _lib.ptr(self.cz), float(self._Pr), _lib.ptr(M), _lib.stream_ptr(stream)))

    def collide(self, t, U, M, f, Q, stream=None):
        """Q = nu (M - f); Q may alias M or f."""
        ns = self._nslice(f)
        self._check_U(U, ns)
        _lib.check(_lib.lib().dgfs_relax_collide(
            self.dt_code, ns, self.Nv, _lib.ptr(U), _lib.ptr(M), _lib.ptr(f),
            float(self._prefactor), float(self._omega), _lib.ptr(Q), _lib.stream_ptr(stream)))

    def _lm(self, dt, A, B, C, s):
        a = _lib.LM()
        a.nstages, a.dt, a.prefac = int(s), float(dt), float(self._prefactor)
        a.omega, a.Pr = float(self._omega), float(self._Pr)
        a.srate = float(self.srate)
        for i in range(s):
            a.a[i], a.b[i] = float(A[i]), float(B[i])
        for i in range(s + 1):
            a.c[i] = float(C[i])
        return a

    def updateMomentLM(self, dt, A, B, C, LU, LM, U, nstages, stream=None):
        """U[nstages] from the histories (updateMom{s}_LM)."""
        s = int(nstages)
        a = self._lm(dt, A, B, C, s)
        ns = U[0].numel() // self.nalph
        for i in range(s):
            a.U[i], a.LU[i], a.LM[i] = U[i].data_ptr(), LU[i].data_ptr(), LM[i].data_ptr()
        a.Unew = U[s].data_ptr()
        _lib.check(_lib.lib().dgfs_imex_lm_update(self.dt_code, ns, self.Nv, _lib.C.byref(a), 0,
                                                  _lib.stream_ptr(stream)))

    def updateDistLM(self, dt, A, B, C, L, M, F, fnew, U, nstages, stream=None):
        """fnew from the histories; M[0..s-1] penalisations, M[s] target of U[s]."""
        s = int(nstages)
        a = self._lm(dt, A, B, C, s)
        ns = self._nslice(fnew)
        for i in range(s):
            a.F[i], a.L[i], a.Pn[i] = F[i].data_ptr(), L[i].data_ptr(), M[i].data_ptr()
        a.Mnew, a.fnew, a.Unew = M[s].data_ptr(), fnew.data_ptr(), U[s].data_ptr()
        _lib.check(_lib.lib().dgfs_imex_lm_update(self.dt_code, ns, self.Nv, _lib.C.byref(a), 1,
                                                  _lib.stream_ptr(stream)))

    def _ars(self, dt, ah, ai, s):
        a = _lib.ARS()
        a.stage, a.dt, a.prefac = int(s), float(dt), float(self._prefactor)
        a.omega, a.Pr = float(self._omega), float(self._Pr)
        a.srate = float(self.srate)
        for i in range(s):
            a.ah[i] = float(ah[i])
            a.ai[i + 1] = float(ai[i])
        return a

    def updateMomentARS(self, dt, ah, ai, LU, U, stream=None):
        """U[s] of stage s = len(ah): ah explicit weights of L_0..L_{s-1}, ai
        implicit weights of stages 1..s (updateMom{s}_ARS)."""
        s = len(ah)
        a = self._ars(dt, ah, ai, s)
        ns = U[0].numel() // self.nalph
        for i in range(s + 1):
            a.U[i] = U[i].data_ptr()
        for i in range(s):
            a.LU[i] = LU[i].data_ptr()
        _lib.check(_lib.lib().dgfs_imex_ars_update(self.dt_code, ns, self.Nv, _lib.C.byref(a),
                                                   0, _lib.stream_ptr(stream)))

    def updateDistARS(self, dt, ah, ai, L, U, M, F, stream=None):
        """F[s] of stage s; M[i-1] is the target of U[i] (i = 1..s)."""
        s = len(ah)
        a = self._ars(dt, ah, ai, s)
        ns = self._nslice(F[0])
        for i in range(s + 1):
            a.U[i], a.F[i] = U[i].data_ptr(), F[i].data_ptr()
        for i in range(s):
            a.L[i] = L[i].data_ptr()
            a.M[i + 1] = M[i].data_ptr()
        _lib.check(_lib.lib().dgfs_imex_ars_update(self.dt_code, ns, self.Nv, _lib.C.byref(a),
                                                   1, _lib.stream_ptr(stream)))


class DGFSBGKDirectGLLScatteringModelAstd(DGFSShakovDirectGLLScatteringModelAstd):
    """BGK penalisation P = nu (M - f) (astd/scattering.py:44-286): the
    Maxwellian of (rho, u, T) as target, nu = prefac rho T^(1-omega) with
    prefac = t0 p0 / mu(T0).  The reference evolves its five conserved
    moments; the 14-moment Shakov update with Pr = 1 gives the same
    distribution (its target then depends on rho, u and T only, and the
    relaxation leaves the conserved moments to the explicit part), so that
    update runs here.  A Pr in the deck is ignored, as in the reference."""
    scattering_model = 'bgk-direct-gll'

    def load_parameters(self):
        super().load_parameters()
        self._Pr = 1.0


class DGFSESBGKDirectGLLScatteringModelAstd(DGFSShakovDirectGLLScatteringModelAstd):
    """ES-BGK penalisation P = nu (G - f) (astd/scattering.py:287-354, kernels
    astd/kernels/scattering/esbgk-direct-gll.mako): G the Gaussian of
    Theta = T/Pr I + 2 (1 - 1/Pr) (P/rho - u u^T) (dgfs_esbgk_target),
    nu = prefac rho T^(1-omega) with prefac = t0 Pr p0 / mu(T0); the
    stresses relax at nu/Pr.  The reference evolves 11 moments; they are the
    first 11 of the 14 kept here (the heat flux is carried along, unused)."""
    scattering_model = 'esbgk-direct-gll'

    def load_parameters(self):
        super().load_parameters()
        self._prefactor *= self._Pr
        self.srate = 1.0 / self._Pr

    def constructMaxwellian(self, t, U, M, stream=None):
        """M = ES-BGK target of U (momentNorm + cmaxwellian)."""
        ns = self._nslice(M)
        self._check_U(U, ns)
        _lib.check(_lib.lib().dgfs_esbgk_target(
            self.dt_code, ns, self.Nv, _lib.ptr(U), _lib.ptr(self.cx), _lib.ptr(self.cy),
            _lib.ptr(self.cz), float(self._Pr), _lib.ptr(M), _lib.stream_ptr(stream)))
