"""IMEX integrators of the asymptotic-preserving path (reference
dgfs1D/astd/integrator.py).

Registry key ``intg_kind``; constructor ``(explicit, psm, shape, dtype, t=,
dt=, f0=, explicitQ=)`` and ``integrate(t, dt, nacptsteps, d_ucoeff)`` as the
reference (astd/abstd.py:465-474).  ``explicit(t, f, out)`` is the transport
operator, ``explicitQ(t, f, out)`` the Boltzmann collision operator and
``psm`` the penalisation model (moments, Shakov target, P = nu (M - f), stage
updates).

* ARS IMEX Runge-Kutta (ars-111, ars-222, ars-443; :48-178).  When a
  Boltzmann operator is present the reference evaluates it and then
  overwrites its register with the Maxwellian before use (:68-77): the
  explicit part of every stage is transport - P(f).  That is kept here,
  without the discarded evaluation.
* IMEX linear multistep (imex-lm-bdf1..4, -cn2, -ad2, -sg2, -ad3, -tvb3;
  :623-849): the s-1 starting values come from ARS-443 with 10 substeps per
  dt (:649-668); explicit parts are transport + Q - P.  Deviation: the
  reference subtracts the penalisation register M[0] from the newest explicit
  part (:727), which holds P(f^{n+1}) only for s = 1; here the newest
  explicit part always subtracts P(f^{n+1}).  The reference also stops after
  its second step (:731); this implementation continues.
* SSPL IMEX Runge-Kutta (sspl-222, sspl-443; :182-278): stage 0 implicit
  only; explicit part = transport; the step ends with the weighted
  combination f^{n+1} = f^n + dt sum_k _w_k L_k + dt sum_k w_k P_k
  (updateDistributionWeight_SSPL), evaluated as collide + axnpby.
* boltz-ars-111 (:491-550): BDF-111 with the explicit part transport + Q - P
  for the distribution (the moments use the transport part alone).
* BDF-k (bdf-111, bdf-222, bdf-333; :281-488): explicit part = transport.
  The update is the linear-multistep one with a = A/A_q, b = G/A_q (newest
  first) and c = (B/A_q, 0, ...) (no penalisation terms), so the same
  kernels evaluate it.  bdf-222/333 start with ARS-222 substeps (nstepsInit
  per dt); the reference's start calls ARS-222's stage loop without its
  tableau arguments (:361-363, :431-433) and cannot run, so the evident intent
  is implemented.  bdf-333 advances two steps at its first call and none at
  its second, its second substep loop at a fixed time (:436-447), as the
  reference.
"""
import numpy as np

from ..transport import axnpby


class DGFSIntegratorAstd:
    intg_kind = None

    def __init__(self, explicit, sm, shape, dtype, device=None, **kwargs):
        import torch
        self.torch = torch
        self._explicit = explicit
        self._sm = sm
        self._shape = shape
        self._dtype = dtype
        self._device = device
        self._explicitQ = kwargs.get('explicitQ')
        K, Ne, Nv = shape
        self._n = K * Ne * Nv
        self._nslice = K * Ne

    def _dist(self):
        return self.torch.zeros(self._n, dtype=self._dtype, device=self._device)

    def _mom(self):
        return self._sm.new_moments(self._nslice)

    def integrate(self, t, dt, nacptsteps, d_ucoeff):
        raise NotImplementedError


class ARSabcIntegratorAstd(DGFSIntegratorAstd):
    _A = A = None

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        ns = len(self._A) - 1
        self._ns = ns
        self.L = [self._dist() for _ in range(ns)]
        self.M = [self._dist() for _ in range(ns)]
        self.F = [self._dist() for _ in range(ns - 1)]
        self.U = [self._mom() for _ in range(ns + 1)]
        self.LU = [self._mom() for _ in range(ns)]

    def integrate(self, t, dt, nacptsteps, d_ucoeff):
        sm, ns = self._sm, self._ns
        L, M, U, LU = self.L, self.M, self.U, self.LU
        F = [d_ucoeff] + self.F + [d_ucoeff]
        sm.moment(t, F[0], U[0])
        for i in range(ns):
            self._explicit(t, F[i], L[i])
            if self._explicitQ is not None:
                sm.constructMaxwellian(t, U[i], M[i])
                sm.collide(t, U[i], M[i], F[i], M[i])
                axnpby(L[i], [(-1.0, M[i])], 1.0)
            sm.moment(t, L[i], LU[i])
            ah, ai = self._A[i + 1][0:i + 1], self.A[i + 1][1:i + 2]
            sm.updateMomentARS(dt, ah, ai, LU[:i + 1], U[:i + 2])
            sm.constructMaxwellian(t, U[i + 1], M[i])
            sm.updateDistARS(dt, ah, ai, L[:i + 1], U[:i + 2], M[:i + 1], F[:i + 2])


class ARS111IntegratorAstd(ARSabcIntegratorAstd):
    intg_kind = 'ars-111'
    _A = [[0., 0.], [1., 0.]]
    A = [[0., 0.], [0., 1.]]


class ARS222IntegratorAstd(ARSabcIntegratorAstd):
    intg_kind = 'ars-222'
    gamma = 1. - (2. ** 0.5) / 2.
    delta = 1. - 1. / (2. * gamma)
    _A = [[0., 0., 0.], [gamma, 0., 0.], [delta, 1 - delta, 0.]]
    A = [[0., 0., 0.], [0., gamma, 0.], [0., 1 - gamma, gamma]]


class ARS443IntegratorAstd(ARSabcIntegratorAstd):
    intg_kind = 'ars-443'
    _A = [[0., 0., 0., 0., 0.], [1. / 2., 0., 0., 0., 0.], [11. / 18., 1. / 18., 0., 0., 0.],
          [5. / 6., -5. / 6., 1. / 2., 0., 0.], [1. / 4., 7. / 4., 3. / 4., -7. / 4., 0.]]
    A = [[0., 0., 0., 0., 0.], [0., 1. / 2., 0., 0., 0.], [0., 1. / 6., 1. / 2., 0., 0.],
         [0., -1. / 2., 1. / 2., 1. / 2., 0.], [0., 3. / 2., -3. / 2., 1. / 2., 1. / 2.]]


class SSPLabcIntegratorAstd(DGFSIntegratorAstd):
    _A = A = _w = w = None

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        ns = self._ns = len(self._A)
        self.L = [self._dist() for _ in range(ns)]
        self.M = [self._dist() for _ in range(ns)]
        self.F = [self._dist() for _ in range(ns)]
        self.U = [self._mom() for _ in range(ns + 1)]
        self.LU = [self._mom() for _ in range(ns)]
        self.P = self._dist()
        self.dL, self.dLU = self._dist(), self._mom()      # weight-0 stand-ins

    def integrate(self, t, dt, nacptsteps, d_ucoeff):
        sm, ns = self._sm, self._ns
        L, M, U, LU = self.L, self.M, self.U, self.LU
        F = [d_ucoeff] + self.F
        sm.moment(t, F[0], U[0])
        sm.updateMomentARS(dt, [0.0], [self.A[0][0]], [self.dLU], U[:2])
        sm.constructMaxwellian(t, U[1], M[0])
        sm.updateDistARS(dt, [0.0], [self.A[0][0]], [self.dL], U[:2], M[:1], F[:2])
        for i in range(1, ns):
            self._explicit(t, F[i], L[i - 1])
            sm.moment(t, L[i - 1], LU[i - 1])
            ah, ai = self._A[i][:i + 1], self.A[i][:i + 1]
            sm.updateMomentARS(dt, ah, ai, LU[:i] + [self.dLU], U[:i + 2])
            sm.constructMaxwellian(t, U[i + 1], M[i])
            sm.updateDistARS(dt, ah, ai, L[:i] + [self.dL], U[:i + 2], M[:i + 1], F[:i + 2])
        self._explicit(t, F[ns], L[ns - 1])
        # f^{n+1} = f^n + dt sum_k (_w_k L_k + w_k nu(U_{k+1}) (M_k - F_{k+1}))
        terms = [(self._w[k] * dt, L[k]) for k in range(ns) if self._w[k] != 0.0]
        for i in range(0, len(terms), 3):
            axnpby(d_ucoeff, terms[i:i + 3], 1.0)
        for k in range(ns):
            if self.w[k] != 0.0:
                sm.collide(t, U[k + 1], M[k], F[k + 1], self.P)
                axnpby(d_ucoeff, [(self.w[k] * dt, self.P)], 1.0)


class SSPL222IntegratorAstd(SSPLabcIntegratorAstd):
    intg_kind = 'sspl-222'
    gamma = 1. - (2. ** 0.5) / 2.
    _A = [[0., 0.], [1., 0.]]
    A = [[gamma, 0.], [1 - 2 * gamma, gamma]]
    _w = [0.5, 0.5]
    w = [0.5, 0.5]


class SSPL443IntegratorAstd(SSPLabcIntegratorAstd):
    intg_kind = 'sspl-443'
    alpha, beta, eta = 0.24169426078821, 0.06042356519705, 0.12915286960590
    _A = [[0., 0., 0., 0.], [0., 0., 0., 0.], [0., 1., 0., 0.], [0., 1. / 4., 1. / 4., 0.]]
    A = [[alpha, 0., 0., 0.], [-alpha, alpha, 0., 0.], [0., 1. - alpha, alpha, 0.],
         [beta, eta, 1. / 2. - beta - eta - alpha, alpha]]
    _w = [0., 1. / 6., 1. / 6., 2. / 3.]
    w = [0., 1. / 6., 1. / 6., 2. / 3.]


class BDFabcIntegratorAstd(DGFSIntegratorAstd):
    A = G = B = None
    nstepsInit = 100

    def __init__(self, explicit, sm, shape, dtype, device=None, **kwargs):
        super().__init__(explicit, sm, shape, dtype, device=device, **kwargs)
        q = self._q = len(self.G)
        Aq = self.A[q]
        self._a = [self.A[q - 1 - k] / Aq for k in range(q)]
        self._b = [self.G[q - 1 - k] / Aq for k in range(q)]
        self._c = [self.B / Aq] + [0.0] * q
        # history, oldest first; U[q] is the new moments, M the new target
        self.F = [self._dist() for _ in range(q)]
        self.L = [self._dist() for _ in range(q)]
        self.U = [self._mom() for _ in range(q + 1)]
        self.LU = [self._mom() for _ in range(q)]
        self.Mn = self._dist()
        self._ars = None
        if q > 1:
            kw = dict(kwargs)
            kw['explicitQ'] = None
            self._ars = ARS222IntegratorAstd(explicit, sm, shape, dtype, device=device, **kw)

    def _record(self, t, f, k):
        self._sm.moment(t, f, self.U[k])
        self._explicit(t, f, self.L[k])
        self._sm.moment(t, self.L[k], self.LU[k])
        self.F[k].copy_(f)

    def _substeps(self, t, dt, f, advance=True):
        tloc, dtloc = t, dt / self.nstepsInit
        for step in range(self.nstepsInit):
            self._ars.integrate(tloc, dtloc, step, f)
            if advance:
                tloc += dtloc
        return tloc

    def _roll(self):
        q = self._q
        for v in (self.F, self.L, self.LU):
            v.append(v.pop(0))
        self.U[:q] = self.U[1:q] + self.U[:1]

    def integrate(self, t, dt, nacptsteps, d_ucoeff):
        sm, q = self._sm, self._q
        if q >= 2 and nacptsteps == 0:
            fn = d_ucoeff.clone()
            tloc = self._substeps(t, dt, d_ucoeff)
            self._record(t, fn, q - 1)
            if q == 3:
                fn.copy_(d_ucoeff)
                self._substeps(tloc, dt, d_ucoeff, advance=False)
                self._roll()
                self._record(t, fn, q - 1)
            return
        if q == 3 and nacptsteps == 1:
            return
        self._roll()
        self._record(t, d_ucoeff, q - 1)
        # LU stands in for the (zero-weight) penalisation moments, F for the
        # (zero-weight) penalisation registers
        sm.updateMomentLM(dt, self._a, self._b, self._c, self.LU, self.LU, self.U, q)
        sm.constructMaxwellian(t, self.U[q], self.Mn)
        sm.updateDistLM(dt, self._a, self._b, self._c, self.L, self.F + [self.Mn], self.F,
                        d_ucoeff, self.U, q)


class Bdf111IntegratorAstd(BDFabcIntegratorAstd):
    intg_kind = 'bdf-111'
    A, G, B = [-1., 1.], [1.], 1.


class Bdf222IntegratorAstd(BDFabcIntegratorAstd):
    intg_kind = 'bdf-222'
    A, G, B = [1. / 3., -4. / 3., 1.], [-2. / 3., 4. / 3.], 2. / 3.


class Bdf333IntegratorAstd(BDFabcIntegratorAstd):
    intg_kind = 'bdf-333'
    A, G, B = [-2. / 11., 9. / 11., -18. / 11., 1.], [6. / 11., -18. / 11., 18. / 11.], 6. / 11.


class BoltzARS111IntegratorAstd(Bdf111IntegratorAstd):
    """boltz-ars-111 (astd/integrator.py:491-550): BDF-111 whose explicit
    part for the distribution is transport + Q(f) - P(f); the moment update
    uses the moments of the transport part alone, taken before Q - P is
    added (:531-537), as the reference."""
    intg_kind = 'boltz-ars-111'

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        if self._explicitQ is None:
            raise ValueError('boltz-ars-111 needs the Boltzmann operator (explicitQ)')
        self._Q, self._M0 = self._dist(), self._dist()

    def _record(self, t, f, k):
        super()._record(t, f, k)
        sm = self._sm
        sm.constructMaxwellian(t, self.U[k], self._M0)
        sm.collide(t, self.U[k], self._M0, f, self._M0)          # P(f)
        self._explicitQ(t, f, self._Q)
        axnpby(self.L[k], [(1.0, self._Q), (-1.0, self._M0)], 1.0)


class IMEXLMsIntegratorAstd(DGFSIntegratorAstd):
    """s-step IMEX linear multistep scheme with Shakov penalisation."""
    A = B = C = None
    nstepsInit = 10

    def __init__(self, explicit, sm, shape, dtype, device=None, **kwargs):
        super().__init__(explicit, sm, shape, dtype, device=device, **kwargs)
        s = self._nstages = len(self.A)
        t, dt, f0 = kwargs.get('t'), kwargs.get('dt'), kwargs.get('f0')
        # history registers, oldest first; M has one extra slot (the new target)
        self.F = [self._dist() for _ in range(s)]
        self.L = [self._dist() for _ in range(s)]
        self.M = [self._dist() for _ in range(s + 1)]
        self.U = [self._mom() for _ in range(s + 1)]
        self.LU = [self._mom() for _ in range(s)]
        self.LM = [self._mom() for _ in range(s)]
        self._tmp = self._dist()
        # starting values: ARS-443 with nstepsInit substeps per dt
        ars = ARS443IntegratorAstd(explicit, sm, shape, dtype, device=device, **kwargs)
        f = f0.clone()
        tloc, dtloc = t, dt / self.nstepsInit
        fn = []
        for _ in range(s):
            fn.append(f.clone())
            for step in range(self.nstepsInit):
                ars.integrate(tloc, dtloc, step, f)
                tloc += dtloc
        del ars
        for n in range(s):
            self._updater(fn[n], n, t + n * dt)

    def _explicit_part(self, t, f, L, P):
        """L = transport(f) [+ Q(f) - P]."""
        self._explicit(t, f, L)
        if self._explicitQ is not None:
            self._explicitQ(t, f, self._tmp)
            axnpby(L, [(1.0, self._tmp), (-1.0, P)], 1.0)

    def _updater(self, dist, idx, t):
        sm = self._sm
        sm.moment(t, dist, self.U[idx])
        sm.constructMaxwellian(t, self.U[idx], self.M[idx])
        sm.collide(t, self.U[idx], self.M[idx], dist, self.M[idx])
        sm.moment(t, self.M[idx], self.LM[idx])
        self._explicit_part(t, dist, self.L[idx], self.M[idx])
        sm.moment(t, self.L[idx], self.LU[idx])
        self.F[idx].copy_(dist)

    @staticmethod
    def _roll(v):
        v.append(v.pop(0))

    def integrate(self, t, dt, nacptsteps, d_ucoeff):
        sm, s = self._sm, self._nstages
        if nacptsteps < s - 1:
            d_ucoeff.copy_(self.F[nacptsteps + 1])
            return
        sm.updateMomentLM(dt, self.A, self.B, self.C, self.LU, self.LM, self.U, s)
        sm.constructMaxwellian(t, self.U[s], self.M[s])
        sm.updateDistLM(dt, self.A, self.B, self.C, self.L, self.M, self.F, d_ucoeff, self.U, s)
        sm.collide(t, self.U[s], self.M[s], d_ucoeff, self.M[s])     # P(f^{n+1})
        sm.moment(t, self.M[s], self.LM[0])
        for v in (self.M, self.U, self.LM, self.F, self.L, self.LU):
            self._roll(v)
        # now M[s-1] = P(f^{n+1}), U[s-1] = U^{n+1}, LM[s-1] = its moments
        self._explicit_part(t, d_ucoeff, self.L[s - 1], self.M[s - 1])
        sm.moment(t, self.L[s - 1], self.LU[s - 1])
        self.F[s - 1].copy_(d_ucoeff)


class IMEXLM_BDF1_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-bdf1'
    A, B, C = [-1.], [1.], [1., 0.]


class IMEXLM_CN2_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-cn2'
    A, B, C = [-1., 0.], [1.5, -0.5], [0.5, 0.5, 0.]


class IMEXLM_AD2_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-ad2'
    A, B, C = [-1., 0.], [1.5, -0.5], [9. / 16., 3. / 8., 1. / 16.]


class IMEXLM_BDF2_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-bdf2'
    A, B, C = [-4. / 3., 1. / 3.], [4. / 3., -2. / 3.], [2. / 3., 0., 0.]


class IMEXLM_SG2_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-sg2'
    A, B, C = [-3. / 4., 0., -1. / 4.], [3. / 2., 0., 0.], [1., 0., 0., 1. / 2.]


class IMEXLM_BDF3_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-bdf3'
    A = [-18. / 11., 9. / 11., -2. / 11.]
    B = [18. / 11., -18. / 11., 6. / 11.]
    C = [6. / 11., 0., 0., 0.]


class IMEXLM_AD3_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-ad3'
    A = [-1., 0., 0.]
    B = [23. / 12., -4. / 3., 5. / 12.]
    C = [4661. / 10000., 15551. / 30000., 1949. / 30000., -1483. / 30000.]


class IMEXLM_TVB3_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-tvb3'
    A = [-3909. / 2048., 1367. / 1024., -873. / 2048.]
    B = [18463. / 12288., -1271. / 768., 8233. / 12288.]
    C = [1089. / 2048., -1139. / 12288., -367. / 6144., 1699. / 12288.]


class IMEXLM_BDF4_IntegratorAstd(IMEXLMsIntegratorAstd):
    intg_kind = 'imex-lm-bdf4'
    A = [-48. / 25., 36. / 25., -16. / 25., 3. / 25.]
    B = [48. / 25., -72. / 25., 48. / 25., -12. / 25.]
    C = [12. / 25., 0., 0., 0., 0.]


__all__ = [n for n in dir() if n.endswith('IntegratorAstd')] + ['np']
