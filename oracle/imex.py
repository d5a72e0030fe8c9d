"""Oracle: IMEX penalisation (14-moment Shakov) and IMEX integrators (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, per slice of a (nslice, n) array:
  moment()             dgfs1D/astd/scattering.py:155-162, mat rows :380-397
  momentNorm           astd/kernels/scattering/shakov-direct-gll.mako:193-230
  cmaxwellian          :232-264
  collide              :266-305
  updateMom{s}_ARS     :326-440      updateDistribution{s}_ARS   :443-484
  updateMom{s}_LM      :548-641      updateDistribution{s}_LM    :644-691
  updateMom{q}_BDF     :17-120       updateDistribution{q}_BDF   :123-159
  updateDistributionWeight{s}_SSPL   :505-541
and the integrators of dgfs1D/astd/integrator.py:
  ARS (a,b,c) stage loop             :48-95 with tables :100-178
  SSPL stage loop                    :182-237 with tables :240-278
  BDF-111/222/333                    :281-488
  IMEX linear multistep              :623-736 with tables :739-849
Documented deviation (see the product's integrator): the newest explicit
part of an s >= 2 multistep scheme subtracts the newest penalisation.
"""
import numpy as np

NALPH = 14


def mat14(cv, cw):
    c2 = (cv * cv).sum(0)
    return np.vstack([np.ones(cv.shape[1]), cv, c2,
                      cv[0] * cv[0], cv[1] * cv[1], cv[2] * cv[2],
                      cv[0] * cv[1], cv[1] * cv[2], cv[2] * cv[0],
                      c2 * cv[0], c2 * cv[1], c2 * cv[2]]) * cw


def moment(f, cv, cw):
    """U = f . mat^T for f of shape (nslice, n)."""
    return f @ mat14(cv, cw).T


def temperature(U):
    return (U[:, 4] - (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0]) / (1.5 * U[:, 0])


def momentnorm(U):
    rho = U[:, 0]
    u = U[:, 1:4] / rho[:, None]
    u2 = (u * u).sum(1)
    T = U[:, 4] / (1.5 * rho) - u2 / 1.5
    P = np.stack([np.stack([U[:, 5], U[:, 8], U[:, 10]], 1),
                  np.stack([U[:, 8], U[:, 6], U[:, 9]], 1),
                  np.stack([U[:, 10], U[:, 9], U[:, 7]], 1)], 1)        # (ns, 3, 3)
    Pu = np.einsum('sij,sj->si', P, u)
    q = U[:, 11:14] - 2.0 * Pu + (rho * u2)[:, None] * u - 1.5 * u * (rho * T)[:, None]
    return rho, u, T, q


def target(U, cv, Pr):
    rho, u, T, q = momentnorm(U)
    C = cv[None, :, :] - u[:, :, None]
    C2 = (C * C).sum(1)
    fe = (rho / (np.pi * T) ** 1.5)[:, None] * np.exp(-C2 / T[:, None])
    qc = (q[:, :, None] * C).sum(1)
    return fe * (1.0 + 2.0 * (1.0 - Pr) / 5.0 * qc / (rho * T * T)[:, None]
                 * (2.0 * C2 / T[:, None] - 5.0))


def nu_of(U, prefac, omega):
    return prefac * U[:, 0] * temperature(U) ** (1.0 - omega)


def collide(U, M, f, prefac, omega):
    return nu_of(U, prefac, omega)[:, None] * (M - f)


def _targets(U, T, Pr):
    r = U[:, 0]
    P = np.stack([0.5 * r * T + U[:, 1] ** 2 / r, 0.5 * r * T + U[:, 2] ** 2 / r,
                  0.5 * r * T + U[:, 3] ** 2 / r, U[:, 1] * U[:, 2] / r,
                  U[:, 2] * U[:, 3] / r, U[:, 3] * U[:, 1] / r], 1)
    m2 = (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / r ** 2
    rows = ((0, 3, 5), (3, 1, 4), (5, 4, 2))
    Q = []
    for i, (q1, q2, q3) in enumerate(rows):
        Pu = U[:, 5 + q1] * U[:, 1] + U[:, 5 + q2] * U[:, 2] + U[:, 5 + q3] * U[:, 3]
        ui = U[:, 1 + i]
        Q.append((1.0 - Pr) * (-2.0 * Pu / r + ui * m2 - 1.5 * ui * T) + (ui * m2 + 2.5 * ui * T))
    return P, np.stack(Q, 1)


def lm_moments(dt, A, B, C, LU, LM, U, s, prefac, omega, Pr, srate=1.0):
    """srate: the stresses relax at srate nu (1 Shakov, 1/Pr ES-BGK)."""
    E = sum(-A[s - 1 - i] * U[i] + dt * B[s - 1 - i] * LU[i] + dt * C[s - i] * LM[i]
            for i in range(s))
    T = temperature(E)
    w = C[0] * dt * prefac * E[:, 0] * T ** (1.0 - omega)
    Un = E.copy()
    P, _ = _targets(E, T, Pr)
    ws = srate * w[:, None]
    Un[:, 5:11] = (E[:, 5:11] + ws * P) / (1.0 + ws)
    _, Q = _targets(Un, T, Pr)
    Un[:, 11:14] = (E[:, 11:14] + w[:, None] * Q) / (1.0 + Pr * w[:, None])
    return Un


def lm_dist(dt, A, B, C, L, Pn, F, Mnew, Unew, s, prefac, omega):
    acc = sum(-A[s - 1 - i] * F[i] + dt * B[s - 1 - i] * L[i] + dt * C[s - i] * Pn[i]
              for i in range(s))
    cdn = (C[0] * dt * nu_of(Unew, prefac, omega))[:, None]
    return (acc + cdn * Mnew) / (1.0 + cdn)


def ars_moments(dt, ah, ai, LU, U, s, prefac, omega, Pr, srate=1.0):
    """U has s entries U_0..U_{s-1}; returns U_s (ai indexed 1..s)."""
    E = U[0] + sum(dt * ah[i] * LU[i] for i in range(s))
    for i in range(1, s):
        Ti = temperature(U[i])
        w = ai[i] * dt * prefac * U[i][:, 0] * Ti ** (1.0 - omega)
        P, Q = _targets(U[i], Ti, Pr)
        E[:, 5:11] += srate * w[:, None] * (P - U[i][:, 5:11])
        E[:, 11:14] += w[:, None] * (Q - Pr * U[i][:, 11:14])
    Ts = temperature(E)
    w = ai[s] * dt * prefac * E[:, 0] * Ts ** (1.0 - omega)
    Un = E.copy()
    P, _ = _targets(E, Ts, Pr)
    ws = srate * w[:, None]
    Un[:, 5:11] = (E[:, 5:11] + ws * P) / (1.0 + ws)
    _, Q = _targets(Un, Ts, Pr)
    Un[:, 11:14] = (E[:, 11:14] + w[:, None] * Q) / (1.0 + Pr * w[:, None])
    return Un


def ars_dist(dt, ah, ai, L, U, M, F, s, prefac, omega):
    """U: U_0..U_s, M: targets indexed 1..s (M[0] unused), F: f_0..f_{s-1}."""
    acc = F[0] + sum(ah[i] * dt * L[i] for i in range(s))
    for i in range(1, s):
        acc = acc + ai[i] * dt * nu_of(U[i], prefac, omega)[:, None] * (M[i] - F[i])
    w = (ai[s] * dt * nu_of(U[s], prefac, omega))[:, None]
    return (acc + w * M[s]) / (1.0 + w)


# ----------------------------------------------------------------- tableaux
ARS = {
    'ars-111': ([[0., 0.], [1., 0.]], [[0., 0.], [0., 1.]]),
    'ars-222': None,
    'ars-443': ([[0., 0., 0., 0., 0.], [1. / 2., 0., 0., 0., 0.],
                 [11. / 18., 1. / 18., 0., 0., 0.], [5. / 6., -5. / 6., 1. / 2., 0., 0.],
                 [1. / 4., 7. / 4., 3. / 4., -7. / 4., 0.]],
                [[0., 0., 0., 0., 0.], [0., 1. / 2., 0., 0., 0.], [0., 1. / 6., 1. / 2., 0., 0.],
                 [0., -1. / 2., 1. / 2., 1. / 2., 0.], [0., 3. / 2., -3. / 2., 1. / 2., 1. / 2.]]),
}
_g = 1.0 - 2.0 ** 0.5 / 2.0
_d = 1.0 - 1.0 / (2.0 * _g)
ARS['ars-222'] = ([[0., 0., 0.], [_g, 0., 0.], [_d, 1 - _d, 0.]],
                  [[0., 0., 0.], [0., _g, 0.], [0., 1 - _g, _g]])

LM = {
    'imex-lm-bdf1': ([-1.], [1.], [1., 0.]),
    'imex-lm-cn2': ([-1., 0.], [1.5, -0.5], [0.5, 0.5, 0.]),
    'imex-lm-ad2': ([-1., 0.], [1.5, -0.5], [9. / 16., 3. / 8., 1. / 16.]),
    'imex-lm-bdf2': ([-4. / 3., 1. / 3.], [4. / 3., -2. / 3.], [2. / 3., 0., 0.]),
    'imex-lm-sg2': ([-3. / 4., 0., -1. / 4.], [3. / 2., 0., 0.], [1., 0., 0., 1. / 2.]),
    'imex-lm-bdf3': ([-18. / 11., 9. / 11., -2. / 11.], [18. / 11., -18. / 11., 6. / 11.],
                     [6. / 11., 0., 0., 0.]),
    'imex-lm-ad3': ([-1., 0., 0.], [23. / 12., -4. / 3., 5. / 12.],
                    [4661. / 10000., 15551. / 30000., 1949. / 30000., -1483. / 30000.]),
    'imex-lm-tvb3': ([-3909. / 2048., 1367. / 1024., -873. / 2048.],
                     [18463. / 12288., -1271. / 768., 8233. / 12288.],
                     [1089. / 2048., -1139. / 12288., -367. / 6144., 1699. / 12288.]),
    'imex-lm-bdf4': ([-48. / 25., 36. / 25., -16. / 25., 3. / 25.],
                     [48. / 25., -72. / 25., 48. / 25., -12. / 25.],
                     [12. / 25., 0., 0., 0., 0.]),
}


# BDF-k in the penalisation form: A (q + 1), G (q), B (astd/integrator.py:285-
# 287, :334-336, :405-407); the reference passes g_i = -G_i to the kernels.
BDF = {
    'bdf-111': ([-1., 1.], [1.], 1.),
    'bdf-222': ([1. / 3., -4. / 3., 1.], [-2. / 3., 4. / 3.], 2. / 3.),
    'bdf-333': ([-2. / 11., 9. / 11., -18. / 11., 1.], [6. / 11., -18. / 11., 18. / 11.],
                6. / 11.),
}


def bdf_moments(dt, A, G, B, U, LU, prefac, omega, Pr, srate=1.0):
    """updateMom{q}_BDF: the conserved moments from the history, then the
    stresses and (with the new stresses) the heat flux relaxed implicitly
    with weight B dt nu(U_new)."""
    q = len(G)
    N = -sum(A[i] * U[i] - G[i] * dt * LU[i] for i in range(q))
    aq = A[q]
    Un = N / aq
    T = temperature(Un)
    w = (B * dt * prefac * Un[:, 0] * T ** (1.0 - omega))[:, None]
    P, _ = _targets(Un, T, Pr)
    Un[:, 5:11] = (N[:, 5:11] + srate * w * P) / (aq + srate * w)
    _, Q = _targets(Un, T, Pr)
    Un[:, 11:14] = (N[:, 11:14] + w * Q) / (aq + Pr * w)
    return Un


def bdf_dist(dt, A, G, B, F, L, Mn, Un, prefac, omega):
    """updateDistribution{q}_BDF."""
    q = len(G)
    N = -sum(A[i] * F[i] - G[i] * dt * L[i] for i in range(q))
    w = (B * dt * nu_of(Un, prefac, omega))[:, None]
    return (N + w * Mn) / (A[q] + w)


# SSPL tableaux (_A explicit, A implicit, _w, w) (astd/integrator.py:240-278)
_gs = 1. - (2. ** 0.5) / 2.
_al, _be, _et = 0.24169426078821, 0.06042356519705, 0.12915286960590
SSPL = {
    'sspl-222': ([[0., 0.], [1., 0.]], [[_gs, 0.], [1 - 2 * _gs, _gs]], [0.5, 0.5], [0.5, 0.5]),
    'sspl-443': ([[0., 0., 0., 0.], [0., 0., 0., 0.], [0., 1., 0., 0.],
                  [0., 1. / 4., 1. / 4., 0.]],
                 [[_al, 0., 0., 0.], [-_al, _al, 0., 0.], [0., 1. - _al, _al, 0.],
                  [_be, _et, 1. / 2. - _be - _et - _al, _al]],
                 [0., 1. / 6., 1. / 6., 2. / 3.], [0., 1. / 6., 1. / 6., 2. / 3.]),
}


def es_target(U, cv, Pr):
    """ES-BGK target of the moments (astd/kernels/scattering/esbgk-direct-gll.mako
    momentNorm :157-219, cmaxwellian :221-255): Theta = T/Pr I + 2 (1 - 1/Pr)
    (P/rho - u u^T), M = rho/(pi^1.5 sqrt det Theta) exp(-C^T Theta^-1 C);
    inverted here with numpy.linalg (the reference writes cofactors)."""
    rho = U[:, 0]
    u = U[:, 1:4] / rho[:, None]
    T = temperature(U)
    idx = [[5, 8, 10], [8, 6, 9], [10, 9, 7]]
    out = np.empty((U.shape[0], cv.shape[1]))
    for s in range(U.shape[0]):
        P = np.array([[U[s, idx[i][j]] for j in range(3)] for i in range(3)])
        Th = T[s] / Pr * np.eye(3) + 2.0 * (1.0 - 1.0 / Pr) * (P / rho[s] - np.outer(u[s], u[s]))
        C = cv - u[s][:, None]
        q = np.einsum('iv,ij,jv->v', C, np.linalg.inv(Th), C)
        out[s] = rho[s] / (np.pi ** 1.5 * np.sqrt(np.linalg.det(Th))) * np.exp(-q)
    return out


class Penal:
    """The penalisation model bound to a grid: prefac, omega, Pr, cv, cw;
    model 'shakov' (14 moments) or 'esbgk' (the ES-BGK target, stresses
    relaxing at nu/Pr, astd/scattering.py:287-354)."""

    def __init__(self, cv, cw, prefac, omega, Pr, model='shakov'):
        self.cv, self.cw, self.prefac, self.omega, self.Pr = cv, cw, prefac, omega, Pr
        self.model = model
        self.srate = 1.0 / Pr if model == 'esbgk' else 1.0

    def moment(self, f):
        return moment(f, self.cv, self.cw)

    def target(self, U):
        if self.model == 'esbgk':
            return es_target(U, self.cv, self.Pr)
        return target(U, self.cv, self.Pr)

    def penal(self, U, f):
        return collide(U, self.target(U), f, self.prefac, self.omega)


def ars_step(kind, pm, explicit, t, dt, f, penalize=True):
    """One ARS step (astd/integrator.py:48-95); explicit(t, f) -> L.  With the
    Boltzmann part present the reference overwrites it by the Maxwellian
    before use (:68-77), so the explicit part is explicit(f) - P(f)
    (penalize=True); penalize=False is the plain IMEX splitting
    explicit + implicit P (used by the order tests)."""
    _A, A = ARS[kind]
    ns = len(_A) - 1
    F = [f]
    U = [pm.moment(f)]
    L, LU, M = [], [], [None]
    for i in range(ns):
        Li = explicit(t, F[i]) - pm.penal(U[i], F[i]) if penalize else explicit(t, F[i])
        L.append(Li)
        LU.append(pm.moment(Li))
        ah, ai = _A[i + 1][0:i + 1], [0.0] + A[i + 1][1:i + 2]
        U.append(ars_moments(dt, ah, ai, LU, U, i + 1, pm.prefac, pm.omega, pm.Pr, pm.srate))
        M.append(pm.target(U[i + 1]))
        F.append(ars_dist(dt, ah, ai, L, U, M, F, i + 1, pm.prefac, pm.omega))
    return F[-1]


class LMIntegrator:
    """IMEX linear multistep (astd/integrator.py:623-736): history from
    ARS-443 substeps (10 per dt), then the s-step update."""
    nsub = 10

    def __init__(self, kind, pm, explicit, boltz, t, dt, f0, penalize_init=True):
        self.A, self.B, self.C = LM[kind]
        self.s = s = len(self.A)
        self.pm, self.explicit, self.boltz = pm, explicit, boltz
        fn, f, tl = [], f0.copy(), t
        for n in range(s):
            fn.append(f.copy())
            for _ in range(self.nsub):
                f = ars_step('ars-443', pm, self._ex_only, tl, dt / self.nsub, f,
                             penalize=penalize_init)
                tl += dt / self.nsub
        self.F, self.U, self.P, self.LM, self.L, self.LU = [], [], [], [], [], []
        for n in range(s):
            self._push(fn[n], t + n * dt)

    def _ex_only(self, t, f):
        return self.explicit(t, f)

    def _push(self, f, t):
        pm = self.pm
        U = pm.moment(f)
        P = pm.penal(U, f)
        L = self.explicit(t, f)
        if self.boltz is not None:
            L = L + self.boltz(t, f) - P
        self.F.append(f.copy())
        self.U.append(U)
        self.P.append(P)
        self.LM.append(pm.moment(P))
        self.L.append(L)
        self.LU.append(pm.moment(L))

    def step(self, t, dt, nacptsteps, f):
        s, pm = self.s, self.pm
        if nacptsteps < s - 1:
            return self.F[nacptsteps + 1].copy()
        Un = lm_moments(dt, self.A, self.B, self.C, self.LU, self.LM, self.U, s,
                        pm.prefac, pm.omega, pm.Pr, pm.srate)
        Mn = pm.target(Un)
        fn = lm_dist(dt, self.A, self.B, self.C, self.L, self.P, self.F, Mn, Un, s,
                     pm.prefac, pm.omega)
        for lst in (self.F, self.U, self.P, self.LM, self.L, self.LU):
            lst.pop(0)
        Pn = collide(Un, Mn, fn, pm.prefac, pm.omega)
        L = self.explicit(t, fn)
        if self.boltz is not None:
            L = L + self.boltz(t, fn) - Pn
        self.F.append(fn.copy())
        self.U.append(Un)
        self.P.append(Pn)
        self.LM.append(pm.moment(Pn))
        self.L.append(L)
        self.LU.append(pm.moment(L))
        return fn


def sspl_step(kind, pm, explicit, t, dt, f):
    """One SSPL step (astd/integrator.py:182-237): stage 0 implicit only, the
    explicit part of the stages is explicit(t, F_i), and
    f^{n+1} = f^n + dt sum_k _w_k L_k + dt sum_k w_k nu(U_{k+1}) (M_{k+1} - F_{k+1})
    (updateDistributionWeight_SSPL); M_k is the target of U_k."""
    _A, A, _w, w = SSPL[kind]
    ns = len(_A)
    zero, zeroU = np.zeros_like(f), np.zeros((f.shape[0], NALPH))
    F, U, M = [f], [pm.moment(f)], [None]
    L, LU = [], []
    U.append(ars_moments(dt, [0.0], [0.0, A[0][0]], [zeroU], U, 1, pm.prefac, pm.omega, pm.Pr,
                             pm.srate))
    M.append(pm.target(U[1]))
    F.append(ars_dist(dt, [0.0], [0.0, A[0][0]], [zero], U, M, F, 1, pm.prefac, pm.omega))
    for i in range(1, ns):
        L.append(explicit(t, F[i]))
        LU.append(pm.moment(L[-1]))
        ah, ai = _A[i][:i + 1], [0.0] + A[i][:i + 1]
        U.append(ars_moments(dt, ah, ai, LU + [zeroU], U, i + 1, pm.prefac, pm.omega, pm.Pr,
                             pm.srate))
        M.append(pm.target(U[i + 1]))
        F.append(ars_dist(dt, ah, ai, L + [zero], U, M, F, i + 1, pm.prefac, pm.omega))
    L.append(explicit(t, F[ns]))
    out = F[0].copy()
    for k in range(ns):
        out = out + _w[k] * dt * L[k] + w[k] * dt * nu_of(U[k + 1], pm.prefac, pm.omega)[:, None] \
            * (M[k + 1] - F[k + 1])
    return out


class BDFIntegrator:
    """bdf-111/222/333 (astd/integrator.py:281-488); explicit part
    explicit(t, f).  bdf-222/333 start with nsub ARS-222 substeps per dt (the
    reference's start calls ARS-222 without its tableau, :361-363, :431-433,
    and fails; this is the evident intent); bdf-333 advances two steps at its
    first call and none at its second, with the second substep loop at a
    fixed time (:436-447), as the reference."""
    nsub = 100

    def __init__(self, kind, pm, explicit, boltz=None):
        """boltz: with kind 'bdf-111', the boltz-ars-111 scheme
        (astd/integrator.py:491-550): the distribution's explicit part is
        explicit + boltz - P(f), its moments those of explicit alone."""
        self.A, self.G, self.B = BDF[kind]
        self.q = len(self.G)
        self.pm, self.explicit, self.boltz = pm, explicit, boltz
        self.hist = []

    def _rec(self, t, f):
        U = self.pm.moment(f)
        L = self.explicit(t, f)
        LU = self.pm.moment(L)
        if self.boltz is not None:
            L = L + self.boltz(t, f) - self.pm.penal(U, f)
        return f.copy(), L, U, LU

    def _ars(self, t, dt, f, advance=True):
        tl = t
        for _ in range(self.nsub):
            f = ars_step('ars-222', self.pm, self.explicit, tl, dt / self.nsub, f, penalize=False)
            if advance:
                tl += dt / self.nsub
        return f, tl

    def step(self, t, dt, nacptsteps, f):
        q, pm = self.q, self.pm
        if q >= 2 and nacptsteps == 0:
            fn = f.copy()
            f, tl = self._ars(t, dt, f)
            self.hist = [self._rec(t, fn)]
            if q == 3:
                fn = f.copy()
                f, _ = self._ars(tl, dt, f, advance=False)
                self.hist.append(self._rec(t, fn))
            return f
        if q == 3 and nacptsteps == 1:
            return f
        self.hist = (self.hist + [self._rec(t, f)])[-q:]
        F = [h[0] for h in self.hist]
        L = [h[1] for h in self.hist]
        U = [h[2] for h in self.hist]
        LU = [h[3] for h in self.hist]
        Un = bdf_moments(dt, self.A, self.G, self.B, U, LU, pm.prefac, pm.omega, pm.Pr, pm.srate)
        Mn = pm.target(Un)
        return bdf_dist(dt, self.A, self.G, self.B, F, L, Mn, Un, pm.prefac, pm.omega)


# --------------------------------------------- binary mixture (no reference)
# The reference's IMEX path is single-species; the mixture's species-wise
# Shakov penalisation is this build's (dgfs1d_gpu_amd/bi/penalty.py).  Here it
# is written in each species' physical variables: for species mass m and raw
# velocity moments of f on the common grid (cw),
#   n = sum f cw, u = sum f c cw / n, T = (2/3) m sum f |c-u|^2 cw / n,
#   q_c = sum f |c-u|^2 (c-u) cw,
#   S = n (m/(pi T))^1.5 exp(-m C^2/T) (1 + 4(1-Pr)/5 m^2 (q_c.C)/(n T^2) (m C^2/T - 5/2)),
#   nu = prefac n T^(1-omega).
# The linear-multistep and ARS moment updates act on the species' mass-scaled
# moment vector D_m U (the moments of the unit-mass problem in c' = sqrt(m) c),
# D_m = m^1.5 diag(1, m^.5 (x3), m, m (x6), m^1.5 (x3)).


def mass_scale(m):
    return m ** 1.5 * np.array([1.0] + [m ** 0.5] * 3 + [m] + [m] * 6 + [m ** 1.5] * 3)


def species_target(U, cv, m, Pr):
    """Shakov target of species mass m from its mass-scaled moments U."""
    Ur = U / mass_scale(m)[None, :]                     # raw moments on the grid
    n = Ur[:, 0]
    u = Ur[:, 1:4] / n[:, None]
    C = cv[None, :, :] - u[:, :, None]
    C2 = (C * C).sum(1)
    T = 2.0 / 3.0 * m * (Ur[:, 4] - n * (u * u).sum(1)) / n
    qdot = (_central3(Ur, n, u)[:, :, None] * C).sum(1)
    fm = (n * (m / (np.pi * T)) ** 1.5)[:, None] * np.exp(-m * C2 / T[:, None])
    return fm * (1.0 + 0.8 * (1.0 - Pr) * (m * m * qdot / (n * T * T)[:, None])
                 * (m * C2 / T[:, None] - 2.5))


def _central3(Ur, n, u):
    """sum f |c-u|^2 (c-u) cw from the raw moments (R3 = sum f |c|^2 c cw,
    P = sum f c c^T cw, tr = sum f |c|^2 cw, j = sum f c cw), by the expansion
    |c-u|^2 (c-u) = |c|^2 c - 2 (c.u) c - |c|^2 u + 2 (c.u) u + |u|^2 c - |u|^2 u."""
    P = np.stack([np.stack([Ur[:, 5], Ur[:, 8], Ur[:, 10]], 1),
                  np.stack([Ur[:, 8], Ur[:, 6], Ur[:, 9]], 1),
                  np.stack([Ur[:, 10], Ur[:, 9], Ur[:, 7]], 1)], 1)
    j = Ur[:, 1:4]
    tr = Ur[:, 4]
    u2 = (u * u).sum(1)
    ju = (j * u).sum(1)
    return (Ur[:, 11:14] - 2.0 * np.einsum('sij,sj->si', P, u) - tr[:, None] * u
            + 2.0 * ju[:, None] * u + u2[:, None] * j - (u2 * n)[:, None] * u)


class SpeciesPenal:
    """Species penalisation: mass m, species prefactor (nu = prefac n T^(1-omega)),
    omega, Pr on the common grid (cv, cw)."""

    def __init__(self, cv, cw, m, prefac, omega, Pr):
        self.cv, self.cw, self.m, self.omega, self.Pr = cv, cw, m, omega, Pr
        self.species_prefac = prefac
        self.prefac = prefac * m ** -1.5          # acting on the scaled density
        self.srate = 1.0

    def moment(self, f):
        return moment(f, self.cv, self.cw) * mass_scale(self.m)[None, :]

    def target(self, U):
        return species_target(U, self.cv, self.m, self.Pr)

    def penal(self, U, f):
        return collide(U, self.target(U), f, self.prefac, self.omega)


class MixPenal:
    """Two SpeciesPenals over a species-major stack of 2 nslice slices; the
    update functions see per-slice prefac and omega arrays."""

    def __init__(self, species, nslice):
        self.sp, self.ns = species, nslice
        self.prefac = np.concatenate([np.full(nslice, s.prefac) for s in species])
        self.omega = np.concatenate([np.full(nslice, s.omega) for s in species])
        if len({s.Pr for s in species}) != 1:
            raise ValueError('one Prandtl number for both species')
        self.Pr = species[0].Pr
        self.srate = 1.0

    def _split(self, x):
        return [x[p * self.ns:(p + 1) * self.ns] for p in range(len(self.sp))]

    def moment(self, f):
        return np.concatenate([s.moment(x) for s, x in zip(self.sp, self._split(f))])

    def target(self, U):
        return np.concatenate([s.target(x) for s, x in zip(self.sp, self._split(U))])

    def penal(self, U, f):
        return collide(U, self.target(U), f, self.prefac, self.omega)


KB = 1.38064852e-23


def vss_viscosity(m, alpha, omega, dRef, Tref, T):
    """Variable-soft-sphere viscosity (Bird, Molecular Gas Dynamics, 1994,
    eq. (3.74)): mu_ref = 5 (alpha+1)(alpha+2) sqrt(m k Tref / pi) /
    (4 alpha (5 - 2 omega)(7 - 2 omega) dRef^2), mu = mu_ref (T/Tref)^omega."""
    mu_ref = 5.0 * (alpha + 1.0) * (alpha + 2.0) * np.sqrt(m * KB * Tref / np.pi) / (
        4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega) * dRef ** 2)
    return mu_ref * (T / Tref) ** omega


def species_prefactor(H0, u0, n0, T0, molar_mass, NA, alpha, omega, dRef, Tref):
    """Shakov frequency prefactor of a pure species in the solver's units:
    nu = prefac n T^(1-omega) with prefac = t0 n0 k T0 / mu(T0), t0 = H0/u0."""
    return H0 / u0 * n0 * KB * T0 / vss_viscosity(molar_mass / NA, alpha, omega, dRef, Tref, T0)
