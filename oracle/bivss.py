"""Oracle: binary-mixture VSS fast-spectral collision operator (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, step for step, the reference pipeline for one species pair (p, q):
  * velocity mesh of the mixture: L, S, R from the largest mass ratio, full
    spherical design with sw = 4 pi / M (bi/velocitymesh.py:28-119)
  * prefactor_pq and gamma_pq = 2 (1 - omega_pq), eta_pq = alpha_pq - 1
    (bi/scattering.py:54-93)
  * a[x,y](l)     = (pi/L) qz_x (sz_y . l)                    (vss.mako:29-48)
  * b_pq[x,y](l)  = sum_pre (1 + sz_y . szpre)^eta fac_b(x)
                        exp(-i m_q/(m_p+m_q) (pi/L) qz_x szpre . l)
    c_pq(l)       = sum_{x,y,pre} (1 + sz_y . szpre)^eta fac_c(x)
                        exp(-i (pi/L) qz_x szpre . l)
    fac_b = swpre qz_x^(gamma+2), fac_c = qw_x sw fac_b      (vss.mako:50-93,
                                                              bi/scattering.py:193-206)
  * fs (bi/scattering.py:261-321): F = FFT(f_p)/n, G = FFT(f_q)/n,
    t1 = exp(+i m_q/(m_p+m_q) a) F, t2 = exp(-i m_p/(m_p+m_q) a) G
    (cosSinMul, vss.mako:117-150), gain = IFFT_u(sum_xy qw_x sw/n b_pq
    FFT(IFFT_u(t1) IFFT_u(t2))) (prod/computeQG, :152-230), loss =
    IFFT_u(c_pq G) (ax2, :232-250), Q += prefac (Re gain - f_p Re loss)
    (output_pq, :252-285).
cuFFT conventions as in oracle/collision.py: forward e^{-i}, inverse e^{+i},
both unnormalised.

Pinned by tests/test_oracle_bivss.py against the physics the operator must
keep: Q_pq vanishes on a common Maxwellian of two unequal masses; the pair
conserves momentum and energy; the equal-mass VHS case (alpha = 1) converges
to the single-species operator of oracle/collision.py as the pre-rule grows.
"""
from math import gamma as _gamma

import numpy as np

from .collision import _fft, _ifft_u, lvec
from .velocitymesh import NA, R0, load_design
from . import quadratures as Q


class BiVelocityMesh:
    """Mixture velocity mesh (bi/velocitymesh.py:28-119).  ``masses`` are the
    molar masses divided by molarMass0 (:50-53)."""

    def __init__(self, H0, T0, rho0, molarMass0, Nv, Nrho, cmax, Tmax, dev,
                 molar_masses, M, ssrule='womersley', quad='jacobi', qa=0.0, qb=0.0,
                 sz_override=None):
        self.H0, self.T0, self.rho0, self.molarMass0 = H0, T0, rho0, molarMass0
        self.n0 = rho0 / molarMass0 * NA                              # :33
        self.u0 = np.sqrt(2.0 * R0 / molarMass0 * T0)                 # :34
        self.Nv, self.vsize, self.Nrho = Nv, Nv ** 3, Nrho
        cmax = cmax / self.u0
        Tmax = Tmax / T0
        self.masses = [m / molarMass0 for m in molar_masses]          # :50-53
        mr = max(self.masses)
        if mr > 4:
            raise ValueError('Mass ratio is too high!')               # :56
        self.L = cmax + dev * np.sqrt(Tmax)                           # :59
        self.S = min(                                                 # :60-63
            self.L * 2.0 / (max(4.0 * mr / (1.0 + mr), 2.0) + np.sqrt(1.0 + mr) + 1.0),
            self.L * 2.0 / (max(4.0 / (1.0 + mr), 2.0) + np.sqrt(1.0 + 1.0 / mr) + 1.0))
        self.R = 2.0 * self.S
        self.cw = (2.0 * self.L / Nv) ** 3
        c0 = np.linspace(-self.L + self.L / Nv, self.L - self.L / Nv, Nv)
        I, J, K = np.meshgrid(np.arange(Nv), np.arange(Nv), np.arange(Nv), indexing='ij')
        self.cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[K.ravel()]])
        rules = {'jacobi': Q.zwgj, 'lobatto': Q.zwglj,
                 'radau-m': Q.zwgrjm, 'radau-p': Q.zwgrjp}
        z, w = rules[quad](Nrho, qa, qb)                              # :101-103
        self.qz = (self.R / 2.0) * (1.0 + z)
        self.qw = (self.R / 2.0) * w
        self.M = M
        if sz_override is not None:
            self.sz = np.asarray(sz_override, dtype=np.float64)
        else:
            self.sz = load_design(ssrule, M)                          # :111-117 (full sphere)
        self.sw = 4.0 * np.pi / M                                     # :118


def vss_prefactor(masses, p, q, n0, H0, T0, alpha, omega, dRef, Tref):
    """invKn_pq and prefactor_pq (bi/scattering.py:78-90)."""
    mp, mq = masses[p], masses[q]
    invKn = (np.sqrt(1.0 + mp / mq) * np.pi * n0 * dRef * dRef
             * (Tref / T0) ** (omega - 0.5) * H0)
    mu = mp * mq / (mp + mq)
    pref = invKn * alpha / (np.sqrt(1.0 + mp / mq) * mu ** (omega - 0.5)
                            * 2.0 ** (1.0 + alpha) * _gamma(2.5 - omega) * np.pi)
    return pref, invKn


class BiPrecomp:
    """a, b_pq, c_pq of one species pair (bi/scattering.py:95-206,
    vss.mako:29-93)."""

    def __init__(self, N, L, qz, qw, sz, sw, szpre, mp, mq, alpha, omega):
        self.N, self.n = N, N ** 3
        self.L, self.qz, self.qw = L, np.asarray(qz), np.asarray(qw)
        self.sz, self.sw = np.asarray(sz), sw
        self.szpre = np.asarray(szpre)
        self.Mpre = len(self.szpre)
        swpre = 4.0 * np.pi / self.Mpre                               # scattering.py:62
        self.M, self.Nrho = self.sz.shape[0], len(self.qz)
        self.mp, self.mq = mp, mq
        gam, eta = 2.0 * (1.0 - omega), alpha - 1.0                   # :71, :74
        l = lvec(N)
        self.l = l
        fac = np.pi / L * self.qz                                     # vss.mako:40
        self.a = fac[:, None, None] * (self.sz @ l)[None]             # (Nrho, M, n)
        B = np.maximum(1.0 + self.sz @ self.szpre.T, 0.0) ** eta      # (M, Mpre)
        spl = self.szpre @ l                                          # (Mpre, n)
        mpq = mq / (mp + mq)                                          # vss.mako:54
        fac_b = swpre * self.qz ** (gam + 2.0)                        # scattering.py:197
        fac_c = self.qw * sw * fac_b                                  # :198
        self.b = np.empty((self.Nrho, self.M, self.n), dtype=np.complex128)
        self.c = np.zeros(self.n, dtype=np.complex128)
        for x in range(self.Nrho):
            eb = np.exp(-1j * mpq * fac[x] * spl)                     # (Mpre, n)
            ec = np.exp(-1j * fac[x] * spl)
            self.b[x] = fac_b[x] * (B @ eb)                           # vss.mako:78-80
            self.c += fac_c[x] * (B @ ec).sum(0)                      # :81-82, :88-89


def q_pq(pre, f, g, prefac):
    """One node: the increment prefac (gain - f loss) added to Q_p by the
    partner species q (bi/scattering.py:261-321)."""
    N, n = pre.N, pre.n
    F = _fft(np.asarray(f, dtype=np.complex128), N) / n               # cosSinMul scale
    G = _fft(np.asarray(g, dtype=np.complex128), N) / n
    a = pre.a.reshape(-1, n)
    m = pre.mp + pre.mq
    t1 = np.exp(1j * a * (pre.mq / m)) * F                            # vss.mako:137-140
    t2 = np.exp(-1j * a * (pre.mp / m)) * G                           # :142-145
    prod = _ifft_u(t1, N) * _ifft_u(t2, N)                            # inverse + prod
    P = _fft(prod, N).reshape(pre.Nrho, pre.M, n)
    w = (pre.qw * pre.sw / n)[:, None, None]                          # computeQG intw
    QG = _ifft_u((w * P * pre.b).sum(axis=(0, 1)), N)
    loss = _ifft_u(pre.c * G, N)                                      # ax2 + inverse
    return prefac * (QG.real - np.asarray(f) * loss.real)             # output_pq


def maxwellian(cv, m, nden, u, T):
    """n (m/(pi T))^1.5 exp(-m |c - u|^2 / T) (bi/initcond.py:60-64)."""
    u = np.asarray(u, dtype=np.float64).reshape(3, 1)
    return nden * (m / np.pi / T) ** 1.5 * np.exp(-m * ((cv - u) ** 2).sum(0) / T)


class BiOracle:
    """FP64 dgfsBi1D operator (bi/bi.py:265-338): species p's RHS
    invMass(transport_p(u_p) + sum_q sum_{m, r, e} Q_pq(U u_p, V u_q)).
    ``transports[p]`` is an oracle.std.StdOracle carrying species p's walls
    (its transport only is used); ``pres[p, q]`` / ``prefac[p, q]`` the
    pair tables.  us is a list of (K, Ne, n) arrays."""

    def __init__(self, transports, pres, prefac, project=None):
        """project = (cv, m0, m1): the collision terms of each node are replaced
        by their mixture conservation projection (oracle.relax.project_mixture)."""
        self.tr, self.pres, self.prefac = transports, pres, prefac
        self.b = transports[0].b
        self.K, self.Ne = self.b.K, transports[0].Ne
        self.project = project

    def collision(self, p, us):
        b, K, Ne = self.b, self.K, self.Ne
        out = np.zeros_like(us[p])
        for m in range(K):
            f = np.einsum('ra,aen->ren', b.U[m].T, us[p])                 # trans_U
            for q in range(len(us)):
                g = np.einsum('rb,ben->ren', b.V[m], us[q])               # trans_V
                for r in b.sig[m]:
                    for e in range(Ne):
                        out[m, e] += q_pq(self.pres[p, q], f[r, e], g[r, e], self.prefac[p, q])
        return out

    def rhs(self, p, t, us):
        b, K = self.b, self.K
        ux = self.tr[p].transport(t, us[p])
        if self.project is None:
            ux += self.collision(p, us)
        else:
            from .relax import project_mixture
            cv, m0, m1 = self.project
            c0, c1 = self.collision(0, us), self.collision(1, us)
            for m in range(K):
                for e in range(self.Ne):
                    c0[m, e], c1[m, e] = project_mixture(c0[m, e], c1[m, e], cv, m0, m1)
            ux += (c0, c1)[p]
        return (b.invmass @ ux.reshape(K, -1)).reshape(ux.shape)


def integrate_bi(op, scheme, t, dt, us):
    """One step of the mixture schemes (bi/integrator.py:48-176); ssp-rk3 with
    the second stage r1 = 0.75 u + 0.25 r0 + 0.25 dt L(r0)."""
    ns = len(us)
    L = lambda tt, state: [op.rhs(p, tt, state) for p in range(ns)]
    if scheme == 'euler':
        k = L(t, us)
        return [us[p] + dt * k[p] for p in range(ns)]
    if scheme == 'ssp-rk2':
        k = L(t, us)
        r0 = [us[p] + dt * k[p] for p in range(ns)]
        k = L(t + dt, r0)
        return [0.5 * us[p] + 0.5 * r0[p] + 0.5 * dt * k[p] for p in range(ns)]
    if scheme == 'ssp-rk3':
        k = L(t, us)
        r0 = [us[p] + dt * k[p] for p in range(ns)]
        k = L(t + dt, r0)
        r1 = [0.75 * us[p] + 0.25 * r0[p] + 0.25 * dt * k[p] for p in range(ns)]
        k = L(t + 0.5 * dt, r1)
        return [us[p] / 3.0 + 2.0 / 3.0 * r1[p] + 2.0 / 3.0 * dt * k[p] for p in range(ns)]
    if scheme == 'lse-rk45':
        a = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
             -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
        b = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
             1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
             2277821191437.0 / 14882151754819.0]
        c = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
             2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]
        r1 = [np.zeros_like(u) for u in us]
        for i in range(5):
            k = L(t + c[i] * dt, us)
            r1 = [a[i] * r1[p] + dt * k[p] for p in range(ns)]
            us = [us[p] + b[i] * r1[p] for p in range(ns)]
        return us
    raise ValueError(scheme)
