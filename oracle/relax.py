"""Oracle: Shakov relaxation term and the conservation projection (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

``q_shakov`` restates, step for step, the reference's per-node
'shakov-direct-gll' fs() (dgfs1D/std/scattering.py:1786-1833) with the
kernels of std/kernels/scattering/shakov-direct-gll.mako:
  flocKern                       :15-29    f = d_f[(modein, elem)]
  mom1 + sums + mom01Norm        :32-61    u_i = sum f c_i / sum f, rho = cw sum f
  mom23Sk + sums + mom23SkNorm   :63-104   T = cw sum f C^2 / (1.5 rho),
                                           q_i = cw sum f C^2 C_i
  equiSkDistCompute              :106-142  f_S
  output                         :145-167  Q += prefac rho / T^(omega-1) (f_S - f)

``q_bgk`` restates 'bgk-direct-gll' (a Maxwellian target) and ``q_esbgk``
'esbgk-direct-gll' (an anisotropic Gaussian target) the same way.

``project_conservative`` is the Euclidean-nearest slice with vanishing
mass, momentum and energy moments (no reference counterpart), computed with
a least-squares solve so that it does not share the product's Gram-inverse
formulation.
"""
import numpy as np


def shakov_moments(f, cv, cw):
    """(rho, u (3,), T, q (3,)) by the reference's two-pass sequence."""
    s0 = f.sum()
    u = np.array([(f * cv[0]).sum(), (f * cv[1]).sum(), (f * cv[2]).sum()]) / s0
    rho = s0 * cw
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    T = (f * C2).sum() * cw / (1.5 * rho)
    q = np.array([(f * C2 * C[i]).sum() for i in range(3)]) * cw
    return rho, u, T, q


def shakov_fS(f, cv, cw, Pr):
    """Shakov target distribution (mako:106-142)."""
    rho, u, T, q = shakov_moments(f, cv, cw)
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    fe = rho / (np.pi * T) ** 1.5 * np.exp(-C2 / T)
    qc = q[0] * C[0] + q[1] * C[1] + q[2] * C[2]
    return fe * (1.0 + (1.0 - Pr) / 5.0 * 4.0 * qc / (rho * T * T) * (C2 / T - 2.5))


def q_shakov(f, cv, cw, prefac, omega, Pr):
    """Increment added to Q[modeout, elem] by one fs() call."""
    rho, _, T, _ = shakov_moments(f, cv, cw)
    return prefac * (rho / T ** (omega - 1.0)) * (shakov_fS(f, cv, cw, Pr) - f)


def q_bgk(f, cv, cw, prefac, omega):
    """One 'bgk-direct-gll' fs() call (std/scattering.py:1313-1369, kernels of
    std/kernels/scattering/bgk-direct-gll.mako): mom1/mom01Norm (:32-61) give
    u = sum f c / sum f and rho = cw sum f, mom2/equiDistInit (:63-92) give
    T = cw sum f |c - u|^2 / (1.5 rho), and output (:95-122) adds
    prefac rho / T^(omega-1) (rho/(pi T)^1.5 exp(-|c - u|^2/T) - f)."""
    s0 = f.sum()
    ux, uy, uz = (f * cv[0]).sum() / s0, (f * cv[1]).sum() / s0, (f * cv[2]).sum() / s0
    rho = s0 * cw
    C2 = (cv[0] - ux) ** 2 + (cv[1] - uy) ** 2 + (cv[2] - uz) ** 2
    T = (f * C2).sum() * cw / (1.5 * rho)
    fe = rho / (np.pi * T) ** 1.5 * np.exp(-C2 / T)
    return prefac * (rho / T ** (omega - 1.0)) * (fe - f)


def q_esbgk(f, cv, cw, prefac, omega, Pr):
    """One 'esbgk-direct-gll' fs() call (std/scattering.py:1540-1584, kernels
    of std/kernels/scattering/esbgk-direct-gll.mako): rho, u, T as BGK
    (:32-94), the discrete Maxwellian fe of them (:97-118), the tensor
    L_ij = cw/rho sum (c_i - u_i)(c_j - u_j) ((1 - 1/Pr) f + (1/Pr) fe)
    (mom2ES, mom2ESNorm :120-162), the Gaussian
    rho / ((2 pi)^1.5 sqrt(det L)) exp(-C^T L^-1 C / 2) (:164-207; inverted
    here with numpy.linalg, the reference writes cofactors), and output
    (:209-229): prefac rho / T^(omega-1) (G - f)."""
    s0 = f.sum()
    u = np.array([(f * cv[0]).sum(), (f * cv[1]).sum(), (f * cv[2]).sum()]) / s0
    rho = s0 * cw
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    T = (f * C2).sum() * cw / (1.5 * rho)
    fe = rho / (np.pi * T) ** 1.5 * np.exp(-C2 / T)
    w = (1.0 - 1.0 / Pr) * f + (1.0 / Pr) * fe
    L = np.array([[(C[i] * C[j] * w).sum() for j in range(3)] for i in range(3)]) * cw / rho
    A = np.linalg.inv(L)
    G = rho / ((2.0 * np.pi) ** 1.5 * np.sqrt(np.linalg.det(L))) * \
        np.exp(-0.5 * np.einsum('iv,ij,jv->v', C, A, C))
    return prefac * (rho / T ** (omega - 1.0)) * (G - f)


def bgk_conservative_target(f, cv, cw):
    """'bgk-gll' target (std/scattering.py:600-706, bgk-gll.mako:83-229): the
    discrete Maxwellian a0 exp(-a1 |C|^2 + a2 Cx + a3 Cy + a4 Cz), C = c - u,
    whose discrete mass, momentum and energy equal f's.  Solved here with
    scipy.optimize.root (MINPACK hybrd) from the reference's starting point
    (rho/(pi T)^1.5, 1/T, 0, 0, 0); the reference iterates Newton steps."""
    from scipy.optimize import root
    s0 = f.sum()
    u = np.array([(f * cv[0]).sum(), (f * cv[1]).sum(), (f * cv[2]).sum()]) / s0
    rho = s0 * cw
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    T = (f * C2).sum() * cw / (1.5 * rho)
    phi = np.vstack([np.ones_like(C2), cv, (cv * cv).sum(0)]) * cw
    m = phi @ f
    ar = abs(rho)
    sc = np.array([ar, ar * np.sqrt(T), ar * np.sqrt(T), ar * np.sqrt(T), ar * T])

    def fe(a):
        return a[0] * np.exp(-a[1] * C2 + a[2] * C[0] + a[3] * C[1] + a[4] * C[2])
    fun = lambda a: (phi @ fe(a) - m) / sc
    sol = root(fun, [rho / (np.pi * T) ** 1.5, 1.0 / T, 0, 0, 0], method='hybr',
               options={'xtol': 1e-15})
    _check_root(sol, fun, 'bgk_conservative_target')
    return fe(sol.x), rho, T


def _check_root(sol, fun, what):
    """MINPACK may stop on 'no further progress' at machine precision: accept a
    solve whose scaled residual is below 1e-11 either way, refuse the rest."""
    res = float(np.abs(fun(sol.x)).max())
    if not np.isfinite(res) or (not sol.success and res > 1e-11) or res > 1e-9:
        raise RuntimeError('%s: root solve failed (%s, residual %.3e)' % (what, sol.message, res))


def q_bgk_conservative(f, cv, cw, prefac, omega):
    """One 'bgk-gll' fs() call: prefac rho / T^(omega-1) (fe - f)."""
    fe, rho, T = bgk_conservative_target(f, cv, cw)
    return prefac * (rho / T ** (omega - 1.0)) * (fe - f)


def esbgk_conservative_target(f, cv, cw, Pr):
    """'esbgk-gll' target (std/scattering.py:752-1185, esbgk-gll.mako:218-416):
    fe the conservative BGK target, T_ij = cw/rho sum C_i C_j ((1 - 1/Pr) f
    + (1/Pr) fe), and the discrete Gaussian in the reference's parameters
    a0 exp(-a1 Cx^2 + a2 Cx - a3 Cy^2 + a4 Cy - a5 Cz^2 + a6 Cz + a7 cx cy
    + a8 cx cz + a9 cy cz) whose mass, momentum and second moments equal
    rho, rho u, rho (u_i u_j + T_ij); solved with scipy.optimize.root."""
    from scipy.optimize import root
    fe, rho, T = bgk_conservative_target(f, cv, cw)
    s0 = f.sum()
    u = np.array([(f * cv[0]).sum(), (f * cv[1]).sum(), (f * cv[2]).sum()]) / s0
    C = cv - u[:, None]
    w = (1.0 - 1.0 / Pr) * f + (1.0 / Pr) * fe
    pr = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    Tij = [(C[i] * C[j] * w).sum() * cw / rho for i, j in pr]
    m = np.array([rho, *(rho * u)] + [rho * (u[i] * u[j] + Tij[k]) for k, (i, j) in enumerate(pr)])
    phi = np.vstack([np.ones(cv.shape[1]), cv] + [cv[i] * cv[j] for i, j in pr]) * cw
    ar = abs(rho)
    sc = np.array([ar] + [ar * np.sqrt(T)] * 3 + [ar * T] * 6)

    def G(a):
        return a[0] * np.exp(-a[1] * C[0] ** 2 + a[2] * C[0] - a[3] * C[1] ** 2 + a[4] * C[1]
                             - a[5] * C[2] ** 2 + a[6] * C[2] + a[7] * cv[0] * cv[1]
                             + a[8] * cv[0] * cv[2] + a[9] * cv[1] * cv[2])
    a0 = [rho / (np.pi * T) ** 1.5, 1 / T, 0, 1 / T, 0, 1 / T, 0, 0, 0, 0]
    fun = lambda a: (phi @ G(a) - m) / sc
    sol = root(fun, a0, method='hybr', options={'xtol': 1e-15})
    _check_root(sol, fun, 'esbgk_conservative_target')
    return G(sol.x), rho, T


def q_esbgk_conservative(f, cv, cw, prefac, omega, Pr):
    """One 'esbgk-gll' fs() call: prefac rho / T^(omega-1) (G - f)."""
    G, rho, T = esbgk_conservative_target(f, cv, cw, Pr)
    return prefac * (rho / T ** (omega - 1.0)) * (G - f)


def shakov_prefactor(H0, u0, T0, n0, muRef, Tref, omega, R0=8.3144598, NA=6.0221409e+23):
    """t0 p0 / mu(T0) with t0 = H0/u0, p0 = n0 R0/NA T0 (std/scattering.py:1645-1661)."""
    t0 = H0 / u0
    visc = muRef * (T0 / Tref) ** omega
    p0 = n0 * R0 / NA * T0
    return t0 * p0 / visc


def invariants(cv):
    return np.vstack([np.ones(cv.shape[1]), cv, (cv * cv).sum(0)])


def project_conservative(Q, cv):
    """Q - Phi^T lam with lam the least-squares coefficients of Q on span(Phi)."""
    phi = invariants(np.asarray(cv, dtype=np.float64))
    lam, *_ = np.linalg.lstsq(phi.T, np.asarray(Q, dtype=np.float64), rcond=None)
    return Q - phi.T @ lam


def project_mixture(Q0, Q1, cv, m0, m1):
    """Binary mixture: the least-norm (dQ0, dQ1) making sum Q0, sum Q1,
    m0 sum c Q0 + m1 sum c Q1 and m0 sum |c|^2 Q0 + m1 sum |c|^2 Q1 vanish,
    by a least-norm solve of the stacked 6 x 2n constraint system
    (numpy.linalg.lstsq), independent of the product's Gram inverse."""
    phi = invariants(np.asarray(cv, dtype=np.float64))
    n = phi.shape[1]
    A = np.zeros((6, 2 * n))
    A[0, :n] = 1.0
    A[1, n:] = 1.0
    A[2:, :n] = m0 * phi[1:]
    A[2:, n:] = m1 * phi[1:]
    q = np.concatenate([np.asarray(Q0, dtype=np.float64), np.asarray(Q1, dtype=np.float64)])
    d, *_ = np.linalg.lstsq(A, -(A @ q), rcond=None)     # least-norm correction
    return Q0 + d[:n], Q1 + d[n:]
