"""Oracle: the linear minmod slope limiter of the IMEX driver (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, element by element:
  * the cell-average and linear-slope operators of the nodal bases
    (basis.py:289-298): V = orthonormal basis at the nodes,
    A = V[:, :1] V^-1[:1], DrLin = D^T V[:, :2] V^-1[:2]
  * limit(): uavg = A u, ulx = invjac DrLin u (astd/abstd.py:402-416)
  * limitLin (astd/kernels/limiters.mako:17-76): first and last element
    copied; an interior element is flagged when
    |a - mm(a - u_0, d+, d-) - u_0| + |a + mm(u_{K-1} - a, d+, d-) - u_{K-1}| > 1e-6
    and then replaced by a + (x_k - x_c) mm(s, d+/h, d-/h).

Pinned by tests/test_oracle_limiter.py: linear data and constants pass
unchanged, cell means are preserved, limited nodal values stay between the
neighbouring cell means, the two end elements are untouched.
"""
import numpy as np

from . import quadratures as Q


def limiter_ops(K, z, D):
    V = Q.ortho_basis_at(K, z).T
    invV = np.linalg.inv(V)
    return V[:, :1] @ invV[:1, :], D.T @ (V[:, :2] @ invV[:2, :])


def minmod(a, b, c):
    s = np.sign(a) + np.sign(b) + np.sign(c)
    m = np.trunc(s / 3.0)
    return np.where(np.abs(m) == 1, m * np.minimum(np.minimum(np.abs(a), np.abs(b)), np.abs(c)),
                    0.0)


def limit_linear(u, z, D, xsol, invjac):
    """u: (K, Ne, n) nodal coefficients; xsol (K, Ne); invjac (Ne,)."""
    K, Ne, n = u.shape
    A, DrLin = limiter_ops(K, z, D)
    uavg = np.einsum('kl,len->ken', A, u)
    ulx = np.einsum('kl,len->ken', DrLin, u) * np.asarray(invjac)[None, :, None]
    out = u.copy()
    for e in range(1, Ne - 1):
        h = 2.0 / invjac[e]
        x0 = xsol[0, e] + h / 2
        a, ap, am = uavg[0, e], uavg[0, e + 1], uavg[0, e - 1]
        dp, dm = ap - a, a - am
        flag = (np.abs(a - minmod(a - u[0, e], dp, dm) - u[0, e])
                + np.abs(a + minmod(u[K - 1, e] - a, dp, dm) - u[K - 1, e])) > 1e-6
        s = minmod(ulx[0, e], dp / h, dm / h)
        for k in range(K):
            out[k, e] = np.where(flag, a + (xsol[k, e] - x0) * s, u[k, e])
    return out


def upwind_ops(K, z, w, D, sigma=-1.0):
    """Upwind SBP operators (basis.py:303-312) in the 'ux = A u' orientation."""
    V = Q.ortho_basis_at(K, z).T
    Vo, _ = np.linalg.qr(V)
    sig = np.zeros(K)
    sig[-1] = sigma
    S = Vo @ np.diag(sig) @ Vo.T
    invP = np.diag(1.0 / w)
    return (D + 0.5 * S @ invP).T, (D - 0.5 * S @ invP).T


def limit_positivity(u, z, D):
    """limitPos (limiters.mako:178-203): theta = min(1, |(a - 1e-13)/(a - min u)|)
    with the absolute value rounded to float, ulim = a + theta (u - a)."""
    K = u.shape[0]
    A, _ = limiter_ops(K, z, D)
    a = np.einsum('kl,len->ken', A, u)[0]
    m = np.minimum(u.min(axis=0), 1e10)
    with np.errstate(divide='ignore', invalid='ignore'):
        q = np.abs(((a - 1e-13) / (a - m)).astype(np.float32)).astype(np.float64)
    th = np.where(np.isnan(q), 1.0, np.minimum(1.0, q))
    return a[None] + th[None] * (u - a[None])
