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
