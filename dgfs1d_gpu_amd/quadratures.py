"""Quadrature rules and 1D polynomial bases used at set-up on the host.

Gauss-Jacobi nodes come from scipy.special.roots_jacobi; Gauss-Lobatto and
Gauss-Radau rules are built from those nodes with the closed-form weights of
the respective rules (Karniadakis & Sherwin, App. B).  Same rules and return
conventions as the reference's dgfs1D/quadratures.py (zwgj, zwglj, zwgrjm,
zwgrjp, ortho_basis_at, nodal_basis_at, jac_*), independent implementation.
"""
from math import lgamma

import numpy as np
from scipy.special import eval_jacobi, roots_jacobi


def _jac(n, a, b, x):
    return eval_jacobi(n, a, b, x) if n >= 0 else np.zeros_like(x)


def _djac(n, a, b, x):
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64))
    return 0.5 * (n + a + b + 1.0) * eval_jacobi(n - 1, a + 1.0, b + 1.0, x)


def zwgj(n, a, b):
    z, w = roots_jacobi(n, a, b)
    return np.asarray(z, dtype=np.float64), np.asarray(w, dtype=np.float64)


def zwglj(n, a, b):
    """Gauss-Lobatto-Jacobi: interior = zeros of P_{n-2}^{(a+1,b+1)}."""
    z = np.empty(n)
    z[0], z[-1] = -1.0, 1.0
    if n > 2:
        z[1:-1] = roots_jacobi(n - 2, a + 1.0, b + 1.0)[0]
    p = _jac(n - 1, a, b, z)
    lfac = ((a + b + 1.0) * np.log(2.0) + lgamma(a + n) + lgamma(b + n)
            - np.log(n - 1.0) - lgamma(n) - lgamma(a + b + n + 1.0))
    w = np.exp(lfac) / (p * p)
    w[0] *= (b + 1.0)
    w[-1] *= (a + 1.0)
    return z, w


def zwgrjm(n, a, b):
    """Gauss-Radau-Jacobi including z = -1: interior = zeros of P_{n-1}^{(a,b+1)}.

    The weights use P_n^{(a,b)} at the nodes, as the reference does
    (dgfs1D/quadratures.py:138); this equals the textbook P_{n-1} form only for
    a = b = 0 (then P_n = -P_{n-1} at the Radau nodes)."""
    z = np.empty(n)
    z[0] = -1.0
    if n > 1:
        z[1:] = roots_jacobi(n - 1, a, b + 1.0)[0]
    p = _jac(n, a, b, z)
    lfac = ((a + b) * np.log(2.0) + lgamma(a + n) + lgamma(b + n)
            - np.log(b + n) - lgamma(n) - lgamma(a + b + n + 1.0))
    w = np.exp(lfac) * (1.0 - z) / (p * p)
    w[0] *= (b + 1.0)
    return z, w


def zwgrjp(n, a, b):
    """Gauss-Radau-Jacobi including z = +1 (weights with P_n, as zwgrjm)."""
    z = np.empty(n)
    z[-1] = 1.0
    if n > 1:
        z[:-1] = roots_jacobi(n - 1, a + 1.0, b)[0]
    p = _jac(n, a, b, z)
    lfac = ((a + b) * np.log(2.0) + lgamma(a + n) + lgamma(b + n)
            - np.log(a + n) - lgamma(n) - lgamma(a + b + n + 1.0))
    w = np.exp(lfac) * (1.0 + z) / (p * p)
    w[-1] *= (a + 1.0)
    return z, w


def jacobi(n, a, b, z):
    """[P_0 .. P_n]^{(a,b)}(z)."""
    z = np.asarray(z, dtype=np.float64)
    return [_jac(i, a, b, z) for i in range(n + 1)]


def jacobi_diff(n, a, b, z):
    z = np.asarray(z, dtype=np.float64)
    return [_djac(i, a, b, z) for i in range(n + 1)]


def ortho_basis_at(order, pts):
    """Orthonormal Legendre basis sqrt(i+1/2) P_i, shape (order, npts)."""
    pts = np.asarray(pts, dtype=np.float64)
    return np.array([np.sqrt(i + 0.5) * _jac(i, 0.0, 0.0, pts) for i in range(order)])


def jac_ortho_basis_at(order, pts):
    pts = np.asarray(pts, dtype=np.float64)
    return np.array([np.sqrt(i + 0.5) * _djac(i, 0.0, 0.0, pts) for i in range(order)])


def nodal_basis_at(order, pts, epts):
    """Lagrange basis through pts at epts, shape (len(epts), order)."""
    return np.linalg.solve(ortho_basis_at(order, pts), ortho_basis_at(order, epts)).T


def jac_nodal_basis_at(order, pts, epts):
    """Derivatives of the Lagrange basis, shape (order, len(epts))."""
    return np.linalg.solve(ortho_basis_at(order, pts), jac_ortho_basis_at(order, epts))
