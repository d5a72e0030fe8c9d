"""Oracle: Jacobi polynomials and Gauss-type quadrature rules (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates ``dgfs1D/quadratures.py`` of the reference:
  * three-term Jacobi recurrence            quadratures.py:7-26
  * derivative via P'_n^{a,b} = (n+a+b+1)/2 P_{n-1}^{a+1,b+1}   :28-35
  * orthonormal Legendre basis sqrt(i+1/2) P_i                  :37-54
  * nodal (Lagrange) basis by solving with the orthonormal Vandermonde :56-62
  * Gauss-Radau (z=-1 / z=+1) by Newton deflation (Polylib)     :67-165
  * Gauss-Jacobi and Gauss-Lobatto-Jacobi by Golub-Welsch        :171-220
"""
from math import gamma, cos, pi

import numpy as np


def jacobi(n, a, b, z):
    """[P_0 .. P_n]^{(a,b)}(z) by the standard recurrence (quadratures.py:7-26)."""
    z = np.asarray(z, dtype=np.float64)
    out = [np.ones_like(z)]
    if n >= 1:
        out.append(0.5 * ((a + b + 2.0) * z + (a - b)))
    for q in range(2, n + 1):
        s = a + b
        c0 = 2.0 * q * (q + s) * (2.0 * q + s - 2.0)
        c1 = (2.0 * q + s - 1.0) * ((2.0 * q + s) * (2.0 * q + s - 2.0) * z + a * a - b * b)
        c2 = 2.0 * (q + a - 1.0) * (q + b - 1.0) * (2.0 * q + s)
        out.append((c1 * out[-1] - c2 * out[-2]) / c0)
    return out


def jacobi_diff(n, a, b, z):
    """[P_0' .. P_n'] (quadratures.py:28-35)."""
    z = np.asarray(z, dtype=np.float64)
    out = [np.zeros_like(z)]
    if n >= 1:
        inner = jacobi(n - 1, a + 1.0, b + 1.0, z)
        out.extend(0.5 * (i + a + b + 2.0) * p for i, p in enumerate(inner))
    return out


def ortho_basis_at(order, pts):
    """Orthonormal Legendre basis, shape (order, npts) (quadratures.py:37-48)."""
    pts = np.asarray(pts, dtype=np.float64)
    P = jacobi(order - 1, 0.0, 0.0, pts)
    return np.array([np.sqrt(i + 0.5) * p for i, p in enumerate(P)])


def jac_ortho_basis_at(order, pts):
    """Derivative of the orthonormal basis, shape (order, npts) (:41-54)."""
    pts = np.asarray(pts, dtype=np.float64)
    dP = jacobi_diff(order - 1, 0.0, 0.0, pts)
    return np.array([np.sqrt(i + 0.5) * p for i, p in enumerate(dP)])


def nodal_basis_at(order, pts, epts):
    """Lagrange basis through ``pts`` evaluated at ``epts``; shape (nepts, order)
    exactly as quadratures.py:56-58 returns it (solve(V(pts), V(epts)).T)."""
    return np.linalg.solve(ortho_basis_at(order, pts), ortho_basis_at(order, epts)).T


def jac_nodal_basis_at(order, pts, epts):
    """Derivative of the Lagrange basis; (order, nepts) (quadratures.py:60-62)."""
    return np.linalg.solve(ortho_basis_at(order, pts), jac_ortho_basis_at(order, epts))


def _newton_roots(n, a, b):
    """Zeros of P_n^{(a,b)} by Newton with deflation (Polylib jacobz, :67-95)."""
    if n == 0:
        return np.zeros(0)
    z = np.zeros(n)
    dth = pi / (2.0 * n)
    prev = 0.0
    for k in range(n):
        r = -cos((2.0 * k + 1.0) * dth)
        if k >= 1:
            r = 0.5 * (r + prev)
        for _ in range(5000):
            p = jacobi(n, a, b, r)[-1]
            dp = jacobi_diff(n, a, b, r)[-1]
            s = sum(1.0 / (r - z[i]) for i in range(k))
            d = -p / (dp - s * p)
            r += d
            if abs(d) < 1e-20:
                break
        z[k] = r
        prev = r
    return z


def zwgrjm(n, a, b):
    """Gauss-Radau-Jacobi including z=-1 (quadratures.py:130-146)."""
    z = np.zeros(n)
    w = np.full(n, 2.0)
    if n >= 1:
        z[0] = -1.0
        z[1:] = _newton_roots(n - 1, a, b + 1.0)
        p = jacobi(n, a, b, z)[-1]
        fac = 2.0 ** (a + b) * gamma(a + n) * gamma(b + n)
        fac /= (b + n) * gamma(n) * gamma(a + b + n + 1.0)
        w = fac * (1.0 - z) / (p * p)
        w[0] *= (b + 1.0)
    return z, w


def zwgrjp(n, a, b):
    """Gauss-Radau-Jacobi including z=+1 (quadratures.py:149-165)."""
    z = np.zeros(n)
    w = np.full(n, 2.0)
    if n >= 1:
        z[n - 1] = 1.0
        z[:-1] = _newton_roots(n - 1, a + 1.0, b)
        p = jacobi(n, a, b, z)[-1]
        fac = 2.0 ** (a + b) * gamma(a + n) * gamma(b + n)
        fac /= (a + n) * gamma(n) * gamma(a + b + n + 1.0)
        w = fac * (1.0 + z) / (p * p)
        w[n - 1] *= (a + 1.0)
    return z, w


def _recurrence(n, a, b):
    """Monic Jacobi recurrence coefficients (alpha_i, beta_i) (:171-191)."""
    alpha = np.zeros(n)
    beta = np.zeros(n)
    s = a + b
    alpha[0] = (b - a) / (s + 2.0)
    beta[0] = 2.0 ** (s + 1.0) * gamma(a + 1.0) * gamma(b + 1.0) / gamma(s + 2.0)
    if n > 1:
        beta[1] = 4.0 * (a + 1.0) * (b + 1.0) / ((s + 3.0) * (s + 2.0) ** 2)
    for i in range(1, n):
        t = 2.0 * i + s
        alpha[i] = (b * b - a * a) / (t * (t + 2.0))
        if i + 1 < n:
            k = i + 1
            t2 = 2.0 * k + s
            beta[k] = 4.0 * k * (k + a) * (k + b) * (k + s) / ((t2 * t2 - 1.0) * t2 * t2)
    return alpha, beta


def _golub_welsch(alpha, beta):
    """Nodes/weights from the Jacobi matrix eigen-decomposition (:193-203)."""
    n = len(alpha)
    off = np.sqrt(beta[1:n])
    J = np.diag(alpha) + np.diag(off, -1) + np.diag(off, 1)
    z, V = np.linalg.eigh(J)
    order = np.argsort(z)
    return z[order], beta[0] * V[0, order] ** 2


def zwgj(n, a, b):
    """Gauss-Jacobi rule (quadratures.py:205-208)."""
    return _golub_welsch(*_recurrence(n, a, b))


def zwglj(n, a, b):
    """Gauss-Lobatto-Jacobi rule via the modified Jacobi matrix (:210-220)."""
    alpha, beta = _recurrence(n, a, b)
    N = n - 2
    s1 = a + b + 1.0
    alpha[n - 1] = (a - b) / (2.0 * N + s1 + 1.0)
    beta[n - 1] = 4.0 * (N + a + 1.0) * (N + b + 1.0) * (N + s1) / (
        (2.0 * N + s1) * (2.0 * N + s1 + 1.0) ** 2)
    return _golub_welsch(alpha, beta)
