"""Pins for the limiter oracle (oracle/limiter.py) by what a TVD slope
limiter must keep: linear data and constants pass unchanged, cell means are
preserved, limited node values stay between the neighbouring cell means, and
the two end elements are untouched.  A planted slip (the slope taken without
1/h) changes linear data."""
import numpy as np
import pytest

from oracle import quadratures as Q
from oracle.limiter import limit_linear, limiter_ops


def _setup(K=3, Ne=8):
    z, w = Q.zwglj(K, 0.0, 0.0)
    D = Q.jac_nodal_basis_at(K, z, z)
    xm = np.linspace(0.0, 1.0, Ne + 1)
    h = np.diff(xm)
    xsol = np.array([0.5 * (xm[j] + xm[j + 1]) + h[j] / 2 * z for j in range(Ne)]).T
    return z, w, D, xsol, 2.0 / h


@pytest.mark.parametrize('K', [2, 3, 4])
def test_linear_and_constant_data_unchanged(K):
    z, w, D, xsol, invjac = _setup(K)
    u = (2.0 + 3.0 * xsol)[:, :, None] * np.array([1.0, -0.5, 0.0])[None, None, :]
    assert np.abs(limit_linear(u, z, D, xsol, invjac) - u).max() < 1e-13
    c = np.full_like(u, 0.7)
    assert np.abs(limit_linear(c, z, D, xsol, invjac) - c).max() < 1e-15


@pytest.mark.parametrize('K', [3, 4])
def test_means_preserved_and_bounded(K):
    z, w, D, xsol, invjac = _setup(K, Ne=10)
    rng = np.random.default_rng(3)
    step = (xsol > 0.45).astype(float)
    u = np.stack([step + 0.05 * rng.standard_normal(step.shape),
                  np.sin(7 * xsol) + 0.2 * rng.standard_normal(step.shape)], axis=-1)
    out = limit_linear(u, z, D, xsol, invjac)
    mean = lambda v: np.einsum('q,qen->en', w, v) / 2
    assert np.abs(mean(out) - mean(u)).max() < 1e-13
    assert np.array_equal(out[:, 0], u[:, 0]) and np.array_equal(out[:, -1], u[:, -1])
    A, _ = limiter_ops(K, z, D)
    a = np.einsum('kl,len->ken', A, u)[0]
    changed = np.abs(out - u).max(axis=0) > 0
    assert changed.any()
    for e in range(1, u.shape[1] - 1):
        lo = np.minimum(np.minimum(a[e - 1], a[e]), a[e + 1]) - 1e-13
        hi = np.maximum(np.maximum(a[e - 1], a[e]), a[e + 1]) + 1e-13
        for v in range(u.shape[2]):
            if changed[e, v]:
                assert np.all(out[:, e, v] >= lo[v]) and np.all(out[:, e, v] <= hi[v])


def test_planted_slope_slip_is_caught():
    z, w, D, xsol, invjac = _setup(3)
    # a ramp with an odd-even oscillation at the nodes: every interior element
    # is flagged and gets the (non-zero) minmod slope of the ramp
    osc = 0.05 * (-1.0) ** np.arange(3)[:, None]
    u = (xsol + osc)[:, :, None]
    good = limit_linear(u, z, D, xsol, invjac)
    assert np.abs(good - u).max() > 1e-3
    bad = limit_linear(u, z, D, xsol, np.ones_like(invjac))   # h and slope without 1/J
    assert np.abs(good - bad).max() > 1e-3
