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


@pytest.mark.parametrize('K', [3, 4, 5])
def test_upwind_operators(K):
    """D+ + D- = 2D; both differentiate polynomials of degree <= K-2 exactly;
    P (D+ - D-)^T = S is symmetric negative semidefinite (dissipative)."""
    from oracle.limiter import upwind_ops
    z, w, D, xsol, invjac = _setup(K)
    Ap, Am = upwind_ops(K, z, w, D)
    assert np.abs((Ap + Am) - 2 * D.T).max() < 1e-13
    for deg in range(K - 1):
        p, dp = z ** deg, deg * z ** max(deg - 1, 0)
        assert np.abs(Ap @ p - dp).max() < 1e-12 and np.abs(Am @ p - dp).max() < 1e-12
    S = np.diag(w) @ (Ap - Am)
    assert np.abs(S - S.T).max() < 1e-13
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).max() < 1e-13


def test_upwind_transport_without_dissipation_is_the_standard_one():
    """With sigma = 0 (D+ = D- = D) the split upwind transport equals the
    standard upwind-flux transport exactly (the flux splitting reproduces
    1/2 |c| (uL - uR) +- 1/2 c (...) term by term)."""
    from oracle.deck import Deck
    from oracle.limiter import upwind_ops
    from tests.test_gpu_solver import _ini
    deck = Deck(_ini({('mesh', 'Ne'): 4}))
    b = deck.basis
    Ap, Am = upwind_ops(b.K, b.z, b.w, b.D, sigma=0.0)
    rng = np.random.default_rng(2)
    u = deck.u0 * (1 + 0.1 * rng.uniform(-1, 1, deck.u0.shape))
    ref = deck.op.transport(0.0, u)
    got = deck.op.transport_upwind(0.0, u, Ap, Am)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
    Ap1, Am1 = upwind_ops(b.K, b.z, b.w, b.D, sigma=-1.0)
    assert np.abs(deck.op.transport_upwind(0.0, u, Ap1, Am1) - ref).max() > 1e-6 * np.abs(ref).max()


def test_positivity_limiter_pins():
    """Cell means kept, an element with a negative node value scaled towards
    its mean until the minimum is non-negative, other elements untouched."""
    from oracle.limiter import limit_positivity
    z, w, D, xsol, invjac = _setup(3, Ne=6)
    rng = np.random.default_rng(5)
    u = rng.uniform(0.1, 1.0, (3, 6, 20))
    u[0, 2, 3] = -0.05                    # one negative node value, mean still positive
    out = limit_positivity(u, z, D)
    mean = lambda v: np.einsum('q,qen->en', w, v) / 2
    assert np.abs(mean(out) - mean(u)).max() < 1e-13
    assert out[:, 2, 3].min() >= -1e-7 and out[:, 2, 3].min() < 1e-6
    untouched = np.ones((6, 20), bool)
    untouched[2, 3] = False
    assert np.abs(out - u)[:, untouched].max() < 1e-6 * np.abs(u).max()


def test_entropy_transport_pins():
    """The log mean: L(a, a) = a, the closed form (a - b)/(ln a - ln b) away
    from a = b, symmetry.  The flux-differencing volume term with the
    arithmetic mean in place of L is exactly -c D u, so the entropy transport
    then equals the standard transport; with L it differs slightly on smooth
    positive data."""
    from oracle.deck import Deck
    from oracle.std import StdOracle
    from tests.test_gpu_solver import _ini
    a = np.array([0.3, 1.0, 2.0, 5e-3])
    b = np.array([0.7, 1.0, 2.0 * (1 + 1e-3), 4e-3])
    L = StdOracle.log_mean(a, b)
    closed = np.where(a == b, a, (a - b) / (np.log(a) - np.log(b)))
    assert np.allclose(L, closed, rtol=1e-12)
    assert np.allclose(StdOracle.log_mean(b, a), L, rtol=1e-15)
    deck = Deck(_ini({('mesh', 'Ne'): 4}))
    rng = np.random.default_rng(4)
    u = deck.u0 * (1 + 0.1 * rng.uniform(-1, 1, deck.u0.shape))
    Dr = deck.basis.deriv                       # the 'ux = A u' derivative (derivMat)
    arith = lambda x, y: 0.5 * (x + y)
    ref = deck.op.transport(0.0, u)
    assert np.abs(deck.op.transport_entropy(0.0, u, Dr, mean=arith) - ref).max() < \
        1e-12 * np.abs(ref).max()
    d = np.abs(deck.op.transport_entropy(0.0, u, Dr) - ref).max() / np.abs(ref).max()
    assert 1e-9 < d < 1e-1
