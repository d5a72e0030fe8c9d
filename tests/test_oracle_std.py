"""Pins for the single-species operator oracle (oracle/std.py) by what the
discretisation must satisfy, independent of its formulas:
  * the DG transport differentiates polynomials of degree < K exactly (copy
    walls: no boundary jumps) and conserves mass up to the boundary flux for
    any (discontinuous) data;
  * the diffuse wall admits no net mass flux through the wall;
  * the explicit integrators have the amplification polynomials of their
    order on u' = lambda u and integrate u' = t^(q-1) exactly at their order
    (stage times);
  * discrete moments of a resolved Maxwellian return its (rho, u, T) with
    zero heat flux and shear stress; those of an anisotropic drifting
    Gaussian return its covariance as T and P (with the factor 2), and a
    Grad-13 perturbation returns its closed-form heat flux.
"""
import numpy as np
import pytest

from oracle import std as OS
from oracle.deck import Deck


def _deck(K=3, Ne=4, walls='dgfs-periodic'):
    from tests.test_gpu_solver import _ini
    return Deck(_ini({('basis', 'K'): K, ('mesh', 'Ne'): Ne, ('soln-bcs-xlo', 'type'): walls,
                      ('soln-bcs-xhi', 'type'): walls}))


@pytest.mark.parametrize('K', [2, 3, 4, 5])
def test_transport_exact_on_polynomials(K):
    d = _deck(K)
    x = d.xsol                                        # (K, Ne) node coordinates
    c = d.vm.cv[0]
    for deg in range(K):
        p, dp = x ** deg, deg * x ** max(deg - 1, 0)
        g = np.exp(-(d.vm.cv ** 2).sum(0))            # any velocity profile
        u = p[:, :, None] * g[None, None, :]
        got = d.op.transport(0.0, u)
        ref = -c[None, None, :] * dp[:, :, None] * g[None, None, :]
        assert np.abs(got - ref).max() < 1e-11 * max(1.0, np.abs(u).max()), deg


@pytest.mark.parametrize('K', [2, 3, 4])
def test_transport_conserves_up_to_boundary_flux(K):
    """sum_e sum_k w_k h_e/2 u_t = -c (u_right - u_left) for any data."""
    d = _deck(K, Ne=5)
    rng = np.random.default_rng(K)
    u = rng.uniform(0.1, 1.0, (K, 5, d.vm.vsize))    # discontinuous at every face
    ut = d.op.transport(0.0, u)
    w = d.basis.w
    h = 1.0 / d.invjac * 2.0
    total = np.einsum('k,e,ken->n', w, h / 2.0, ut)
    c = d.vm.cv[0]
    ref = -c * (u[K - 1, -1] - u[0, 0])
    assert np.abs(total - ref).max() < 1e-12 * np.abs(ref).max()


def test_diffuse_wall_has_no_net_mass_flux():
    d = _deck()
    vm = d.vm
    c = vm.cv[0]
    rng = np.random.default_rng(0)
    interior = OS.maxwellian(vm.cv, 1.0, 0.1, 0.2, 0.0, 1.2) * rng.uniform(0.9, 1.1, vm.vsize)
    for nl, wall_T in ((-1.0, 0.9), (1.0, 1.1)):
        f0 = OS.maxwellian(vm.cv, 1.0, 0.0, 0.3, 0.0, wall_T)
        w = OS.Wall('diffuse', nl, vm.cv, vm.cw, f0=f0)
        ghost = w.apply(interior)
        # physical upwind state at the wall: leaving the domain -> interior
        leaving = (c * nl) > 0
        face = np.where(leaving, interior, ghost)
        flux = (c * face).sum() * vm.cw
        assert abs(flux) < 1e-13 * (np.abs(c) * face).sum() * vm.cw


def test_expression_wall_matches_fixed_wall_and_carries_no_mass():
    """The expression wall (std/bc.py:136-198) uses the unnormalised
    exp(-|c - u(t)|^2 / T(t)); its density factor from the zero-mass-flux
    condition makes the ghost state independent of that normalisation: with
    constant expressions it equals the fixed diffuse wall's ghost, and at any
    time its net mass flux vanishes."""
    d = _deck()
    vm = d.vm
    c = vm.cv[0]
    rng = np.random.default_rng(1)
    interior = OS.maxwellian(vm.cv, 1.0, 0.1, 0.2, 0.0, 1.2) * rng.uniform(0.9, 1.1, vm.vsize)
    for nl in (-1.0, 1.0):
        fixed = OS.Wall('diffuse', nl, vm.cv, vm.cw,
                        f0=OS.maxwellian(vm.cv, 1.0, 0.0, 0.3, 0.0, 0.95))
        expr = OS.Wall('diffuse-expr', nl, vm.cv, vm.cw, state=lambda t: (0.0, 0.3, 0.0, 0.95))
        g1, g2 = fixed.apply(interior), expr.apply(interior, 0.7)
        assert np.abs(g1 - g2).max() < 1e-14 * np.abs(g1).max()
        moving = OS.Wall('diffuse-expr', nl, vm.cv, vm.cw,
                         state=lambda t: (0.0, np.sin(t), 0.1 * t, 1.0 + 0.2 * t))
        for t in (0.0, 0.4, 1.3):
            ghost = moving.apply(interior, t)
            face = np.where((c * nl) > 0, interior, ghost)
            assert abs((c * face).sum()) < 1e-13 * (np.abs(c) * face).sum()


class _Lin:
    def __init__(self, lam):
        self.lam = lam

    def rhs(self, t, u):
        return self.lam * u


class _Poly:
    def __init__(self, q):
        self.q = q

    def rhs(self, t, u):
        return np.full_like(u, t ** (self.q - 1))


@pytest.mark.parametrize('scheme,order', [('euler', 1), ('ssp-rk2', 2), ('ssp-rk3', 3),
                                          ('lse-rk45', 4)])
def test_integrators(scheme, order):
    from math import factorial
    z = -0.05
    R = OS.integrate(_Lin(z), scheme, 0.0, 1.0, np.ones(1))[0]
    taylor = sum(z ** k / factorial(k) for k in range(order + 1))
    # the z^(order+1) coefficient of a many-stage scheme need not be 1/(order+1)!
    # (LSERK45: about 1/200); one order less would miss by z^order / order! >> this
    assert abs(R - taylor) < 1e-2 * abs(z) ** (order + 1)
    # u' = t^(q-1), exact solution t^q / q: exact for q <= order
    for q in range(1, order + 1):
        t0, dt = 0.3, 0.2
        u1 = OS.integrate(_Poly(q), scheme, t0, dt, np.zeros(1))[0]
        assert abs(u1 - ((t0 + dt) ** q - t0 ** q) / q) < 1e-14, (q, u1)


def test_moments_of_a_maxwellian():
    from tests.test_gpu_solver import _ini
    d = Deck(_ini({('velocity-mesh', 'Nv'): 24}))
    vm, b = d.vm, d.basis
    f = OS.maxwellian(vm.cv, 0.8, 0.1, -0.2, 0.0, 1.1)
    u = np.broadcast_to(np.einsum('kq,q->k', b.fwd, np.ones(b.Nq))[:, None, None] * f,
                        (d.K, d.Ne, vm.vsize)).copy()
    m = OS.moments(u, b.interp, vm)
    assert np.allclose(m[..., 0], 0.8, rtol=1e-6)
    assert np.allclose(m[..., 1], 0.1, atol=1e-7) and np.allclose(m[..., 2], -0.2, atol=1e-7)
    assert np.allclose(m[..., 3], 1.1, rtol=1e-6)
    assert np.abs(m[..., 4:6]).max() < 1e-7 and np.abs(m[..., 8]).max() < 1e-7


def _gaussian(cv, rho, u, Sig):
    """rho times the normalised Gaussian of mean u and covariance Sig."""
    d = cv - np.asarray(u, float).reshape(3, 1)
    Si = np.linalg.inv(Sig)
    e = np.einsum('iv,ij,jv->v', d, Si, d)
    return rho * np.exp(-0.5 * e) / ((2 * np.pi) ** 1.5 * np.sqrt(np.linalg.det(Sig)))


def _moments_of(f, Nv=32):
    from tests.test_gpu_solver import _ini
    d = Deck(_ini({('velocity-mesh', 'Nv'): Nv}))
    vm, b = d.vm, d.basis
    u = np.broadcast_to(np.einsum('kq,q->k', b.fwd, np.ones(b.Nq))[:, None, None] * f(vm.cv),
                        (d.K, d.Ne, vm.vsize)).copy()
    return OS.moments(u, b.interp, vm)


def test_moments_of_an_anisotropic_drifting_gaussian():
    """Closed forms for f = rho N(u, Sig) (u_z = 0): T = (2/3) tr Sig and
    P_ij = 2 rho Sig_ij (std/moments.py:58-61 carries the factor 2), q = 0.
    A dropped factor 2 on P, a P built from absolute instead of peculiar
    velocities, or a T without the 2/3 all fail."""
    rho, u = 0.7, (0.25, -0.35, 0.0)
    Sig = np.array([[0.55, 0.12, 0.0], [0.12, 0.40, 0.0], [0.0, 0.0, 0.45]])
    m = _moments_of(lambda cv: _gaussian(cv, rho, u, Sig))
    tol = 3e-8          # the grid's truncation at |c_i| = 5 (N = 32)
    assert np.abs(m[..., 0] - rho).max() < tol
    assert np.abs(m[..., 1] - u[0]).max() < tol and np.abs(m[..., 2] - u[1]).max() < tol
    assert np.abs(m[..., 3] - 2.0 / 3.0 * np.trace(Sig)).max() < tol
    assert np.abs(m[..., 4:6]).max() < tol
    assert np.abs(m[..., 6] - 2 * rho * Sig[0, 0]).max() < tol
    assert np.abs(m[..., 7] - 2 * rho * Sig[1, 1]).max() < tol
    assert np.abs(m[..., 8] - 2 * rho * Sig[0, 1]).max() < tol


def test_moments_heat_flux_of_a_grad13_perturbation():
    """f = M (1 + a (c.e)(|c|^2/T - 5/2)) around a resting Maxwellian
    M = rho/(pi T)^{3/2} exp(-|c|^2/T) (variance T/2 per axis): the
    perturbation carries no mass, momentum, energy or stress, and its heat
    flux is q = int f |c|^2 c = (5/4) rho a T^2 e (from E[c_x^2|c|^2] = 5 s^4,
    E[c_x^2|c|^4] = 35 s^6, s^2 = T/2)."""
    rho, T, a = 0.9, 1.1, 0.08
    e = np.array([0.6, 0.8, 0.0])

    def f(cv):
        M = OS.maxwellian(cv, rho, 0.0, 0.0, 0.0, T)
        ce = e @ cv
        return M * (1 + a * ce * ((cv * cv).sum(0) / T - 2.5))

    m = _moments_of(f)
    tol = 3e-8          # the grid's truncation at |c_i| = 5 (N = 32)
    assert np.abs(m[..., 0] - rho).max() < tol
    assert np.abs(m[..., 1:3]).max() < tol
    assert np.abs(m[..., 3] - T).max() < tol
    q = 1.25 * rho * a * T * T
    assert np.abs(m[..., 4] - q * e[0]).max() < tol
    assert np.abs(m[..., 5] - q * e[1]).max() < tol
    assert np.abs(m[..., 6] - rho * T).max() < tol and np.abs(m[..., 7] - rho * T).max() < tol
    assert np.abs(m[..., 8]).max() < tol
