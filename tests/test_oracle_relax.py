"""Pins of the Shakov and projection oracles (CPU) by facts the mathematics fixes.

Shakov: on a grid wide enough for the Gaussian tails, the discrete moments of
f_S equal (rho, u, T) of f and its heat flux equals (1 - Pr) q -- for the
reference's non-dimensionalisation (variance T/2 per axis) this fixes the
factor 4 (1 - Pr)/5 and the 5/2 of the correction; a Maxwellian is a fixed
point (Q_S = 0); Q_S conserves mass, momentum and energy.  Planted mistakes
in those constants fail the moment test.
Projection: the output has vanishing invariant moments, differs from the input
by an element of span{1, c, |c|^2}, returns inputs orthogonal to that span
unchanged and annihilates elements of the span.
"""
import numpy as np
import pytest

from oracle import relax as OR


def grid(N=32, L=7.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(c0, c0, c0, indexing='ij')
    cv = np.stack([I.ravel(), J.ravel(), K.ravel()])
    return cv, (2 * L / N) ** 3


def maxw(cv, rho, u, T):
    u = np.asarray(u, float).reshape(3, 1)
    return rho / (np.pi * T) ** 1.5 * np.exp(-((cv - u) ** 2).sum(0) / T)


def bimodal(cv):
    return 0.6 * maxw(cv, 1.0, (0.3, 0.1, 0.0), 0.8) + 0.5 * maxw(cv, 0.9, (-0.2, 0.25, 0.1), 1.3)


def moments_of(g, cv, cw):
    """Direct discrete moments (rho, u, T, q) of any g (independent loop)."""
    rho = g.sum() * cw
    u = (g[None, :] * cv).sum(1) * cw / rho
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    T = 2.0 / 3.0 * (g * C2).sum() * cw / rho
    q = (g[None, :] * C2[None, :] * C).sum(1) * cw
    return rho, u, T, q


@pytest.mark.parametrize('Pr', [2.0 / 3.0, 0.5, 1.0])
def test_shakov_target_moments(Pr):
    cv, cw = grid()
    f = bimodal(cv)
    rho, u, T, q = moments_of(f, cv, cw)
    assert np.abs(q).max() > 1e-2                       # non-trivial heat flux
    fS = OR.shakov_fS(f, cv, cw, Pr)
    rS, uS, TS, qS = moments_of(fS, cv, cw)
    assert abs(rS - rho) / rho < 1e-11
    assert np.abs(uS - u).max() < 1e-11
    assert abs(TS - T) / T < 1e-11
    assert np.abs(qS - (1.0 - Pr) * q).max() < 1e-10 * np.abs(q).max() + 1e-14


def test_shakov_moment_state_matches_direct():
    cv, cw = grid()
    f = bimodal(cv)
    rho, u, T, q = OR.shakov_moments(f, cv, cw)
    r2, u2, T2, q2 = moments_of(f, cv, cw)
    assert abs(rho - r2) / r2 < 1e-14 and np.abs(u - u2).max() < 1e-14
    assert abs(T - T2) / T2 < 1e-13 and np.abs(q - q2).max() < 1e-13


def _mutant_fS(f, cv, cw, Pr, k4=4.0, half5=2.5):
    rho, u, T, q = OR.shakov_moments(f, cv, cw)
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    fe = rho / (np.pi * T) ** 1.5 * np.exp(-C2 / T)
    qc = (q[:, None] * C).sum(0)
    return fe * (1.0 + (1.0 - Pr) / 5.0 * k4 * qc / (rho * T * T) * (C2 / T - half5))


@pytest.mark.parametrize('k4,half5', [(2.0, 2.5), (4.0, 5.0), (4.0, 1.5)])
def test_planted_shakov_constant_fails(k4, half5):
    cv, cw = grid()
    f = bimodal(cv)
    _, _, T, q = moments_of(f, cv, cw)
    Pr = 2.0 / 3.0
    _, uS, TS, qS = moments_of(_mutant_fS(f, cv, cw, Pr, k4, half5), cv, cw)
    bad = np.abs(qS - (1.0 - Pr) * q).max() > 1e-6 * np.abs(q).max() or \
        abs(TS - T) / T > 1e-6
    assert bad


def test_shakov_maxwellian_fixed_point():
    cv, cw = grid()
    f = maxw(cv, 1.3, (0.2, -0.1, 0.05), 0.9)
    Q = OR.q_shakov(f, cv, cw, prefac=2.0, omega=0.81, Pr=2.0 / 3.0)
    assert np.abs(Q).max() < 1e-10 * np.abs(f).max()


def test_shakov_conserves():
    cv, cw = grid()
    f = bimodal(cv)
    Q = OR.q_shakov(f, cv, cw, prefac=1.7, omega=0.81, Pr=2.0 / 3.0)
    phi = OR.invariants(cv)
    m = phi @ Q * cw
    scale = np.abs(phi) @ np.abs(Q) * cw
    assert np.abs(m / scale).max() < 1e-11


def test_shakov_collision_frequency_scaling():
    # Q_S(f) carries rho T^{1-omega}: doubling rho at fixed shape doubles the
    # frequency and the amplitude (Q(2f) = 4 Q(f)).
    cv, cw = grid()
    f = bimodal(cv)
    Q1 = OR.q_shakov(f, cv, cw, 1.0, 0.7, 2.0 / 3.0)
    Q2 = OR.q_shakov(2.0 * f, cv, cw, 1.0, 0.7, 2.0 / 3.0)
    assert np.abs(Q2 - 4.0 * Q1).max() < 1e-12 * np.abs(Q2).max()


def test_projection_properties():
    rng = np.random.default_rng(5)
    cv, _ = grid(N=12, L=5.0)
    phi = OR.invariants(cv)
    Q = rng.standard_normal(cv.shape[1])
    Qc = OR.project_conservative(Q, cv)
    assert np.abs(phi @ Qc).max() < 1e-12 * np.abs(phi).max() * np.abs(Q).sum()
    # the correction lies in span(phi)
    d = Q - Qc
    lam, *_ = np.linalg.lstsq(phi.T, d, rcond=None)
    assert np.abs(phi.T @ lam - d).max() < 1e-12 * np.abs(d).max()
    # idempotent, annihilates the span, keeps the orthogonal complement
    assert np.abs(OR.project_conservative(Qc, cv) - Qc).max() < 1e-12 * np.abs(Qc).max()
    span = phi.T @ rng.standard_normal(5)
    assert np.abs(OR.project_conservative(span, cv)).max() < 1e-10 * np.abs(span).max()
    # minimal norm: any other conservative candidate is farther from Q
    other = Qc + 1e-3 * OR.project_conservative(rng.standard_normal(Q.shape), cv)
    assert np.linalg.norm(Q - other) > np.linalg.norm(Q - Qc)


def test_bgk_pins():
    """oracle.relax.q_bgk ('bgk-direct-gll'): a Maxwellian is a fixed point;
    the target f + Q/nu keeps the node's rho, u and T (independent moment
    loop) and has zero heat flux; mass, momentum and energy are conserved;
    Q(2f) = 4 Q(f); and the term equals Shakov's at Pr = 1 (the two oracle
    restatements, from different reference kernels, agree)."""
    cv, cw = grid()
    m = maxw(cv, 1.3, (0.2, -0.1, 0.05), 0.9)
    assert np.abs(OR.q_bgk(m, cv, cw, 2.0, 0.81)).max() < 1e-10 * np.abs(m).max()
    f = bimodal(cv)
    rho, u, T, q = moments_of(f, cv, cw)
    assert np.abs(q).max() > 1e-2
    nu = rho * T ** (1.0 - 0.81)
    target = f + OR.q_bgk(f, cv, cw, 1.0, 0.81) / nu
    rB, uB, TB, qB = moments_of(target, cv, cw)
    assert abs(rB - rho) / rho < 1e-11 and np.abs(uB - u).max() < 1e-11
    assert abs(TB - T) / T < 1e-11 and np.abs(qB).max() < 1e-10 * np.abs(q).max()
    Q = OR.q_bgk(f, cv, cw, 1.7, 0.81)
    phi = OR.invariants(cv)
    assert np.abs((phi @ Q) / (np.abs(phi) @ np.abs(Q))).max() < 1e-11
    Q2 = OR.q_bgk(2.0 * f, cv, cw, 1.7, 0.81)
    assert np.abs(Q2 - 4.0 * Q).max() < 1e-12 * np.abs(Q2).max()
    QS = OR.q_shakov(f, cv, cw, 1.7, 0.81, 1.0)
    assert np.abs(Q - QS).max() < 1e-13 * np.abs(Q).max()


def _dev(X):
    return X - np.trace(X) / 3.0 * np.eye(3)


def _stress(g, cv, cw, u):
    C = cv - np.asarray(u)[:, None]
    return np.array([[(g * C[i] * C[j]).sum() for j in range(3)] for i in range(3)]) * cw


def test_esbgk_pins():
    """oracle.relax.q_esbgk ('esbgk-direct-gll'): a Maxwellian is a fixed point;
    at Pr = 1 it is the BGK term; mass, momentum and energy are conserved; and
    the deviatoric stress relaxes at nu/Pr (the model's Prandtl number), the
    fact the (1 - 1/Pr, 1/Pr) weights of the tensor fix: a planted swap of the
    two weights fails it."""
    cv, cw = grid()
    Pr, om, pf = 2.0 / 3.0, 0.81, 1.7
    m = maxw(cv, 1.3, (0.2, -0.1, 0.05), 0.9)
    assert np.abs(OR.q_esbgk(m, cv, cw, 2.0, om, Pr)).max() < 1e-10 * np.abs(m).max()
    f = bimodal(cv)
    QB = OR.q_bgk(f, cv, cw, pf, om)
    assert np.abs(OR.q_esbgk(f, cv, cw, pf, om, 1.0) - QB).max() < 1e-10 * np.abs(QB).max()
    Q = OR.q_esbgk(f, cv, cw, pf, om, Pr)
    phi = OR.invariants(cv)
    assert np.abs((phi @ Q) / (np.abs(phi) @ np.abs(Q))).max() < 1e-11
    rho, u, T, _ = moments_of(f, cv, cw)
    nu = pf * rho * T ** (1.0 - om)
    Sf, SQ = _dev(_stress(f, cv, cw, u)), _dev(_stress(Q, cv, cw, u))
    assert np.abs(Sf).max() > 1e-2 * rho * T                     # anisotropic stress
    assert np.abs(SQ + nu / Pr * Sf).max() < 1e-9 * nu * np.abs(Sf).max()

    def swapped(f, Pr):
        s0 = f.sum()
        uu = np.array([(f * cv[i]).sum() for i in range(3)]) / s0
        r = s0 * cw
        C = cv - uu[:, None]
        C2 = (C * C).sum(0)
        TT = (f * C2).sum() * cw / (1.5 * r)
        fe = r / (np.pi * TT) ** 1.5 * np.exp(-C2 / TT)
        w = (1.0 / Pr) * f + (1.0 - 1.0 / Pr) * fe
        L = np.array([[(C[i] * C[j] * w).sum() for j in range(3)] for i in range(3)]) * cw / r
        G = r / ((2 * np.pi) ** 1.5 * np.sqrt(np.linalg.det(L))) * \
            np.exp(-0.5 * np.einsum('iv,ij,jv->v', C, np.linalg.inv(L), C))
        return pf * r * TT ** (1.0 - om) * (G - f)
    SX = _dev(_stress(swapped(f, Pr), cv, cw, u))
    assert np.abs(SX + nu / Pr * Sf).max() > 1e-2 * nu * np.abs(Sf).max()


def test_bgk_conservative_pins():
    """oracle.relax.bgk_conservative_target ('bgk-gll'): its discrete mass,
    momentum and energy equal f's to round-off (the defining equations, on a
    grid too coarse for the plain Maxwellian to conserve them); a discrete
    Maxwellian of that exact form is returned unchanged; on a resolved grid it
    tends to the plain BGK target."""
    cv, cw = grid(N=12, L=4.0)                         # coarse: the plain Maxwellian leaks
    f = bimodal(cv)
    fe, rho, T = OR.bgk_conservative_target(f, cv, cw)
    phi = OR.invariants(cv)
    mf, me = phi @ f * cw, phi @ fe * cw
    assert np.abs(me - mf).max() < 1e-12 * np.abs(mf).max()
    Qb = OR.q_bgk(f, cv, cw, 1.0, 0.81)
    plain = phi @ (f + Qb / (rho * T ** 0.19)) * cw
    assert np.abs(plain - mf).max() > 1e-7 * np.abs(mf).max()   # the direct target does not
    C = cv - np.array([0.1, -0.2, 0.05])[:, None]
    g = 0.8 * np.exp(-1.3 * (C * C).sum(0) + 0.2 * C[0] - 0.1 * C[2])
    ge, _, _ = OR.bgk_conservative_target(g, cv, cw)
    assert np.abs(ge - g).max() < 1e-10 * g.max()
    cv2, cw2 = grid()
    f2 = bimodal(cv2)
    Q1 = OR.q_bgk_conservative(f2, cv2, cw2, 1.7, 0.81)
    Q2 = OR.q_bgk(f2, cv2, cw2, 1.7, 0.81)
    assert np.abs(Q1 - Q2).max() < 1e-8 * np.abs(Q2).max()


def test_bgk_conservative_negative_node():
    """A task's f may enter with coefficient -1 (basis U/V terms): the
    conservative BGK term is quadratic, Q(-f) = Q(f), and the target of -f is
    minus the target of f."""
    cv, cw = grid(N=12, L=4.0)
    f = bimodal(cv)
    fe, _, _ = OR.bgk_conservative_target(f, cv, cw)
    fm, _, _ = OR.bgk_conservative_target(-f, cv, cw)
    assert np.abs(fm + fe).max() < 1e-12 * fe.max()
    Q1 = OR.q_bgk_conservative(f, cv, cw, 1.3, 0.81)
    Q2 = OR.q_bgk_conservative(-f, cv, cw, 1.3, 0.81)
    assert np.abs(Q1 - Q2).max() < 1e-12 * np.abs(Q1).max()


def test_esbgk_conservative_pins():
    """oracle.relax.esbgk_conservative_target ('esbgk-gll'): the Gaussian's
    mass, momentum and second moments equal rho, rho u and rho (u u + T_ij)
    of the ES tensor (a strongly anisotropic node); at Pr = 1 the tensor is
    the conservative BGK target's and the Gaussian is that target; the term
    is quadratic, Q(-f) = Q(f)."""
    cv, cw = grid(N=16, L=5.0)
    C = cv - np.array([0.2, -0.1, 0.05])[:, None]
    f = np.exp(-C[0] ** 2 / 0.7 - C[1] ** 2 / 1.3 - C[2] ** 2 / 1.0 - 0.3 * C[0] * C[1])
    Pr = 2.0 / 3.0
    G, rho, T = OR.esbgk_conservative_target(f, cv, cw, Pr)
    fe, _, _ = OR.bgk_conservative_target(f, cv, cw)
    u = np.array([(f * cv[i]).sum() for i in range(3)]) / f.sum()
    Cu = cv - u[:, None]
    w = (1 - 1 / Pr) * f + (1 / Pr) * fe
    for i, j in [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]:
        want = rho * u[i] * u[j] + (Cu[i] * Cu[j] * w).sum() * cw
        got = (cv[i] * cv[j] * G).sum() * cw
        assert abs(got - want) < 1e-11 * rho * T, (i, j)
    assert abs(G.sum() * cw - rho) < 1e-12 * rho
    assert np.abs(cv @ G * cw - rho * u).max() < 1e-11 * rho
    G1, _, _ = OR.esbgk_conservative_target(f, cv, cw, 1.0)
    assert np.abs(G1 - fe).max() < 1e-9 * fe.max()
    Q1 = OR.q_esbgk_conservative(f, cv, cw, 1.2, 0.81, Pr)
    Q2 = OR.q_esbgk_conservative(-f, cv, cw, 1.2, 0.81, Pr)
    assert np.abs(Q1 - Q2).max() < 1e-10 * np.abs(Q1).max()
