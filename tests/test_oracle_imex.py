"""Pins of the IMEX oracle (CPU): closed-form moments of Maxwellians, the
central heat flux of the 14-moment normalisation, and the convergence order
of every ARS and linear-multistep tableau on a relaxation problem with an
exact solution (explicit part -mu (f - M0), implicit Shakov/BGK part
nu (M0 - f) with Pr = 1, so that f(t) = M0 + (f0 - M0) exp(-(mu + nu) t))."""
import numpy as np
import pytest

from oracle import imex as OI


def grid(N=12, L=6.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(c0, c0, c0, indexing='ij')
    return np.stack([I.ravel(), J.ravel(), K.ravel()]), (2 * L / N) ** 3


def maxw(cv, rho, u, T):
    u = np.asarray(u, float).reshape(3, 1)
    return rho / (np.pi * T) ** 1.5 * np.exp(-((cv - u) ** 2).sum(0) / T)


def test_moments_of_maxwellian_closed_form():
    cv, cw = grid(24, 7.0)
    rho, u, T = 1.3, np.array([0.3, -0.2, 0.1]), 0.9
    U = OI.moment(maxw(cv, rho, u, T)[None, :], cv, cw)[0]
    u2 = (u * u).sum()
    exp = np.concatenate([[rho], rho * u, [rho * (u2 + 1.5 * T)],
                          rho * (u * u + T / 2), rho * np.array([u[0] * u[1], u[1] * u[2],
                                                                 u[2] * u[0]]),
                          rho * u * (u2 + 2.5 * T)])
    assert np.abs(U - exp).max() < 1e-8 * np.abs(exp).max()      # grid quadrature of the tails
    r, uu, TT, q = OI.momentnorm(U[None, :])
    assert abs(r[0] - rho) < 1e-8 and np.abs(uu[0] - u).max() < 1e-8
    assert abs(TT[0] - T) < 1e-8 and np.abs(q[0]).max() < 1e-8
    # the target of a Maxwellian's moments is that Maxwellian, for any Pr
    for Pr in (2.0 / 3.0, 1.0):
        M = OI.target(U[None, :], cv, Pr)[0]
        assert np.abs(M - maxw(cv, rho, u, T)).max() < 1e-7 * M.max()


def test_heat_flux_normalisation_is_central():
    cv, cw = grid(24, 7.0)
    f = 0.6 * maxw(cv, 1.0, (0.3, 0.1, 0.0), 0.8) + 0.5 * maxw(cv, 0.9, (-0.2, 0.25, 0.1), 1.3)
    U = OI.moment(f[None, :], cv, cw)
    rho, u, T, q = OI.momentnorm(U)
    C = cv - u[0][:, None]
    C2 = (C * C).sum(0)
    assert abs(rho[0] - f.sum() * cw) < 1e-13
    assert abs(T[0] - 2.0 / 3.0 * (f * C2).sum() * cw / rho[0]) < 1e-12
    qd = (f[None, :] * C2[None, :] * C).sum(1) * cw
    assert np.abs(q[0] - qd).max() < 1e-12


def _relax_problem():
    # fine enough that the discrete Maxwellian's moments equal its parameters
    # to ~1e-10, so the conserved moments and nu stay fixed along the solution
    cv, cw = grid(24, 7.0)
    f0 = 0.6 * maxw(cv, 1.0, (0.3, 0.1, 0.0), 0.8) + 0.5 * maxw(cv, 0.9, (-0.2, 0.25, 0.1), 1.3)
    U0 = OI.moment(f0[None, :], cv, cw)
    M0 = OI.target(U0, cv, 1.0)                       # Maxwellian of f0's conserved moments
    pm = OI.Penal(cv, cw, prefac=1.0, omega=0.5, Pr=1.0)
    nu = OI.nu_of(U0, 1.0, 0.5)[0]
    mu = 0.7
    explicit = lambda t, f: -mu * (f - M0)
    exact = lambda t: M0 + (f0[None, :] - M0) * np.exp(-(mu + nu) * t)
    return pm, explicit, exact, f0[None, :], nu


@pytest.mark.parametrize('kind,order', [('ars-111', 1), ('ars-222', 2), ('ars-443', 3)])
def test_ars_order(kind, order):
    pm, explicit, exact, f0, nu = _relax_problem()
    tend = 0.4 / nu
    errs = []
    for nsteps in (8, 16):
        dt, f = tend / nsteps, f0.copy()
        for k in range(nsteps):
            f = OI.ars_step(kind, pm, explicit, k * dt, dt, f, penalize=False)
        errs.append(np.abs(f - exact(tend)).max())
    rate = np.log2(errs[0] / errs[1])
    print(kind, errs, rate)
    assert rate > order - 0.3


@pytest.mark.parametrize('kind,order', [('imex-lm-bdf1', 1), ('imex-lm-cn2', 2),
                                        ('imex-lm-ad2', 2), ('imex-lm-bdf2', 2),
                                        ('imex-lm-sg2', 2), ('imex-lm-bdf3', 3),
                                        ('imex-lm-ad3', 3), ('imex-lm-tvb3', 3),
                                        ('imex-lm-bdf4', 4)])
def test_lm_order(kind, order):
    pm, explicit, exact, f0, nu = _relax_problem()
    tend = 0.5 / nu
    errs = []
    for nsteps in (16, 32):
        dt = tend / nsteps
        it = OI.LMIntegrator(kind, pm, explicit, None, 0.0, dt, f0.copy(), penalize_init=False)
        f = f0.copy()
        for k in range(nsteps):
            f = it.step(k * dt, dt, k, f)
        errs.append(np.abs(f - exact(tend)).max())
    rate = np.log2(errs[0] / errs[1])
    print(kind, errs, rate)
    assert rate > order - 0.3


def test_planted_lm_coefficient_fails():
    """A wrong BDF2 coefficient (4/3 -> 5/4) destroys the order (mutant
    planted in a copy of the oracle's table)."""
    pm, explicit, exact, f0, nu = _relax_problem()
    saved = OI.LM['imex-lm-bdf2']
    OI.LM['imex-lm-bdf2'] = ([-5. / 4., 1. / 3.], [4. / 3., -2. / 3.], [2. / 3., 0., 0.])
    try:
        tend = 0.5 / nu
        errs = []
        for nsteps in (16, 32):
            dt = tend / nsteps
            it = OI.LMIntegrator('imex-lm-bdf2', pm, explicit, None, 0.0, dt, f0.copy(),
                                 penalize_init=False)
            f = f0.copy()
            for k in range(nsteps):
                f = it.step(k * dt, dt, k, f)
            errs.append(np.abs(f - exact(tend)).max())
        assert not np.log2(errs[0] / errs[1]) > 1.7
    finally:
        OI.LM['imex-lm-bdf2'] = saved


@pytest.mark.parametrize('kind,order', [('sspl-222', 2), ('sspl-443', 3)])
def test_sspl_order(kind, order):
    pm, explicit, exact, f0, nu = _relax_problem()
    tend = 0.4 / nu
    errs = []
    for nsteps in (8, 16):
        dt, f = tend / nsteps, f0.copy()
        for k in range(nsteps):
            f = OI.sspl_step(kind, pm, explicit, k * dt, dt, f)
        errs.append(np.abs(f - exact(tend)).max())
    rate = np.log2(errs[0] / errs[1])
    print(kind, errs, rate)
    assert rate > order - 0.3


@pytest.mark.parametrize('kind,order', [('bdf-111', 1), ('bdf-222', 2), ('bdf-333', 3)])
def test_bdf_order(kind, order):
    pm, explicit, exact, f0, nu = _relax_problem()
    tend = 0.5 / nu
    errs = []
    for nsteps in (16, 32):
        dt = tend / nsteps
        it = OI.BDFIntegrator(kind, pm, explicit)
        it.nsub = 20
        f = f0.copy()
        for k in range(nsteps):
            f = it.step(k * dt, dt, k, f)
        errs.append(np.abs(f - exact(tend)).max())
    rate = np.log2(errs[0] / errs[1])
    print(kind, errs, rate)
    assert rate > order - 0.3


@pytest.mark.parametrize('kind', ['bdf-111', 'bdf-222', 'bdf-333'])
def test_bdf_is_a_linear_multistep_update(kind):
    """The BDF-k update equals the linear-multistep update with
    a = A/A_q, b = G/A_q (newest first) and c = (B/A_q, 0, ...) and no
    penalisation terms: the mapping the product's bdf integrators use."""
    rng = np.random.default_rng(3)
    cv, cw = grid(8, 5.0)
    pm = OI.Penal(cv, cw, prefac=0.8, omega=0.7, Pr=2. / 3.)
    A, G, B = OI.BDF[kind]
    q = len(G)
    f = [maxw(cv, 1.0 + 0.1 * k, (0.1 * k, -0.05, 0.0), 0.9 + 0.05 * k)[None, :] *
         (1 + 0.05 * rng.uniform(-1, 1, cv.shape[1])) for k in range(q)]
    L = [0.1 * rng.uniform(-1, 1, f[0].shape) * f[0] for _ in range(q)]
    U = [pm.moment(x) for x in f]
    LU = [pm.moment(x) for x in L]
    dt = 0.05
    a = [A[q - 1 - k] / A[q] for k in range(q)]
    b = [G[q - 1 - k] / A[q] for k in range(q)]
    c = [B / A[q]] + [0.0] * q
    zeroU = [np.zeros_like(x) for x in U]
    Ub = OI.bdf_moments(dt, A, G, B, U, LU, pm.prefac, pm.omega, pm.Pr)
    Ul = OI.lm_moments(dt, a, b, c, LU, zeroU, U, q, pm.prefac, pm.omega, pm.Pr)
    assert np.abs(Ub - Ul).max() < 1e-13 * np.abs(Ub).max()
    Mn = pm.target(Ub)
    fb = OI.bdf_dist(dt, A, G, B, f, L, Mn, Ub, pm.prefac, pm.omega)
    fl = OI.lm_dist(dt, a, b, c, L, [np.zeros_like(x) for x in f], f, Mn, Ub, q, pm.prefac,
                    pm.omega)
    assert np.abs(fb - fl).max() < 1e-13 * np.abs(fb).max()


def _es_grid(N=24, L=6.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(c0, c0, c0, indexing='ij')
    return np.stack([I.ravel(), J.ravel(), K.ravel()]), (2 * L / N) ** 3


def test_es_target_pins():
    """oracle.imex.es_target (astd esbgk-direct-gll): a Maxwellian's moments
    give that Maxwellian back; for anisotropic moments the target keeps rho,
    rho u and E, and its stress deviator is (1 - 1/Pr) times the input's."""
    cv, cw = _es_grid()
    Pr = 2.0 / 3.0
    u = np.array([0.2, -0.1, 0.05])
    C = cv - u[:, None]
    m = 1.3 / (np.pi * 0.9) ** 1.5 * np.exp(-(C * C).sum(0) / 0.9)
    Um = OI.moment(m[None], cv, cw)
    assert np.abs(OI.es_target(Um, cv, Pr)[0] - m).max() < 1e-10 * m.max()
    Tx, Ty, Tz = 0.8, 1.3, 1.0
    f = 1.1 / (np.pi ** 1.5 * np.sqrt(Tx * Ty * Tz)) * np.exp(
        -C[0] ** 2 / Tx - C[1] ** 2 / Ty - C[2] ** 2 / Tz - 0.3 * C[0] * C[1])
    U = OI.moment(f[None], cv, cw)
    G = OI.es_target(U, cv, Pr)
    UG = OI.moment(G, cv, cw)
    assert np.abs(UG[0, :5] - U[0, :5]).max() < 1e-10 * np.abs(U[0, :5]).max()

    def dev(V):
        r = V[0, 0]
        uu = V[0, 1:4] / r
        P = np.array([[V[0, i] for i in (5, 8, 10)], [V[0, i] for i in (8, 6, 9)],
                      [V[0, i] for i in (10, 9, 7)]]) - r * np.outer(uu, uu)
        return P - np.trace(P) / 3 * np.eye(3)
    assert np.abs(dev(U)).max() > 0.05
    assert np.abs(dev(UG) - (1 - 1 / Pr) * dev(U)).max() < 1e-9 * np.abs(dev(U)).max()


def test_es_relaxation_rate():
    """Space-homogeneous ES-BGK relaxation (explicit part 0): rho, u, T stay
    fixed and the stress deviator decays as exp(-nu t / Pr) (the model's
    Prandtl number).  ARS-222 converges to that closed form at second order;
    with the moment stresses relaxing at nu (srate 1, the Shakov rate) the
    error is an order of magnitude larger."""
    cv, cw = _es_grid()
    Pr = 2.0 / 3.0
    C = cv
    f0 = 1.0 / (np.pi ** 1.5 * np.sqrt(0.8 * 1.3 * 0.9)) * np.exp(
        -C[0] ** 2 / 0.8 - C[1] ** 2 / 1.3 - C[2] ** 2 / 0.9)
    U0 = OI.moment(f0[None], cv, cw)
    nu = U0[0, 0] * OI.temperature(U0)[0] ** (1.0 - 0.81)
    zero = lambda t, f: np.zeros_like(f)

    def dev(V):
        P = np.diag(V[0, 5:8]) - V[0, 5:8].sum() / 3 * np.eye(3)
        return P
    tend = 1.0
    exact = dev(U0) * np.exp(-nu * tend / Pr)
    errs = []
    for n in (10, 20, 40):
        pm = OI.Penal(cv, cw, 1.0, 0.81, Pr, model='esbgk')
        f = f0[None].copy()
        for k in range(n):
            f = OI.ars_step('ars-222', pm, zero, k * tend / n, tend / n, f, penalize=False)
        errs.append(np.abs(dev(OI.moment(f, cv, cw)) - exact).max())
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5
    assert errs[2] < 1e-3 * np.abs(dev(U0)).max()
    pm = OI.Penal(cv, cw, 1.0, 0.81, Pr, model='esbgk')
    pm.srate = 1.0
    f = f0[None].copy()
    for k in range(40):
        f = OI.ars_step('ars-222', pm, zero, k * tend / 40, tend / 40, f, penalize=False)
    # the distribution follows the target, so the slip shows as a first-order
    # error (1e-4 here) instead of the scheme's second-order one (3e-6)
    assert np.abs(dev(OI.moment(f, cv, cw)) - exact).max() > 10 * errs[2]


# ------------------------------------------------ binary-mixture penalisation
def _species_maxw(cv, m, n, u, T):
    C = cv - np.asarray(u, float).reshape(3, 1)
    return n * (m / (np.pi * T)) ** 1.5 * np.exp(-m * (C * C).sum(0) / T)


def _species_moments(h, cv, cw, m):
    n = (h * cw).sum()
    u = (cv * h).sum(1) * cw / n
    C = cv - u[:, None]
    C2 = (C * C).sum(0)
    return n, u, 2.0 / 3.0 * m * (h * C2).sum() * cw / n, (h * C2 * C).sum(1) * cw


@pytest.mark.parametrize('m', [1.0, 2.1])
def test_species_target_moments(m):
    """The species' Shakov target (mass m): a Maxwellian of mass m is its
    own target; for a state with heat flux q the target keeps n, u and T and
    carries (1 - Pr) q -- the Shakov model's defining moments -- with the
    temperature and heat flux of a mass-m particle (a wrong power of m fails)."""
    cv, cw = grid(28, 6.0)
    n, u, T = 0.8, np.array([0.2, -0.1, 0.05]), 1.1
    f = _species_maxw(cv, m, n, u, T)
    sp = OI.SpeciesPenal(cv, cw, m, prefac=1.0, omega=0.8, Pr=2.0 / 3.0)
    assert np.abs(sp.target(sp.moment(f[None]))[0] - f).max() < 1e-6 * f.max()
    C = cv - u[:, None]
    g = f * (1.0 + (C * np.array([0.05, 0.02, -0.03])[:, None]).sum(0)
             * (m * (C * C).sum(0) / T - 2.5))
    S = sp.target(sp.moment(g[None]))[0]
    a, b = _species_moments(g, cv, cw, m), _species_moments(S, cv, cw, m)
    assert abs(a[0] - b[0]) < 1e-10 and np.abs(a[1] - b[1]).max() < 1e-10
    assert abs(a[2] - b[2]) < 1e-10 and abs(a[2] - T) < 1e-6
    assert np.abs(b[3] - (1.0 - 2.0 / 3.0) * a[3]).max() < 1e-8 * np.abs(a[3]).max()
    if m == 1.0:                      # unit mass: the single-species target
        assert np.abs(OI.target(OI.moment(g[None], cv, cw), cv, 2.0 / 3.0)[0] - S).max() \
            < 1e-14 * S.max()


def test_vss_viscosity_against_tabulated_values():
    """Bird (1994), Table A1 at 273 K: argon 2.117e-5 Pa s (VHS d = 4.17e-10,
    omega = 0.81), krypton 2.328e-5 (VHS d = 4.76e-10, omega = 0.80); the VSS
    argon parameters (alpha = 1.40, d = 4.11e-10) give the same viscosity."""
    NA = 6.0221409e23
    mAr, mKr = 39.948e-3 / NA, 83.80e-3 / NA
    assert abs(OI.vss_viscosity(mAr, 1.0, 0.81, 4.17e-10, 273.0, 273.0) / 2.117e-5 - 1) < 2e-3
    assert abs(OI.vss_viscosity(mKr, 1.0, 0.80, 4.76e-10, 273.0, 273.0) / 2.328e-5 - 1) < 2e-3
    assert abs(OI.vss_viscosity(mAr, 1.40, 0.81, 4.11e-10, 273.0, 273.0) / 2.117e-5 - 1) < 2e-3
    # (T/Tref)^omega
    r = OI.vss_viscosity(mAr, 1.0, 0.81, 4.17e-10, 273.0, 546.0) / \
        OI.vss_viscosity(mAr, 1.0, 0.81, 4.17e-10, 273.0, 273.0)
    assert abs(r - 2.0 ** 0.81) < 1e-12


@pytest.mark.parametrize('kind,order', [('imex-lm-bdf2', 2), ('imex-lm-bdf3', 3)])
def test_mixture_relaxation_order(kind, order):
    """Two species (masses 1 and 2.1, different prefactors and omegas) relax
    by their own BGK penalisations (Pr = 1, no transport): n, u and T of each
    species are invariants, so nu_p is constant and
    f_p(t) = M_p + (f_p(0) - M_p) exp(-nu_p t) with M_p the mass-m_p Maxwellian
    of f_p(0)'s n, u, T.  The linear-multistep integrator on the mixture stack
    converges to it at the scheme's order (a wrong mass scaling of nu or of the
    target fails)."""
    cv, cw = grid(24, 6.0)      # resolves the heavier species' Maxwellian to 1e-8
    ms, pre, om = (1.0, 2.1), (1.3, 0.8), (0.81, 0.75)
    sps = [OI.SpeciesPenal(cv, cw, ms[p], pre[p], om[p], 1.0) for p in range(2)]
    pm = OI.MixPenal(sps, 1)
    f0, Ms, nus = [], [], []
    for p, (n, u, T) in enumerate(((1.0, (0.1, 0, 0), 1.0), (0.4, (-0.05, 0.1, 0), 1.2))):
        f = _species_maxw(cv, ms[p], n, u, T)
        C = cv - np.asarray(u, float).reshape(3, 1)
        g = f * (1.0 + 0.2 * np.cos(C[0] + 0.5 * C[1]))
        nn, uu, TT, _ = _species_moments(g, cv, cw, ms[p])
        f0.append(g)
        Ms.append(_species_maxw(cv, ms[p], nn, uu, TT))
        nus.append(pre[p] * nn * TT ** (1.0 - om[p]))
    F0 = np.stack(f0)
    M = np.stack(Ms)
    nu = np.array(nus)[:, None]
    tend = 0.5 / max(nus)
    exact = M + (F0 - M) * np.exp(-nu * tend)
    zero = lambda t, f: np.zeros_like(f)
    errs = []
    for nsteps in (16, 32):
        dt = tend / nsteps
        it = OI.LMIntegrator(kind, pm, zero, None, 0.0, dt, F0.copy(), penalize_init=False)
        f = F0.copy()
        for k in range(nsteps):
            f = it.step(k * dt, dt, k, f)
        errs.append(np.abs(f - exact).max() / np.abs(exact).max())
    rate = np.log2(errs[0] / errs[1])
    print(kind, errs, rate)
    assert rate > order - 0.3 and errs[1] < 1e-3
