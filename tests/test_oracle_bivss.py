"""Pins for the binary-mixture VSS oracle (oracle/bivss.py).

The reference ships no golden vectors for the mixture operator, so the oracle
is pinned by what the mathematics fixes:
  * equal masses, alpha = 1: the pre-rule kernels b, c are quadratures of the
    closed-form sinc kernels of the single-species operator, and the whole
    operator converges to oracle/collision.py's q_vhs as Mpre grows;
  * a common Maxwellian of two unequal masses is an equilibrium of every pair
    (Q_pq small, and much smaller than with the mass fractions exchanged);
  * the pair (p, q) exchanges momentum and energy without creating either,
    and each species keeps its mass.
"""
import numpy as np
import pytest

from oracle import collision as OC
from oracle import quadratures as Q
from oracle.bivss import BiPrecomp, BiVelocityMesh, maxwellian, q_pq, vss_prefactor
from oracle.velocitymesh import load_design


def _radial(L, Nrho):
    R = 2 * L * 2 / (3 + np.sqrt(2))
    z, w = Q.zwgj(Nrho, 0.0, 0.0)
    return R / 2 * (1 + z), R / 2 * w


def _grid(N, L):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing='ij')
    return np.stack([c0[I.ravel()], c0[J.ravel()], c0[K.ravel()]])


def test_equal_mass_hard_sphere_converges_to_single_species():
    N, L, Nrho = 12, 4.0, 6
    qz, qw = _radial(L, Nrho)
    full = load_design('womersley', 12)          # antipodal: full[6:] = -full[:6]
    assert np.allclose(full[6:], -full[:6])
    pre_s = OC.Precomp(N, L, qz, qw, full[:6], 2 * np.pi / 6, 1.0)
    cv = _grid(N, L)
    f = maxwellian(cv, 1.0, 1.0, [0.3, 0, 0], 0.9) * (1 + 0.2 * np.cos(cv[1]))
    g = maxwellian(cv, 1.0, 0.7, [-0.2, 0.1, 0], 1.2)
    qs = OC.q_vhs(pre_s, f, g, 1.0)
    errs = []
    for Mpre in (32, 70, 192):
        pre_b = BiPrecomp(N, L, qz, qw, full, 4 * np.pi / 12, load_design('womersley', Mpre),
                          1.0, 1.0, 1.0, 0.5)
        errs.append(np.abs(q_pq(pre_b, f, g, 1.0) - qs).max() / np.abs(qs).max())
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5


def test_equal_mass_prefactor_matches_single_species():
    n0, H0, T0, om, d, Tr = 2.0e20, 1e-3, 273.0, 0.81, 4.11e-10, 273.0
    pb, kb = vss_prefactor([1.0, 1.0], 0, 1, n0, H0, T0, 1.0, om, d, Tr)
    ps, ks = OC.vhs_prefactor(H0, n0, T0, om, d, Tr)
    assert pb == pytest.approx(ps, rel=1e-14)
    assert kb == pytest.approx(ks, rel=1e-14)


@pytest.fixture(scope='module')
def mixture():
    # argon / krypton-like mass ratio of the reference's Fourier example
    vm = BiVelocityMesh(1e-3, 273.0, 1.1140244750324529e-4, 39.96e-3, 16, 8, 0.0, 273.0, 4.0,
                        [39.96e-3, 83.77e-3], 12)
    szpre = load_design('womersley', 192)
    m = vm.masses
    params = {(0, 0): (1.4, 0.81), (0, 1): (1.36, 0.805), (1, 0): (1.36, 0.805),
              (1, 1): (1.32, 0.80)}
    pres = {k: BiPrecomp(16, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, szpre, m[k[0]], m[k[1]], *v)
            for k, v in params.items()}
    return vm, pres


def test_mesh_follows_mass_ratio(mixture):
    vm, _ = mixture
    mr = vm.masses[1]
    S1 = vm.L * 2 / (max(4 * mr / (1 + mr), 2) + np.sqrt(1 + mr) + 1)
    S2 = vm.L * 2 / (max(4 / (1 + mr), 2) + np.sqrt(1 + 1 / mr) + 1)
    assert vm.R == pytest.approx(2 * min(S1, S2))
    assert vm.sw == pytest.approx(4 * np.pi / 12)
    assert np.allclose(np.linalg.norm(vm.sz, axis=1), 1.0)


def test_common_maxwellian_is_equilibrium(mixture):
    vm, pres = mixture
    m, cv = vm.masses, vm.cv
    eq = [maxwellian(cv, m[0], 1.0, [0.1, 0, 0], 1.3), maxwellian(cv, m[1], 0.5, [0.1, 0, 0], 1.3)]
    neq = [maxwellian(cv, m[0], 1.0, [0.3, 0, 0], 0.8),
           maxwellian(cv, m[1], 0.5, [-0.2, 0.1, 0], 1.3)]
    for (p, q), pre in pres.items():
        qe = np.abs(q_pq(pre, eq[p], eq[q], 1.0)).max()
        assert qe < 2e-2 * eq[p].max(), (p, q)
        if p != q:
            qn = np.abs(q_pq(pre, neq[p], neq[q], 1.0)).max()
            assert qe < 2e-2 * qn, (p, q, qe, qn)
    # exchanging the two mass fractions (a plausible index slip) breaks it
    pre = pres[0, 1]
    swapped = type('P', (), dict(N=pre.N, n=pre.n, a=pre.a, mp=pre.mq, mq=pre.mp, Nrho=pre.Nrho,
                                 M=pre.M, qw=pre.qw, sw=pre.sw, b=pre.b, c=pre.c))()
    qs = np.abs(q_pq(swapped, eq[0], eq[1], 1.0)).max()
    assert qs > 50 * np.abs(q_pq(pre, eq[0], eq[1], 1.0)).max()


def test_pair_conserves_momentum_energy_and_species_mass(mixture):
    vm, pres = mixture
    m, cv = vm.masses, vm.cv
    f = [maxwellian(cv, m[0], 1.0, [0.3, 0, 0], 0.8), maxwellian(cv, m[1], 0.5, [-0.2, 0.1, 0], 1.3)]
    q01 = q_pq(pres[0, 1], f[0], f[1], 1.0)
    q10 = q_pq(pres[1, 0], f[1], f[0], 1.0)
    c2 = (cv ** 2).sum(0)
    # each side transfers O(1) momentum and energy; the sums cancel
    mom01, mom10 = m[0] * (cv * q01).sum(1), m[1] * (cv * q10).sum(1)
    scale_m = np.abs(mom01).max()
    assert scale_m > 1.0
    assert np.abs(mom01 + mom10).max() < 5e-3 * scale_m
    en01, en10 = m[0] * (c2 * q01).sum(), m[1] * (c2 * q10).sum()
    assert abs(en01 + en10) < 5e-3 * m[0] * np.abs(c2 * q01).sum()
    for qq in (q01, q10):
        assert abs(qq.sum()) < 1e-3 * np.abs(qq).sum()


def test_mixture_projection_pins(mixture):
    """oracle.relax.project_mixture: the six mixture constraints vanish,
    projecting twice changes nothing, the correction lies in the span of the
    constraint rows (least norm), and on real collision terms it is a small
    change.  A planted slip (species 0's mass used in both momentum rows)
    leaves the mixture momentum unbalanced."""
    from oracle.relax import project_mixture
    vm, pres = mixture
    m, cv = vm.masses, vm.cv
    f = [maxwellian(cv, m[0], 1.0, [0.3, 0, 0], 0.8), maxwellian(cv, m[1], 0.5, [-0.2, 0.1, 0], 1.3)]
    Q0 = q_pq(pres[0, 0], f[0], f[0], 1.0) + q_pq(pres[0, 1], f[0], f[1], 1.0)
    Q1 = q_pq(pres[1, 0], f[1], f[0], 1.0) + q_pq(pres[1, 1], f[1], f[1], 1.0)
    P0, P1 = project_mixture(Q0, Q1, cv, m[0], m[1])
    c2 = (cv ** 2).sum(0)

    def resid(a, b, ma=m[0], mb=m[1]):
        return np.array([a.sum(), b.sum(), *(ma * (cv * a).sum(1) + mb * (cv * b).sum(1)),
                         ma * (c2 * a).sum() + mb * (c2 * b).sum()])

    scale = np.abs(resid(np.abs(Q0), np.abs(Q1))).max()
    assert np.abs(resid(P0, P1)).max() < 1e-13 * scale
    assert np.abs(resid(Q0, Q1)).max() > 1e-6 * scale           # the raw terms do not conserve
    R0, R1 = project_mixture(P0, P1, cv, m[0], m[1])
    assert np.abs(R0 - P0).max() < 1e-13 * np.abs(P0).max()
    assert np.abs(R1 - P1).max() < 1e-13 * np.abs(P1).max()
    # least norm: the correction is a combination of the constraint rows
    n = cv.shape[1]
    A = np.zeros((6, 2 * n))
    phi = np.vstack([np.ones(n), cv, c2])
    A[0, :n], A[1, n:] = 1.0, 1.0
    A[2:, :n], A[2:, n:] = m[0] * phi[1:], m[1] * phi[1:]
    d = np.concatenate([P0 - Q0, P1 - Q1])
    y, *_ = np.linalg.lstsq(A.T, d, rcond=None)
    assert np.abs(A.T @ y - d).max() < 1e-10 * np.abs(d).max()
    assert np.abs(d).max() < 1e-2 * max(np.abs(Q0).max(), np.abs(Q1).max())
    # planted slip: species 0's mass on both momentum rows
    S0, S1 = project_mixture(Q0, Q1, cv, m[0], m[0])
    assert np.abs(resid(S0, S1)[2:5]).max() > 1e3 * np.abs(resid(P0, P1)[2:5]).max()


def test_antipodal_rules_give_finite_kernels():
    """The same antipodal design as both the M rule and the pre-rule: the
    exactly opposite pairs give (1 + s.s') = 0 (possibly rounded below zero);
    the kernels stay finite."""
    full = load_design('womersley', 12)
    qz, qw = _radial(4.0, 3)
    pre = BiPrecomp(8, 4.0, qz, qw, full, 4 * np.pi / 12, full, 1.0, 2.0, 1.4, 0.8)
    assert np.isfinite(pre.b).all() and np.isfinite(pre.c).all()
