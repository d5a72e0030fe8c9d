"""Pins the FP64 collision oracle (oracle/collision.py) without the reference:

* pipeline == brute-force spectral sum (no FFT of products) at N = 4, 6, 8,
  for 'vhs' (bilinear) and 'vhs-gll' (complex squares) -- pins step order,
  1/n scalings, the aliased index arithmetic, cos/sin vs complex square;
* Q(f, f) of 'vhs' equals 'vhs-gll' exactly (t1 t2 + t3 t4 with f = g);
* bilinearity in (f, g) and the prefactor scaling;
* mass / momentum / energy of Q vanish to spectral accuracy, and Q(M, M) of a
  centred Maxwellian is small, on a grid that resolves the Maxwellian;
* reflection covariance c -> -c (Q(Rf, Rg) = R Q(f, g)).
A planted mistake (e.g. dropping the 1/n of the forward transform or the sign
of one phase) fails the brute-force comparison; see test_planted_mistake.
"""
import numpy as np
import pytest

from oracle import collision as OC
from oracle import velocitymesh as OVM


def _setup(N, M, Nrho, omega=0.81):
    vm = OVM.VelocityMesh(1e-3, 273.0, 2.3e-4, 39.96e-3, N, Nrho, 0.0, 273.0, 5.0, M)
    pre = OC.Precomp(N, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, 2 * (1 - omega))
    return vm, pre


def _f(vm, u, T, seed):
    rng = np.random.default_rng(seed)
    c = vm.cv - np.array(u)[:, None]
    return np.exp(-(c * c).sum(0) / T) * (1 + 0.05 * rng.uniform(-1, 1, vm.vsize))


@pytest.mark.parametrize('N,M,Nrho', [(4, 6, 3), (6, 3, 4), (8, 6, 4)])
def test_pipeline_equals_bruteforce(N, M, Nrho):
    vm, pre = _setup(N, M, Nrho)
    f = _f(vm, (0.3, 0, 0), 0.8, 1)
    g = _f(vm, (-0.3, 0.2, 0), 1.2, 2)
    ref = OC.q_bruteforce(pre, f, g, 0.7)
    assert np.abs(OC.q_vhs(pre, f, g, 0.7) - ref).max() < 1e-13 * np.abs(ref).max()
    refg = OC.q_bruteforce(pre, f, f, 0.7)
    assert np.abs(OC.q_vhs_gll(pre, f, 0.7) - refg).max() < 1e-13 * np.abs(refg).max()


def test_vhs_with_f_equal_g_is_vhs_gll():
    vm, pre = _setup(8, 6, 4)
    f = _f(vm, (0.2, 0.1, 0), 1.0, 3)
    assert np.array_equal(OC.q_vhs(pre, f, f, 1.0), OC.q_vhs_gll(pre, f, 1.0))


def test_bilinear_and_prefactor():
    vm, pre = _setup(8, 6, 4)
    f1, f2 = _f(vm, (0.2, 0, 0), 1.0, 4), _f(vm, (0, 0.3, 0), 0.9, 5)
    g = _f(vm, (0, 0, 0.1), 1.1, 6)
    lhs = OC.q_vhs(pre, 2.0 * f1 - 0.5 * f2, g, 1.0)
    rhs = 2.0 * OC.q_vhs(pre, f1, g, 1.0) - 0.5 * OC.q_vhs(pre, f2, g, 1.0)
    assert np.abs(lhs - rhs).max() < 1e-13 * np.abs(lhs).max()
    assert np.allclose(OC.q_vhs(pre, f1, g, 3.0), 3.0 * OC.q_vhs(pre, f1, g, 1.0), rtol=1e-14,
                       atol=0)


def test_conservation_and_equilibrium():
    vm, pre = _setup(24, 6, 12)
    c = vm.cv
    M0 = np.exp(-(c * c).sum(0) / 1.0) / np.pi ** 1.5
    f = 0.5 * np.exp(-((c - np.array([[0.5], [0], [0]])) ** 2).sum(0) / 0.8) \
        + 0.5 * np.exp(-((c + np.array([[0.5], [0], [0]])) ** 2).sum(0) / 1.2)
    Q = OC.q_vhs(pre, f, f, 1.0)
    scale = np.abs(Q).max() * vm.vsize * vm.cw
    for phi in (np.ones_like(c[0]), c[0], c[1], c[2], (c * c).sum(0)):
        assert abs(np.sum(Q * phi) * vm.cw) < 1e-3 * scale * max(1.0, np.abs(phi).max())
    Qeq = OC.q_vhs(pre, M0, M0, 1.0)
    assert np.abs(Qeq).max() < 2e-2 * np.abs(Q).max()


def test_reflection_covariance():
    vm, pre = _setup(6, 3, 4)
    N = vm.Nv
    f = _f(vm, (0.3, 0.1, -0.2), 0.9, 7)
    g = _f(vm, (-0.1, 0.2, 0.1), 1.1, 8)
    R = lambda h: h.reshape(N, N, N)[::-1, ::-1, ::-1].ravel()
    # cell-centred symmetric grid: c -> -c is index reversal; the half-sphere
    # set is antipodal, so the operator is covariant up to the Nyquist modes
    a = OC.q_bruteforce(pre, R(f), R(g), 1.0)
    b = R(OC.q_bruteforce(pre, f, g, 1.0))
    assert np.abs(a - b).max() < 5e-2 * np.abs(b).max()


def test_planted_mistake(monkeypatch):
    """Plausible pipeline mistakes must fail the brute-force pin: a normalised
    inverse transform (missing the cuFFT factor n) and a conjugated product."""
    vm, pre = _setup(6, 3, 4)
    f, g = _f(vm, (0.3, 0, 0), 0.8, 1), _f(vm, (-0.3, 0.2, 0), 1.2, 2)
    ref = OC.q_bruteforce(pre, f, g, 1.0)
    good = OC.q_vhs(pre, f, g, 1.0)
    assert np.abs(good - ref).max() < 1e-13 * np.abs(ref).max()
    orig = OC._ifft_u
    monkeypatch.setattr(OC, '_ifft_u', lambda x, N: orig(x, N) / N ** 3)
    bad1 = OC.q_vhs(pre, f, g, 1.0)
    monkeypatch.setattr(OC, '_ifft_u', orig)
    assert np.abs(bad1 - ref).max() > 1e-3 * np.abs(ref).max()
    # sign error in the cos/sin product sum (cplxMulAdd subtracting)
    N, n = pre.N, pre.n
    FTf, FTg = OC._fft(f.astype(complex), N) / n, OC._fft(g.astype(complex), N) / n
    a = pre.a.reshape(-1, n)
    t5 = OC._ifft_u(np.cos(a) * FTf, N) * OC._ifft_u(np.cos(a) * FTg, N) \
        - OC._ifft_u(np.sin(a) * FTf, N) * OC._ifft_u(np.sin(a) * FTg, N)
    t1 = OC._fft(t5, N).reshape(pre.Nrho, pre.M, n)
    QG = OC._ifft_u((pre.weight()[:, None, :] * t1).sum(axis=(0, 1)), N)
    bad2 = QG.real - f * OC._ifft_u(pre.bb2 * FTg, N).real
    assert np.abs(bad2 - ref).max() > 1e-3 * np.abs(ref).max()


def bkw(c2, t):
    """Bobylev-Krook-Wu solution of the space-homogeneous Boltzmann equation
    for Maxwell molecules with kernel 1/(4 pi) (std/initcond.py:121-141 for f),
    and its time derivative, K = 1 - exp(-t/6), dK/dt = (1 - K)/6."""
    K = 1.0 - np.exp(-t / 6.0)
    A = 1.0 / (2.0 * (2.0 * np.pi * K) ** 1.5)
    E = np.exp(-c2 / (2.0 * K))
    f = A * E * ((5.0 * K - 3.0) / K + (1.0 - K) / K ** 2 * c2)
    dfdK = f * (-1.5 / K + c2 / (2.0 * K * K)) + A * E * (3.0 / K ** 2 + c2 * (K - 2.0) / K ** 3)
    return f, dfdK * (1.0 - K) / 6.0


def test_bkw_closed_form_pins_normalisation():
    """gamma = 0 and prefactor 1/(4 pi) ('maxwell-gll'): the operator applied to
    f_BKW(t) is d f_BKW/dt, to spectral accuracy that improves with N.  This
    fixes every constant of the pipeline at once (1/n scalings, radial and
    sphere weights, bb1, bb2): a factor-2 slip in the sphere weight fails."""
    errs = []
    for N, Nrho in ((16, 8), (24, 12)):
        vm = OVM.VelocityMesh(1e-3, 273.0, 2.3e-4, 39.96e-3, N, Nrho, 0.0, 273.0, 6.0, 6)
        pre = OC.Precomp(N, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, 0.0)
        c2 = (vm.cv ** 2).sum(0)
        e = 0.0
        for t in (6.0, 7.0):
            f, dfdt = bkw(c2, t)
            q = OC.q_vhs_gll(pre, f, 1.0 / (4.0 * np.pi))
            e = max(e, np.abs(q - dfdt).max() / np.abs(dfdt).max())
            if N == 24:
                slip = OC.Precomp(N, vm.L, vm.qz, vm.qw, vm.sz, 2.0 * vm.sw, 0.0)
                qs = OC.q_vhs_gll(slip, f, 1.0 / (4.0 * np.pi))
                assert np.abs(qs - dfdt).max() > 0.5 * np.abs(dfdt).max()
        errs.append(e)
    assert errs[1] < 5e-4 and errs[1] < errs[0] / 10


def test_reference_state_form_is_exact():
    """The reference-state form Q(f0, f0) + G(h, f + f0) - h L[f] - f0 L[h]
    (h = f - f0) reproduces Q(f, f) for any f0, also far from f: it rests only
    on the gain being a symmetric bilinear form, which the brute-force sum
    pins above.  A sign slip on either loss correction (or taking L[s] for
    L[f]) breaks it by O(|h| |L|)."""
    vm, pre = _setup(8, 6, 4)
    f = _f(vm, (0.3, 0.1, 0), 0.9, 3)
    for f0 in (_f(vm, (0.25, 0.1, 0), 0.95, 4), _f(vm, (-0.4, 0.3, 0.1), 1.3, 5)):
        Q0 = OC.q_vhs(pre, f0, f0, 0.7)
        ref = OC.q_vhs(pre, f, f, 0.7)
        got = OC.q_reference(pre, f, f0, Q0, 0.7)
        assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
        h = f - f0
        bad = Q0 + OC.q_vhs(pre, h, f + f0, 0.7) + 0.7 * (h * OC.loss_field(pre, f0)
                                                         + f0 * OC.loss_field(pre, h))
        assert np.abs(bad - ref).max() > 1e-6 * np.abs(ref).max()
    # with f0 = f the form returns Q0 exactly
    Qf = OC.q_vhs(pre, f, f, 0.7)
    assert np.array_equal(OC.q_reference(pre, f, f.copy(), Qf, 0.7), Qf)
