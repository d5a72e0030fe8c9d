"""Quadrature rules and DG bases: oracle pinned by closed forms / library
routines, product set-up code checked against the oracle (two independent
implementations of dgfs1D/quadratures.py and dgfs1D/basis.py)."""
import numpy as np
import pytest

from oracle import quadratures as OQ
from oracle.basis import OracleBasis

KINDS = ['nodal-sem-gll', 'nodal-gll', 'modal-orthonormal-legendre',
         'modal-modified-karniadakis', 'nodal-sem-gll-2']


@pytest.mark.parametrize('n', range(1, 12))
def test_gauss_legendre_matches_numpy(n):
    z, w = OQ.zwgj(n, 0.0, 0.0)
    zr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(z, zr, atol=1e-14) and np.allclose(w, wr, atol=1e-14)


def test_gll_closed_forms():
    z, w = OQ.zwglj(3, 0.0, 0.0)
    assert np.allclose(z, [-1, 0, 1], atol=1e-15) and np.allclose(w, [1 / 3, 4 / 3, 1 / 3])
    z, w = OQ.zwglj(4, 0.0, 0.0)
    assert np.allclose(z, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1])
    assert np.allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6])


@pytest.mark.parametrize('rule,deg', [('zwgj', lambda n: 2 * n - 1), ('zwglj', lambda n: 2 * n - 3),
                                      ('zwgrjm', lambda n: 2 * n - 2),
                                      ('zwgrjp', lambda n: 2 * n - 2)])
def test_exactness(rule, deg):
    for n in range(2, 9):
        z, w = getattr(OQ, rule)(n, 0.0, 0.0)
        for k in range(deg(n) + 1):
            exact = (1 - (-1) ** (k + 1)) / (k + 1)
            assert abs(np.sum(w * z ** k) - exact) < 1e-12


@pytest.mark.parametrize('rule', ['zwgj', 'zwglj', 'zwgrjm', 'zwgrjp'])
@pytest.mark.parametrize('a,b', [(0.0, 0.0), (1.0, 0.0), (0.5, -0.5), (2.0, 1.0)])
def test_product_quadratures_match_oracle(rule, a, b):
    from dgfs1d_gpu_amd import quadratures as PQ
    for n in range(2, 10):
        zo, wo = getattr(OQ, rule)(n, a, b)
        zp, wp = getattr(PQ, rule)(n, a, b)
        assert np.abs(zo - zp).max() < 1e-12
        assert np.abs(wo - wp).max() < 1e-11 * np.abs(wo).max()


@pytest.mark.parametrize('kind', KINDS)
@pytest.mark.parametrize('K', [2, 3, 4, 5, 7])
def test_product_basis_matches_oracle(kind, K):
    from dgfs1d_gpu_amd.basis import Basis
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.nputil import subclass_where
    p = subclass_where(Basis, basis_kind=kind)(
        Dictionary('[basis]\nkind=%s\nK=%d\nNqr=6\n' % (kind, K)))
    o = OracleBasis(kind, K, Nqr=6)
    for a, b in ((p.mat_fwd, o.fwd), (p.mat_bwd, o.bwd), (p.mat_face, o.face),
                 (p.mat_deriv, o.deriv), (p.mat_invmass, o.invmass), (p.gLD, o.gLD),
                 (p.gRD, o.gRD), (p.interpMat, o.interp)):
        assert np.abs(a - b).max() < 1e-11 * max(1.0, np.abs(b).max())
    assert [list(s) for s in p.sigModes] == [list(s) for s in o.sig]
    for m in range(K):
        Pm = sum((np.outer(p.U[m][:, r], p.V[m][r]) for r in p.sigModes[m]), np.zeros((K, K)))
        Om = sum((np.outer(o.U[m][:, r], o.V[m][r]) for r in o.sig[m]), np.zeros((K, K)))
        assert np.abs(Pm - Om).max() < 1e-12


@pytest.mark.parametrize('kind', KINDS)
def test_derivative_exact_on_polynomials(kind):
    """D applied to the coefficients of x^j (j < K) gives those of j x^{j-1}
    at the solution points (oracle basis)."""
    K = 4
    o = OracleBasis(kind, K, Nqr=6)
    z = o.z
    for j in range(K):
        coeff = o.fwd @ (z ** j)                  # coefficients of x^j
        d = o.deriv @ coeff                       # derivative operator (strong form)
        if kind in ('nodal-sem-gll', 'nodal-gll'):
            vals = d
        else:
            continue                              # modal: weak-form operator, not pointwise
        exact = j * z ** (j - 1) if j else np.zeros_like(z)
        assert np.abs(vals - exact).max() < 1e-12


def test_sig_mode_counts():
    """nodal-sem-gll has one significant mode per m (one Q per node); the
    modified Karniadakis basis needs close to K^2 (README.md:15)."""
    assert sum(len(s) for s in OracleBasis('nodal-sem-gll', 5).sig) == 5
    n = sum(len(s) for s in OracleBasis('modal-modified-karniadakis', 4).sig)
    assert 12 <= n <= 16
