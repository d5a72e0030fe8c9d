"""GPU parity of the linear minmod limiter (dgfs_limit_linear, reference
astd/kernels/limiters.mako:17-76) against the FP64 oracle (oracle/limiter.py,
pinned by tests/test_oracle_limiter.py), and the IMEX driver's
[time-integrator] limiter option.  Tolerances: FP64 1e-13, FP32 1e-5 of
max|u|."""
import numpy as np
import pytest

from oracle.limiter import limit_linear

pytestmark = pytest.mark.gpu


def _field(xsol, n, seed):
    rng = np.random.default_rng(seed)
    K, Ne = xsol.shape
    osc = 0.05 * (-1.0) ** np.arange(K)[:, None, None]
    ramp = xsol[:, :, None] * rng.uniform(-1, 1, n)[None, None, :]
    step = (xsol[:, :, None] > 0.1 * rng.uniform(-3, 3, n)[None, None, :]).astype(float)
    return 1.0 + ramp + 0.3 * step + osc * rng.uniform(0, 1, n)[None, None, :]


@pytest.mark.parametrize('kind,K', [('nodal-sem-gll', 3), ('nodal-sem-gll', 4), ('nodal-gll', 3)])
@pytest.mark.parametrize('prec', ['double', 'single'])
def test_limiter_entry(cuda, kind, K, prec):
    import torch
    from dgfs1d_gpu_amd.basis import Basis
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.mesh import Mesh
    from dgfs1d_gpu_amd.nputil import subclass_where
    from dgfs1d_gpu_amd.transport import Limiter
    from tests.test_gpu_solver import _ini
    cfg = Dictionary(_ini({('basis', 'kind'): kind, ('basis', 'K'): K, ('mesh', 'Ne'): 9,
                           ('config', 'precision'): prec}))
    b = subclass_where(Basis, basis_kind=kind)(cfg)
    mesh = Mesh(cfg)
    xm, jac = mesh.xmesh, mesh.jac
    xsol = np.array([0.5 * (xm[j] + xm[j + 1]) + jac[j] * b.z for j in range(mesh.Ne)]).T
    n = 64
    u = _field(xsol, n, 4)
    td = torch.float64 if prec == 'double' else torch.float32
    if prec == 'single':
        u = u.astype(np.float32).astype(np.float64)
    lim = Limiter(b, xsol, mesh.invjac[:, 0], n, td, cuda)
    d = torch.tensor(u.ravel(), dtype=td, device=cuda)
    out = torch.empty_like(d)
    lim(d, out)
    lim(d, d)                                           # in place
    torch.cuda.synchronize()
    ref = limit_linear(u, b.z, b.Dn, xsol, mesh.invjac[:, 0])
    assert np.abs(ref - u).max() > 1e-3                 # the limiter acts on this field
    tol = 1e-13 if prec == 'double' else 1e-5
    for got in (out, d):
        err = np.abs(got.double().cpu().numpy().reshape(u.shape) - ref).max() / np.abs(u).max()
        print('limiter', kind, K, prec, err, flush=True)
        assert err < tol


def test_astd_limiter_option(cuda):
    """[time-integrator] limiter = 1: each IMEX step is followed by the limiter."""
    import torch
    from dgfs1d_gpu_amd.transport import Limiter
    from tests.test_gpu_imex import _solver, _text
    base = _text('ars-222').replace('Ne = 3', 'Ne = 6')
    a = _solver(cuda, base.replace('scheme = ars-222', 'scheme = ars-222\nlimiter = 1'))
    b = _solver(cuda, base)
    assert a.limiter is not None and b.limiter is None
    # an odd-even oscillation at the nodes, so that the limiter acts
    K, Ne, n = a.K, a.Ne, a.Nv
    osc = 1.0 + 0.2 * (-1.0) ** np.arange(K)[:, None, None] * np.ones((1, Ne, 1))
    u0 = (a.u.double().cpu().numpy().reshape(K, Ne, n) * osc).ravel()
    for s in (a, b):
        s.u.copy_(torch.tensor(u0, dtype=s.tdtype, device=cuda))
    lim = Limiter(b.basis, b.xsol, b.mesh.invjac[:, 0], b.Nv, b.tdtype, cuda)
    for k in range(2):
        a.step()
        b.step()
        before = b.u.clone()
        lim(b.u, b.u)
        torch.cuda.synchronize()
        assert float((before - b.u).abs().max()) > 0          # the limiter acted
        err = float((a.u - b.u).abs().max() / b.u.abs().max())
        print('astd limiter', k, err, flush=True)
        assert err < 1e-14
