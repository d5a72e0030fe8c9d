"""GPU parity of the upwind summation-by-parts IMEX driver (dgfsAstdu1D,
reference astd/abstdu.py) against the FP64 oracle: the split upwind transport
(oracle.std.StdOracle.transport_upwind, pinned by its sigma = 0 equivalence
to the standard transport), IMEX steps with it as the explicit part, the
positivity limiter (oracle.limiter.limit_positivity) and the driver's limiter
option.  Tolerance: FP64 relative 1e-11 (limiters 1e-13)."""
import numpy as np
import pytest

from oracle import imex as OI
from oracle.deck import Deck
from oracle.limiter import limit_positivity, upwind_ops

pytestmark = pytest.mark.gpu


def _text(scheme='ars-222', over=None):
    from tests.test_gpu_imex import EXTRA
    from tests.test_gpu_solver import _ini
    o = {('time-integrator', 'scheme'): scheme, ('scattering-model', 'muRef'): 2.117e-5,
         ('mesh', 'Ne'): 5}
    o.update(over or {})
    return _ini(o) + EXTRA


def _solver(cuda, text):
    from dgfs1d_gpu_amd.astd.abstdu import DGFSAstdu1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    return DGFSAstdu1D(Dictionary(text), device=cuda, writers=False)


@pytest.mark.parametrize('over', [{}, {('basis', 'K'): 4},
                                  {('soln-bcs-xlo', 'type'): 'dgfs-periodic',
                                   ('soln-bcs-xhi', 'type'): 'dgfs-periodic'}])
def test_upwind_transport(cuda, over):
    import torch
    text = _text(over=over)
    deck = Deck(text)
    s = _solver(cuda, text)
    b = deck.basis
    Ap, Am = upwind_ops(b.K, b.z, b.w, b.D)
    rng = np.random.default_rng(1)
    u = deck.u0 * (1 + 0.1 * rng.uniform(-1, 1, deck.u0.shape))
    d = torch.tensor(u.ravel(), dtype=s.tdtype, device=cuda)
    out = torch.zeros_like(d)
    s.transport(0.0, d, out)
    torch.cuda.synchronize()
    ref = deck.op.transport_upwind(0.0, u, Ap, Am)
    err = np.abs(out.cpu().numpy().reshape(ref.shape) - ref).max() / np.abs(ref).max()
    print('upwind transport', over, err, flush=True)
    assert err < 1e-11


def test_upwind_ars_steps(cuda):
    from tests.test_gpu_imex import _oracle
    text = _text('ars-222')
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, _, _ = _oracle(deck)
    b, op = deck.basis, deck.op
    K, Ne, n = deck.K, deck.Ne, deck.vm.vsize
    Ap, Am = upwind_ops(b.K, b.z, b.w, b.D)

    def explicit(t, f):
        ux = op.transport_upwind(t, f.reshape(K, Ne, n), Ap, Am)
        return (b.invmass @ ux.reshape(K, -1)).reshape(K * Ne, n)

    f = deck.u0.reshape(K * Ne, n).copy()
    t = 0.0
    for k in range(3):
        f = OI.ars_step('ars-222', pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    print('upwind ars-222', err, flush=True)
    assert err < 1e-11


def test_positivity_limiter_and_option(cuda):
    import torch
    from dgfs1d_gpu_amd.transport import Limiter, PositivityLimiter
    text = _text('ars-222')
    a = _solver(cuda, text.replace('scheme = ars-222', 'scheme = ars-222\nlimiter = 1'))
    bs = _solver(cuda, text)
    assert a.poslimiter is not None and bs.poslimiter is None
    K, Ne, n = a.K, a.Ne, a.Nv
    # entry: a field with negative node values
    rng = np.random.default_rng(7)
    u = rng.uniform(0.05, 1.0, (K, Ne, n))
    u[0, 1:4, ::7] = -0.02
    pl = PositivityLimiter(a.basis, Ne, n, a.tdtype, cuda)
    d = torch.tensor(u.ravel(), dtype=a.tdtype, device=cuda)
    pl(d, d)
    torch.cuda.synchronize()
    ref = limit_positivity(u, a.basis.z, a.basis.Dn)
    err = np.abs(d.cpu().numpy().reshape(u.shape) - ref).max() / np.abs(u).max()
    print('positivity limiter', err, flush=True)
    assert err < 1e-13
    # option: the step is followed by the linear, then the positivity limiter
    osc = 1.0 + 0.2 * (-1.0) ** np.arange(K)[:, None, None] * np.ones((1, Ne, 1))
    u0 = (a.u.double().cpu().numpy().reshape(K, Ne, n) * osc).ravel()
    for s in (a, bs):
        s.u.copy_(torch.tensor(u0, dtype=s.tdtype, device=cuda))
    lin = Limiter(bs.basis, bs.xsol, bs.mesh.invjac[:, 0], n, bs.tdtype, cuda)
    for k in range(2):
        a.step()
        bs.step()
        lin(bs.u, bs.u)
        pl(bs.u, bs.u)
        torch.cuda.synchronize()
        e = float((a.u - bs.u).abs().max() / bs.u.abs().max())
        print('upwind limiter option', k, e, flush=True)
        assert e < 1e-14
