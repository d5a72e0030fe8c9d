"""GPU parity of the entropy-stable IMEX driver (dgfsAstde21D, reference
astd/abstde2.py) against the FP64 oracle: the flux-differencing transport
(oracle.std.StdOracle.transport_entropy, pinned by its arithmetic-mean
equivalence to the standard transport and the log-mean closed form), IMEX
steps with it as the explicit part, and the driver's limiter option.
Tolerance: FP64 relative 1e-11."""
import numpy as np
import pytest

from oracle import imex as OI
from oracle.deck import Deck
from tests.test_gpu_upwind import _text

pytestmark = pytest.mark.gpu


def _solver(cuda, text):
    from dgfs1d_gpu_amd.astd.abstde2 import DGFSAstde21D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    return DGFSAstde21D(Dictionary(text), device=cuda, writers=False)


@pytest.mark.parametrize('over', [{}, {('basis', 'K'): 4}, {('basis', 'kind'): 'nodal-gll'},
                                  {('soln-bcs-xlo', 'type'): 'dgfs-periodic',
                                   ('soln-bcs-xhi', 'type'): 'dgfs-periodic'}])
def test_entropy_transport(cuda, over):
    import torch
    text = _text(over=over)
    deck = Deck(text)
    s = _solver(cuda, text)
    rng = np.random.default_rng(1)
    u = deck.u0 * (1 + 0.1 * rng.uniform(-1, 1, deck.u0.shape))
    d = torch.tensor(u.ravel(), dtype=s.tdtype, device=cuda)
    out = torch.zeros_like(d)
    s.transport(0.0, d, out)
    torch.cuda.synchronize()
    ref = deck.op.transport_entropy(0.0, u, s.basis.derivMat)
    err = np.abs(out.cpu().numpy().reshape(ref.shape) - ref).max() / np.abs(ref).max()
    print('entropy transport', over, err, flush=True)
    assert err < 1e-11


def test_entropy_ars_steps(cuda):
    from tests.test_gpu_imex import _oracle
    text = _text('ars-222')
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, _, _ = _oracle(deck)
    b, op = deck.basis, deck.op
    K, Ne, n = deck.K, deck.Ne, deck.vm.vsize
    Dr = s.basis.derivMat

    def explicit(t, f):
        ux = op.transport_entropy(t, f.reshape(K, Ne, n), Dr)
        return (b.invmass @ ux.reshape(K, -1)).reshape(K * Ne, n)

    f = deck.u0.reshape(K * Ne, n).copy()
    t = 0.0
    for k in range(3):
        f = OI.ars_step('ars-222', pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    print('entropy ars-222', err, flush=True)
    assert err < 1e-11


def test_entropy_limiter_option(cuda):
    import torch
    from dgfs1d_gpu_amd.transport import Limiter, PositivityLimiter
    text = _text('ars-222')
    a = _solver(cuda, text.replace('scheme = ars-222', 'scheme = ars-222\nlimiter = 1'))
    bs = _solver(cuda, text)
    assert a.poslimiter is not None and bs.poslimiter is None
    K, Ne, n = a.K, a.Ne, a.Nv
    osc = 1.0 + 0.2 * (-1.0) ** np.arange(K)[:, None, None] * np.ones((1, Ne, 1))
    u0 = (a.u.double().cpu().numpy().reshape(K, Ne, n) * osc).ravel()
    for s in (a, bs):
        s.u.copy_(torch.tensor(u0, dtype=s.tdtype, device=cuda))
    lin = Limiter(bs.basis, bs.xsol, bs.mesh.invjac[:, 0], n, bs.tdtype, cuda)
    pos = PositivityLimiter(bs.basis, Ne, n, bs.tdtype, cuda)
    for k in range(2):
        a.step()
        bs.step()
        lin(bs.u, bs.u)
        pos(bs.u, bs.u)
        torch.cuda.synchronize()
        e = float((a.u - bs.u).abs().max() / bs.u.abs().max())
        print('entropy limiter option', k, e, flush=True)
        assert e < 1e-14
