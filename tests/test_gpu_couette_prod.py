"""The production Couette deck (Ne=4, K=3, N=32^3, M=12, N_rho=16, SSP-RK2).

* One FP64 RHS of the full operator (transport, diffuse walls, 12 batched
  collisions, FP64 conservation projection) against the FP64 oracle at full
  size: 1e-11 relative.
* The FP32 spectral path (the deck's reference-state form, FP64 reference
  collision refreshed every few RHS calls) against the FP64 path after the
  deck's 52 SSP-RK2 steps from the same state: all nine moments (n, u_x, u_y,
  T, q_x, q_y, P_xx, P_yy, P_xy) within 1e-6 relative to the field's maximum
  (u_x, which stays near zero, relative to the wall speed), and the collision
  output of the FP64 state conserving mass, momentum and energy to 1e-12
  after projection.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   'dgfs1d_gpu_amd', 'configs', 'couette_prod.ini')


def _solver(cuda, precision):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    cfg = Dictionary(open(CFG).read())
    cfg.set('config', 'precision', precision)
    return DGFSStd1D(cfg, device=cuda, writers=False)


def test_prod_rhs_fp64_vs_oracle(cuda):
    import torch
    from oracle import collision as OC
    from oracle.deck import Deck
    text = open(CFG).read().replace('precision = float', 'precision = double')
    deck = Deck(text)
    s = _solver(cuda, 'double')
    rng = np.random.default_rng(7)
    u = deck.u0 * (1.0 + 0.05 * rng.uniform(-1, 1, deck.u0.shape))
    s.u.copy_(torch.tensor(u.ravel(), dtype=torch.float64, device=cuda))
    out = torch.zeros_like(s.u)
    s.rhs(0.0, s.u, out)
    OC.set_fft_workers(min(16, os.cpu_count() or 1))
    try:
        ref = deck.op.rhs(0.0, u)
    finally:
        OC.set_fft_workers(None)
    err = float(np.abs(out.cpu().numpy().reshape(ref.shape) - ref).max() / np.abs(ref).max())
    print('prod rhs fp64 vs oracle', err)
    assert err < 1e-11


def test_prod_fp32_moments_and_conservation(cuda):
    import torch
    from dgfs1d_gpu_amd.transport import Moments
    from oracle import relax as OR
    s64 = _solver(cuda, 'double')
    s32 = _solver(cuda, 'float')
    assert s32.sm._ref is not None            # the deck runs the reference-state form
    for _ in range(52):
        s64.step()
        s32.step()
    mom = Moments(s64.basis.interpMat, s64.K, s64.Ne, s64.vm, cuda)
    m64 = mom(s64.state_fp64()).reshape(-1, 9)
    m32 = mom(s32.state_fp64()).reshape(-1, 9)      # FP32 state (hi + lo), FP64 moments
    names = ('rho', 'ux', 'uy', 'T', 'qx', 'qy', 'Pxx', 'Pyy', 'Pxy')
    uwall = 168.52743267214112 / s64.vm.u0()
    rel = {}
    for i, nm in enumerate(names):
        scale = uwall if nm == 'ux' else np.abs(m64[:, i]).max()
        rel[nm] = float(np.abs(m32[:, i] - m64[:, i]).max() / scale)
    print('fp32 vs fp64 moments after 52 steps', rel)
    for nm in names:
        assert rel[nm] < 1e-6, nm
    Q = torch.zeros_like(s64.u)
    s64.sm.collide(s64.u, Q)
    q = Q.cpu().numpy().reshape(s64.K * s64.Ne, -1)
    phi = OR.invariants(s64.vm.cv())
    cons = float((np.abs(phi @ q.T) / (np.abs(phi) @ np.abs(q).T)).max())
    print('fp64 collision invariant moments', cons)
    assert cons < 1e-12
