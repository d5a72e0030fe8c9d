"""Small Couette deck (couette_small.ini: Ne=2, K=3, N=16, M=6, N_rho=8,
VHS omega=0.81, Kn=0.1, diffuse walls u_y=-/+0.5 u0, Euler dt=1e-3, FP64):
52 GPU steps against the oracle's golden trajectory
(tests/golden/couette_small_oracle.npz, written by make_golden.py).
Tolerances: moments 1e-10 relative per field scale, final coefficients
1e-10 relative to max|u|, per-step residual 1e-8 relative."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_couette_small_52_steps(cuda):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    from dgfs1d_gpu_amd.transport import Moments
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'couette_small_oracle.npz'))
    cfg = Dictionary.load(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'couette_small.ini'))
    s = DGFSStd1D(cfg, device=cuda, writers=False)
    mom = Moments(s.basis.interpMat, s.K, s.Ne, s.vm, cuda)
    got_m = []
    for step in range(1, 53):
        s.step()
        if step in g['steps']:
            got_m.append(mom(s.u))
    got_m = np.array(got_m)
    scale = np.abs(g['moments']).max(axis=(0, 1, 2))
    scale[scale == 0] = 1.0
    merr = (np.abs(got_m - g['moments']).max(axis=(0, 1, 2)) / scale)
    uerr = np.abs(s.u.cpu().numpy().reshape(g['u_final'].shape) - g['u_final']).max() \
        / np.abs(g['u_final']).max()
    print('moment errors per field', merr, 'u err', uerr)
    assert merr.max() < 1e-10
    assert uerr < 1e-10


def test_couette_small_residual(cuda):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    from dgfs1d_gpu_amd.transport import Residual
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'couette_small_oracle.npz'))
    cfg = Dictionary.load(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'couette_small.ini'))
    s = DGFSStd1D(cfg, device=cuda, writers=False)
    res = Residual(cuda)
    prev = s.u.clone()
    out = []
    for _ in range(52):
        s.step()
        out.append(res(s.u, prev))
        prev.copy_(s.u)
    err = np.abs(np.array(out) - g['residual']) / g['residual']
    print('residual rel err max', err.max())
    assert err.max() < 1e-8
