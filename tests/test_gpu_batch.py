"""Batched uneven workload: many problems in flat tensors with one collision
launch sequence per velocity grid, HIP-graph capture and the q-group
autotuner.

* Batched SSP-RK2 steps equal the problems stepped one by one (FP64, 1e-12).
* Every replay of the captured step equals an eager step from the same state
  (several replays, eager work and synchronisations in between).
* The autotuner keeps results unchanged (any q-group count is exact up to
  summation order: FP64 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [dict(kind='couette', Ne=2, K=3, N=8, uw=0.4, Tl=1.1, Tr=1.1),
         dict(kind='fourier', Ne=3, K=3, N=12, uw=0.0, Tl=0.9, Tr=1.3),
         dict(kind='couette', Ne=2, K=4, N=8, uw=-0.7, Tl=1.2, Tr=1.2),
         dict(kind='fourier', Ne=2, K=3, N=12, uw=0.0, Tl=1.4, Tr=0.85)]


def _batch(cuda, precision='double'):
    from dgfs1d_gpu_amd.batch import BatchedStd1D
    return BatchedStd1D(SPECS, cuda, precision=precision, dt=1e-3)


def test_batched_equals_individual(cuda):
    b = _batch(cuda)
    for _ in range(3):
        b.step()
    for k, s in enumerate(b.solvers):
        s.dt = b.dt
        for _ in range(3):
            s.step()
        got = b.solver_state(k).cpu().numpy()
        ref = s.u.cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print('problem', k, b.specs[k]['N'], err)
        assert err < 1e-12


def test_graph_replay_matches_eager(cuda):
    import torch
    b = _batch(cuda)
    b.capture(warmup=1)
    for rep in range(3):
        u0 = b.u.clone()
        t0 = b.time
        b.graph.replay()
        torch.cuda.synchronize()
        g = b.u.clone()
        b.u.copy_(u0)
        b.odes.integrate(t0, b.dt, b.u)                    # eager step from the same state
        torch.cuda.synchronize()
        err = float((g - b.u).abs().max() / b.u.abs().max())
        print('replay', rep, err)
        assert err == 0.0
        b.time = t0 + b.dt


def test_autotune_keeps_results(cuda):
    import torch
    b = _batch(cuda)
    ref = torch.zeros_like(b.u)
    for g in b.groups:
        g['plan'](b.u[g['lo']:g['hi']], ref[g['lo']:g['hi']])
    res = b.autotune(candidates=(1, 2, 3, 6), reps=2)
    print('autotune', res)
    got = torch.zeros_like(b.u)
    for g in b.groups:
        g['plan'](b.u[g['lo']:g['hi']], got[g['lo']:g['hi']])
    err = float((got - ref).abs().max() / ref.abs().max())
    assert err < 1e-12
