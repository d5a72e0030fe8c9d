"""The learned-closure workload on the GPU: the Mach-2 hard-sphere normal
shock (Ne=32, K=3, N=32^3, M=12, N_rho=16, SSP-RK2 dt=5e-4; after the
reference's examples/std/normalshock/dgfs1D.ini) with the shipped network
(dgfs1d_gpu_amd/data/closure_hs_n32.npz).

* The FP64 spectral path, the FP32 spectral path, the FP32 learned closure
  and the FP32 closure without its network (Shakov + lagged spectral
  correction) advance the same initial profile 300 steps.  The closure runs
  with the bench's settings (100-call spectral warm start, checks every 10
  calls on 25 % of the nodes, tolerance 0.05, network every 4th call).  The
  moments of every FP32 run agree with the FP64 run within 1e-3 of each
  field's natural size (rho, |u|, T, rho T^{3/2} for q, rho T for P).
* What the network gains: on that run's final state (a Mach-2, T1 = 290 K
  shock, not a training state), Q_Shakov + R_net against the spectral Q, as
  relative L2 per node, is compared with Q_Shakov alone (the model the
  network corrects).  The network lowers the median error; the figures are
  printed.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, 'dgfs1d_gpu_amd', 'data', 'closure_hs_n32.npz')
STEPS = 300


def _closure_extra(use_net=True):
    return ('closure = learned\nclosure-weights = %s\nclosure-check-every = 10\n'
            'closure-tol = 0.05\nclosure-sample = 0.25\nclosure-start = 100\n'
            'closure-net-every = 4\nclosure-use-net = %s\n'
            % (WEIGHTS, 'true' if use_net else 'false'))


def _solver(cuda, precision, extra='', T1=223.0):
    from dgfs1d_gpu_amd.closure_train import shock_deck
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    text = shock_deck(mach=2.0, T1=T1, precision=precision, extra=extra)
    return DGFSStd1D(Dictionary(text), device=cuda, writers=False)


def _natural_error(m, ref):
    """max |m - ref| per field over the nodes, relative to the field's natural
    size: rho, |u|, T, rho T^{3/2} (q), rho T (P)."""
    rho, T = ref[:, 0], ref[:, 3]
    speed = np.sqrt(ref[:, 1] ** 2 + ref[:, 2] ** 2)
    scale = [rho.max(), speed.max(), speed.max(), T.max(), (rho * T ** 1.5).max(),
             (rho * T ** 1.5).max(), (rho * T).max(), (rho * T).max(), (rho * T).max()]
    names = ('rho', 'ux', 'uy', 'T', 'qx', 'qy', 'Pxx', 'Pyy', 'Pxy')
    return {nm: float(np.abs(m[:, i] - ref[:, i]).max() / scale[i]) for i, nm in enumerate(names)}


def test_shock_closure_moments_against_fp64(cuda):
    from dgfs1d_gpu_amd.transport import Moments
    runs = {'fp64': _solver(cuda, 'double'), 'fp32 spectral': _solver(cuda, 'float'),
            'fp32 learned': _solver(cuda, 'float', _closure_extra(True)),
            'fp32 no network': _solver(cuda, 'float', _closure_extra(False))}
    for name, s in runs.items():
        print('advancing', name, flush=True)
        for _ in range(STEPS):
            s.step()
    s64 = runs['fp64']
    mom = Moments(s64.basis.interpMat, s64.K, s64.Ne, s64.vm, cuda)
    ref = mom(s64.u).reshape(-1, 9).copy()
    for name in ('fp32 spectral', 'fp32 learned', 'fp32 no network'):
        s = runs[name]
        m = mom(s.state_fp64()).reshape(-1, 9)
        err = _natural_error(m, ref)
        cl = getattr(s.sm, 'closure', None)
        extra = '' if cl is None else ' (nodes on the closure %d of %d, fallbacks %d)' % (
            int(cl.learned.sum()), cl.ns, cl.fallbacks)
        print(name, 'vs fp64 after %d steps' % STEPS, err, extra, flush=True)
        for nm, e in err.items():
            assert e < 1e-3, (name, nm, e)


def test_shock_network_gain_over_shakov(cuda):
    """Q_Shakov + R_net versus Q_Shakov alone, against the spectral Q."""
    import torch
    s = _solver(cuda, 'float', _closure_extra(True), T1=290.0)
    for _ in range(STEPS):
        s.step()
    cl = s.sm.closure
    nodes = np.arange(cl.ns, dtype=np.int32)
    idx = torch.tensor(nodes, device=cuda)
    tab = cl._tab(nodes, cl.shakov.tdtype)
    u = s.u
    q_net, q_sk, q_sp = (torch.zeros_like(u) for _ in range(3))
    cl._model(u, q_net, nodes, idx, tab)          # Q_Shakov + R_net
    cl.shakov(u, q_sk, tasks=tab)                 # Q_Shakov
    cl._spectral(u, q_sp, nodes)                  # spectral Q
    n = cl.G ** 3
    qs = q_sp.view(cl.ns, n).double()
    den = qs.norm(dim=1)
    e_net = ((q_net.view(cl.ns, n).double() - qs).norm(dim=1) / den).cpu().numpy()
    e_sk = ((q_sk.view(cl.ns, n).double() - qs).norm(dim=1) / den).cpu().numpy()
    med_net, med_sk = float(np.median(e_net)), float(np.median(e_sk))
    print('relative L2 of Q against the spectral operator, median over %d nodes: '
          'Shakov %.4f, Shakov + network %.4f (%.1f %% lower); nodes improved %d'
          % (cl.ns, med_sk, med_net, 100 * (1 - med_net / med_sk), int((e_net < e_sk).sum())))
    assert med_net < med_sk
