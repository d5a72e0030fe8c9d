"""GPU parity of the Shakov relaxation term and the FP64 conservation projection.

dgfs_shakov (batched tasks, multi-term f combinations, several tasks per output
slice) against oracle.relax.q_shakov; dgfs_project_conservative against
oracle.relax.project_conservative and against the moment condition itself.
Tolerances: FP64 1e-12 relative to max|Q|; FP32 2e-5 -- FP32 inputs and outputs
with FP64 moments and FP64 evaluation; projected FP64 moments below
1e-12 of sum |phi||Q|.
"""
import numpy as np
import pytest

from oracle import relax as OR

pytestmark = pytest.mark.gpu


def _grid(N, L=5.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(c0, c0, c0, indexing='ij')
    return np.stack([I.ravel(), J.ravel(), K.ravel()]), (2 * L / N) ** 3


def _maxw(cv, rho, u, T):
    u = np.asarray(u, float).reshape(3, 1)
    return rho / (np.pi * T) ** 1.5 * np.exp(-((cv - u) ** 2).sum(0) / T)


def _state(cv, K, Ne, seed):
    rng = np.random.default_rng(seed)
    u = np.empty((K, Ne, cv.shape[1]))
    for k in range(K):
        for e in range(Ne):
            a = rng.uniform(-0.3, 0.3, 3)
            u[k, e] = 0.6 * _maxw(cv, 1.0, a, rng.uniform(0.7, 1.2)) + \
                0.4 * _maxw(cv, 1.1, -a, rng.uniform(0.9, 1.4))
    return u


def _dev(x, td, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(x).ravel(), dtype=td, device=dev)


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_shakov_batched_vs_oracle(cuda, dtype):
    import torch
    from dgfs1d_gpu_amd.relax import Shakov
    td = torch.float64 if dtype == 'f64' else torch.float32
    N, K, Ne = 12, 3, 2
    cv, cw = _grid(N)
    u = _state(cv, K, Ne, seed=11)
    prefac, omega, Pr = 3.7, 0.81, 2.0 / 3.0
    # tasks: single-slice, two-term combinations, two tasks on one output slice
    tasks = [(0, [(0, 1.0)]), (1, [(1, 1.0)]), (2, [(2, 0.7), (4, 0.3)]),
             (3, [(3, 1.0)]), (3, [(5, 0.5), (1, 0.5)]), (5, [(5, 1.0)])]
    sh = Shakov(N ** 3, _dev(cv[0], td, cuda), _dev(cv[1], td, cuda), _dev(cv[2], td, cuda),
                cw, prefac, omega, Pr, td, cuda)
    sh.set_tasks(K, Ne, [(o, f, f) for o, f in tasks])
    Q0 = np.random.default_rng(2).standard_normal(u.shape) * 1e-3
    Qd = _dev(Q0, td, cuda)
    sh(_dev(u, td, cuda), Qd)
    got = Qd.cpu().numpy().reshape(K * Ne, -1)
    ur = u.reshape(K * Ne, -1)
    if dtype == 'f32':
        ur = ur.astype(np.float32).astype(np.float64)
        cvr = cv.astype(np.float32).astype(np.float64)
        Q0 = Q0.astype(np.float32).astype(np.float64)
    else:
        cvr = cv
    ref = Q0.reshape(K * Ne, -1).copy()
    for o, terms in tasks:
        f = sum(c * ur[s] for s, c in terms)
        ref[o] += OR.q_shakov(f, cvr, cw, prefac, omega, Pr)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print('shakov', dtype, err)
    assert err < (1e-12 if dtype == 'f64' else 2e-5)


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_projection_vs_oracle(cuda, dtype):
    import torch
    from dgfs1d_gpu_amd.relax import ConservativeProjection
    td = torch.float64 if dtype == 'f64' else torch.float32
    N, ns = 16, 5
    cv, _ = _grid(N)
    rng = np.random.default_rng(4)
    Q = rng.standard_normal((ns, N ** 3)) * _maxw(cv, 1.0, (0, 0, 0), 1.0)[None, :]
    Q += 0.3 * np.stack([_maxw(cv, 1.0, (0.1 * s, 0, 0), 1.0) for s in range(ns)])
    P = ConservativeProjection(cv, _dev(cv[0], td, cuda), _dev(cv[1], td, cuda),
                               _dev(cv[2], td, cuda), td, cuda)
    Qd = _dev(Q, td, cuda)
    base = rng.standard_normal(Q.shape)
    out = _dev(base, td, cuda)
    P(Qd, out, ns, accumulate=True)
    got = out.cpu().numpy().reshape(ns, -1) - (base if dtype == 'f64'
                                               else base.astype(np.float32).astype(np.float64))
    cvr = cv if dtype == 'f64' else cv.astype(np.float32).astype(np.float64)
    Qr = Q if dtype == 'f64' else Q.astype(np.float32).astype(np.float64)
    ref = np.stack([OR.project_conservative(Qr[s], cvr) for s in range(ns)])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print('projection', dtype, err)
    assert err < (1e-12 if dtype == 'f64' else 2e-5)
    if dtype == 'f64':
        # in place (out aliases Q, accumulate = 0): moments vanish to 1e-12
        P(Qd, Qd, ns, accumulate=False)
        qc = Qd.cpu().numpy().reshape(ns, -1)
        phi = OR.invariants(cv)
        scale = np.abs(phi) @ np.abs(qc).T
        m = np.abs(phi @ qc.T) / scale
        print('projected moments', m.max())
        assert m.max() < 1e-12
