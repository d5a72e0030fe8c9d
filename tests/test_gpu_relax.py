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


def test_axnpby_compensated(cuda):
    """Double-float stage update: matches FP64 arithmetic on hi + lo to ~1e-13
    and keeps increments far below FP32 resolution that plain FP32 loses."""
    import torch
    from dgfs1d_gpu_amd.transport import axnpby, axnpby_compensated
    rng = np.random.default_rng(1)
    n = 100003
    hi = [torch.tensor(rng.uniform(0.5, 2, n), dtype=torch.float32, device=cuda) for _ in range(3)]
    lo = [torch.tensor(rng.uniform(-1, 1, n) * 1e-8, dtype=torch.float32, device=cuda)
          for _ in range(3)]
    d = rng.standard_normal(n)
    x0 = torch.tensor(d, dtype=torch.float32, device=cuda)
    x0l = torch.zeros_like(x0)
    a = (0.7, 1.0 / 3.0, -2.5e-4, 0.123456789)
    ref = a[0] * x0.double().cpu().numpy() + sum(
        a[i + 1] * (hi[i].double().cpu().numpy() + lo[i].double().cpu().numpy()) for i in range(3))
    axnpby_compensated(x0, x0l, [(a[i + 1], hi[i], lo[i]) for i in range(3)], a[0])
    got = x0.double().cpu().numpy() + x0l.double().cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print('compensated axnpby', err)
    assert err < 1e-12
    # 4000 increments of 1e-9 on 1.0: plain FP32 stays at 1.0, compensated keeps them
    u = torch.ones(1024, dtype=torch.float32, device=cuda)
    ul = torch.zeros_like(u)
    up = torch.ones_like(u)
    inc = torch.full_like(u, 1e-9)
    for _ in range(4000):
        axnpby_compensated(u, ul, [(1.0, inc, None)], 1.0)
        axnpby(up, [(1.0, inc)], 1.0)
    tot = (u.double() + ul.double()).cpu().numpy()
    assert np.abs(tot - (1.0 + 4000 * float(np.float32(1e-9)))).max() < 1e-10
    assert np.all(up.cpu().numpy() == 1.0)


def _newton_model(cuda, td, N, target, Pr=0.6, check=True, grid_n=None):
    from dgfs1d_gpu_amd.relax import Shakov
    cv, cw = _grid(N)
    sh = Shakov(N ** 3, _dev(cv[0], td, cuda), _dev(cv[1], td, cuda), _dev(cv[2], td, cuda),
                cw, 1.0, 0.81, Pr, td, cuda, target=target, check=check, grid_n=grid_n)
    assert sh.grid_n == (N if grid_n is None else grid_n)   # the test grid is a tensor grid
    return sh, cv, cw


@pytest.mark.parametrize('grid_n', [None, 0], ids=['tensor', 'points'])
@pytest.mark.parametrize('target', ['bgk', 'esbgk-conservative'])
def test_newton_targets_node_by_node(cuda, target, grid_n):
    """'bgk-gll' (std/scattering.py:445-706) and 'esbgk-gll' (:752-1185): the
    device Newton solves (the BGK sums factored on the tensor grid, or summed
    over the points with each task's sums spread over many work-groups)
    against the oracle's MINPACK solves node by node (FP64, 12 tasks as the
    Couette deck's K = 3, Ne = 4), and the moments of the device targets
    against the conditions that define them: mass, momentum and energy of f
    (bgk), mass, momentum and rho (u_i u_j + T_ij) (esbgk)."""
    import torch
    td = torch.float64
    N, K, Ne, Pr = 16, 3, 4, 0.6
    sh, cv, cw = _newton_model(cuda, td, N, target, Pr, grid_n=grid_n)
    u = _state(cv, K, Ne, seed=21)
    tasks = [(s, [(s, 1.0)]) for s in range(K * Ne)]
    sh.set_tasks(K, Ne, [(o, f, f) for o, f in tasks])
    Qd = torch.zeros(K * Ne * N ** 3, dtype=td, device=cuda)
    sh(_dev(u, td, cuda), Qd)
    assert not sh.status().any()
    got = Qd.cpu().numpy().reshape(K * Ne, -1)
    ur = u.reshape(K * Ne, -1)
    pr = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    worst, worst_m = 0.0, 0.0
    for s in range(K * Ne):
        f = ur[s]
        if target == 'bgk':
            tgt, rho, T = OR.bgk_conservative_target(f, cv, cw)
        else:
            tgt, rho, T = OR.esbgk_conservative_target(f, cv, cw, Pr)
        rate = rho / T ** (0.81 - 1.0)
        ref = rate * (tgt - f)
        worst = max(worst, np.abs(got[s] - ref).max() / np.abs(ref).max())
        G = f + got[s] / rate                       # the device's target
        uu = np.array([(f * cv[i]).sum() for i in range(3)]) / f.sum()
        if target == 'bgk':
            phi = np.vstack([np.ones(cv.shape[1]), cv, (cv * cv).sum(0)])
            want = phi @ f * cw
            sc = np.array([rho] + [rho * np.sqrt(T)] * 3 + [rho * T])
        else:
            C = cv - uu[:, None]
            fe, _, _ = OR.bgk_conservative_target(f, cv, cw)
            w = (1.0 - 1.0 / Pr) * f + (1.0 / Pr) * fe
            phi = np.vstack([np.ones(cv.shape[1]), cv] + [cv[i] * cv[j] for i, j in pr])
            Tij = [(C[i] * C[j] * w).sum() * cw / rho for i, j in pr]
            want = np.array([rho, *(rho * uu)] +
                            [rho * (uu[i] * uu[j] + Tij[k]) for k, (i, j) in enumerate(pr)])
            sc = np.array([rho] + [rho * np.sqrt(T)] * 3 + [rho * T] * 6)
        worst_m = max(worst_m, float((np.abs(phi @ G * cw - want) / sc).max()))
    print(target, 'Q vs oracle', worst, 'target moments', worst_m)
    assert worst < 1e-10
    assert worst_m < 1e-11


@pytest.mark.parametrize('target', ['bgk', 'esbgk-conservative'])
def test_newton_tensor_matches_points_fp32(cuda, target):
    """The tensor-grid Newton kernels (one-dimensional moments for BGK,
    column-wise Horner sums with the node-rounding correction for ES-BGK)
    against the point-sum kernels on an FP32 grid whose spacing is not a
    binary fraction (so the stored nodes are not exactly uniform), N = 24,
    12 tasks: the collision terms agree to FP32 output rounding, and the
    device targets keep f's moments."""
    import torch
    from dgfs1d_gpu_amd.relax import Shakov
    td = torch.float32
    N, K, Ne, Pr = 24, 3, 4, 2.0 / 3.0
    cv, cw = _grid(N, L=5.2371)
    cv32 = cv.astype(np.float32)
    dev = [_dev(cv32[i], td, cuda) for i in range(3)]
    u = _state(cv, K, Ne, seed=7).astype(np.float32)
    out = {}
    for gn in (None, 0):
        sh = Shakov(N ** 3, *dev, cw, 1.0, 0.81, Pr, td, cuda, target=target, grid_n=gn)
        assert sh.grid_n == (N if gn is None else 0)
        sh.set_tasks(K, Ne, [(s_, [(s_, 1.0)], [(s_, 1.0)]) for s_ in range(K * Ne)])
        Q = torch.zeros(K * Ne * N ** 3, dtype=td, device=cuda)
        sh(_dev(u, td, cuda), Q)
        out[gn] = Q.double().cpu().numpy().reshape(K * Ne, -1)
    err = np.abs(out[None] - out[0]).max() / np.abs(out[0]).max()
    phi = np.vstack([np.ones(cv.shape[1]), cv32.astype(np.float64),
                     (cv32.astype(np.float64) ** 2).sum(0)]) * cw
    mom = max(np.abs(phi @ out[g].T).max() / (np.abs(phi) @ np.abs(out[g].T)).max()
              for g in out)
    print(target, 'tensor vs points', err, 'moments of Q', mom)
    assert err < 2e-6 and mom < 1e-6


@pytest.mark.parametrize('grid_n', [None, 0], ids=['tensor', 'points'])
@pytest.mark.parametrize('target', ['bgk', 'esbgk-conservative'])
def test_newton_nan_raises(cuda, target, grid_n):
    """A task whose slice is zero has no temperature: its Newton solve meets a
    NaN, the status says so and the call raises, as the reference's
    RuntimeError('NaN encountered') (std/scattering.py:689, :1084).  The other
    tasks of the same call are solved as usual."""
    import torch
    td = torch.float64
    N, K, Ne = 8, 1, 3
    sh, cv, cw = _newton_model(cuda, td, N, target, check=False, grid_n=grid_n)
    u = _state(cv, K, Ne, seed=5)
    u[0, 1] = 0.0
    sh.set_tasks(K, Ne, [(s, [(s, 1.0)], [(s, 1.0)]) for s in range(Ne)])
    Qd = torch.zeros(K * Ne * N ** 3, dtype=td, device=cuda)
    sh(_dev(u, td, cuda), Qd)
    st = sh.status()
    print(target, 'status', st)
    assert st[1] & 1 and st[0] == 0 and st[2] == 0
    with pytest.raises(RuntimeError, match='NaN encountered'):
        sh.check_status()
    sh.check = True
    with pytest.raises(RuntimeError, match='NaN encountered'):
        sh(_dev(u, td, cuda), Qd)
