"""GPU parity of the binary-mixture VSS operator against the FP64 oracle.

The oracle (oracle/bivss.py) restates bi/scattering.py:95-321 and
bi/kernels/scattering/vss.mako; it is pinned by the physics checks of
tests/test_oracle_bivss.py.  Tolerances (relative to max|Q_pq|): FP64 1e-11,
FP32 3e-5.  Each case runs both orderings of a pair with unequal masses and
unequal VSS parameters, bilinear tasks with several terms, and shared output
slices (grouped accumulation).
"""
import numpy as np
import pytest

from oracle import bivss as OB
from oracle.velocitymesh import load_design

pytestmark = pytest.mark.gpu


def _mesh(N, M, Nrho):
    return OB.BiVelocityMesh(1e-3, 273.0, 1.1140244750324529e-4, 39.96e-3, N, Nrho, 0.0, 273.0,
                             4.0, [39.96e-3, 83.77e-3], M)


def _fields(vm, K, Ne, seed):
    rng = np.random.default_rng(seed)
    out = []
    for p, m in enumerate(vm.masses):
        fs = []
        for s in range(K * Ne):
            u = (0.2 - 0.15 * s + 0.1 * p, 0.05 * s, 0.0)
            T = 0.9 + 0.1 * s + 0.2 * p
            f = OB.maxwellian(vm.cv, m, 1.0 - 0.3 * p, u, T)
            fs.append(f * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, f.shape)))
        out.append(np.stack(fs))
    return out


TASKS = [(0, [(0, 1.0)], [(1, 1.0)]), (1, [(2, 0.5), (3, -0.25)], [(1, 1.0), (0, 0.3)]),
         (1, [(3, 1.0)], [(2, -1.0)]), (3, [(1, 0.7)], [(1, 0.7), (2, 0.1)]),
         (2, [(2, 1.0)], [(2, 1.0)])]


def _run(cuda, N, M, Nrho, Mpre, dtype, p, q, tasks=TASKS, K=2, Ne=2, ntask_max=None):
    import torch
    from dgfs1d_gpu_amd import _lib
    from dgfs1d_gpu_amd.collision import BiCollisionPlan
    vm = _mesh(N, M, Nrho)
    szpre = load_design('womersley', Mpre)
    params = {(0, 0): (1.4, 0.81), (0, 1): (1.36, 0.805), (1, 0): (1.36, 0.805),
              (1, 1): (1.32, 0.80)}
    al, om = params[p, q]
    m = vm.masses
    prefac = 0.41
    pre = OB.BiPrecomp(N, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, szpre, m[p], m[q], al, om)
    u = _fields(vm, K, Ne, 7)
    n = N ** 3
    q0 = np.random.default_rng(5).standard_normal((K * Ne, n)) * 1e-3
    plan = BiCollisionPlan(N, M, Nrho, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, szpre, m[p], m[q], al,
                           om, prefac, dtype, ntask_max or len(tasks), device=cuda)
    plan.set_tasks(K, Ne, tasks)
    td = torch.float64 if dtype == _lib.F64 else torch.float32
    up = torch.tensor(u[p], dtype=td, device=cuda)
    uq = torch.tensor(u[q], dtype=td, device=cuda)
    Qd = torch.tensor(q0, dtype=td, device=cuda)
    plan(up, uq, Qd)
    torch.cuda.synchronize()
    got = Qd.double().cpu().numpy()
    ref = q0.copy()
    for out, ft, gt in tasks:
        f = sum(c * u[p][s] for s, c in ft)
        g = sum(c * u[q][s] for s, c in gt)
        ref[out] += OB.q_pq(pre, f, g, prefac)
    scale = np.abs(ref - q0).max()
    info = plan.query()
    plan.close()
    return np.abs(got - ref).max() / scale, info


@pytest.mark.parametrize('N,M,Nrho,Mpre', [(8, 6, 4, 32), (12, 12, 6, 48), (16, 6, 8, 32)])
@pytest.mark.parametrize('pq', [(0, 1), (1, 0), (1, 1)])
def test_bi_collision_fp64(cuda, N, M, Nrho, Mpre, pq):
    from dgfs1d_gpu_amd import _lib
    err, info = _run(cuda, N, M, Nrho, Mpre, _lib.F64, *pq)
    print('bi fp64', N, M, Nrho, Mpre, pq, err, info, flush=True)
    assert err < 1e-11


@pytest.mark.parametrize('N,M,Nrho,Mpre', [(24, 6, 4, 32), (32, 6, 4, 32)])
@pytest.mark.parametrize('pq', [(0, 1), (1, 1)])
def test_bi_collision_fp64_slab_path(cuda, N, M, Nrho, Mpre, pq):
    """Grids whose gain cube needs a residue split run the two-kernel slab
    path (k_bi_yz, k_bi_xprod)."""
    from dgfs1d_gpu_amd import _lib
    err, info = _run(cuda, N, M, Nrho, Mpre, _lib.F64, *pq)
    print('bi fp64 slab', N, pq, err, info, flush=True)
    assert info['split'] > 1
    assert err < 1e-11


@pytest.mark.parametrize('N,M,Nrho,Mpre', [(8, 6, 4, 32), (24, 6, 6, 32), (32, 6, 4, 32)])
def test_bi_collision_fp32(cuda, N, M, Nrho, Mpre):
    from dgfs1d_gpu_amd import _lib
    err, info = _run(cuda, N, M, Nrho, Mpre, _lib.F32, 0, 1)
    print('bi fp32', N, M, Nrho, Mpre, err, info, flush=True)
    assert err < 3e-5


def test_bi_collision_fp64_n24_fourier_example(cuda):
    """The reference's binary Fourier example grid (Nv=24, M=12, Nrho=6,
    Mpre=192; examples/bi/fourier/dgfs1D.ini), one node of each pair."""
    from dgfs1d_gpu_amd import _lib
    tasks = [(0, [(0, 1.0)], [(0, 1.0)])]
    for pq in ((0, 0), (0, 1)):
        err, info = _run(cuda, 24, 12, 6, 192, _lib.F64, *pq, tasks=tasks, K=1, Ne=1)
        print('bi fp64 n24', pq, err, info, flush=True)
        assert err < 1e-11


def test_bi_collision_chunked(cuda, monkeypatch):
    """A scratch budget below one chunk of all tasks splits the launch into
    several chunks (task-table offsets per chunk); results are unchanged."""
    from dgfs1d_gpu_amd import _lib
    monkeypatch.setenv('DGFS_BI_SCRATCH_MB', '3')      # 2 tasks per chunk at N=16 / 24
    for N, dt in ((16, _lib.F64), (24, _lib.F64), (16, _lib.F32)):
        err, info = _run(cuda, N, 6, 4, 32, dt, 0, 1)
        print('bi chunked', N, dt, err, info, flush=True)
        assert info['chunk'] < len(TASKS)
        assert err < (1e-11 if dt == _lib.F64 else 3e-5)


def test_bi_collision_fp32_n48(cuda):
    """N = 48 (the mixture grid of the reference's largest published runs),
    FP32 on the slab path; FP64 at N = 48 is refused with an error."""
    from dgfs1d_gpu_amd import _lib
    err, info = _run(cuda, 48, 6, 2, 32, _lib.F32, 0, 1, tasks=TASKS[:2])
    print('bi fp32 n48', err, info, flush=True)
    assert info['split'] > 1
    assert err < 3e-5
    with pytest.raises(_lib.DGFSError, match='N=48 is built for DGFS_F32 only'):
        _run(cuda, 48, 6, 2, 32, _lib.F64, 0, 1, tasks=TASKS[:1])
