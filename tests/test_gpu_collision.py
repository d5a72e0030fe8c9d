"""GPU parity of the batched collision operator against the FP64 oracle.

The oracle (oracle/collision.py) restates std/scattering.py:251-313 ('vhs-gll')
and :346-429 ('vhs'); it is itself pinned by the brute-force spectral sum
(tests/test_oracle_collision.py).  Tolerances (relative to max|Q|):
FP64 1e-11, FP32 2e-5.
"""
import numpy as np
import pytest

from oracle import collision as OC
from oracle import velocitymesh as OVM

pytestmark = pytest.mark.gpu


def _mesh(N, M, Nrho):
    return OVM.VelocityMesh(H0=1e-3, T0=273.0, rho0=2.3e-4, molarMass0=39.96e-3, Nv=N,
                            Nrho=Nrho, cmax=0.0, Tmax=273.0, dev=5.0, M=M)


def _field(vm, seed, u=(0.3, 0.0, 0.0), T=0.8):
    rng = np.random.default_rng(seed)
    c = vm.cv - np.array(u)[:, None]
    f = np.exp(-(c * c).sum(0) / T) / (np.pi * T) ** 1.5
    return f * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, f.shape))


def _run(cuda, N, M, Nrho, dtype, variant, tasks_spec, K=2, Ne=2, omega=0.81):
    import torch
    from dgfs1d_gpu_amd import _lib
    from dgfs1d_gpu_amd.collision import CollisionPlan
    vm = _mesh(N, M, Nrho)
    gamma = 2.0 * (1.0 - omega)
    pre = OC.Precomp(N, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, gamma)
    prefac = 0.37
    n = N ** 3
    u = np.stack([_field(vm, 10 + s, u=(0.3 - 0.2 * s, 0.1 * s, 0.0), T=0.8 + 0.1 * s)
                  for s in range(K * Ne)])
    q0 = np.random.default_rng(99).standard_normal((K * Ne, n)) * 1e-3
    plan = CollisionPlan(N, M, Nrho, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, gamma, prefac,
                         dtype, variant, ntask_max=len(tasks_spec), device=cuda)
    plan.set_tasks(K, Ne, tasks_spec)
    td = torch.float64 if dtype == _lib.F64 else torch.float32
    ud = torch.tensor(u, dtype=td, device=cuda)
    Qd = torch.tensor(q0, dtype=td, device=cuda)
    plan(ud, Qd)
    torch.cuda.synchronize()
    got = Qd.double().cpu().numpy()
    ref = q0.copy()
    for out, ft, gt in tasks_spec:
        f = sum(c * u[s] for s, c in ft)
        g = sum(c * u[s] for s, c in gt)
        if variant == _lib.VHS:
            ref[out] += OC.q_vhs(pre, f, g, prefac)
        else:
            ref[out] += OC.q_vhs_gll(pre, f, prefac)
    scale = np.abs(ref - q0).max()
    err = np.abs(got - ref).max() / scale
    plan.close()
    return err


NODAL = [(s, [(s, 1.0)], [(s, 1.0)]) for s in range(4)]
BILINEAR = [(0, [(0, 1.0)], [(1, 1.0)]), (1, [(2, 0.5), (3, -0.25)], [(1, 1.0), (0, 0.3)]),
            (1, [(3, 1.0)], [(2, -1.0)]), (3, [(1, 0.7)], [(1, 0.7), (2, 0.1)])]


@pytest.mark.parametrize('N,M,Nrho', [(4, 6, 4), (6, 6, 4), (8, 6, 4), (16, 6, 8), (32, 6, 4)])
@pytest.mark.parametrize('variant', [0, 1])
def test_collision_fp64_nodal(cuda, N, M, Nrho, variant):
    from dgfs1d_gpu_amd import _lib
    err = _run(cuda, N, M, Nrho, _lib.F64, variant, NODAL)
    print('fp64 nodal', N, variant, err)
    assert err < 1e-11


@pytest.mark.parametrize('N,M,Nrho', [(6, 3, 4), (8, 6, 4), (12, 6, 6)])
def test_collision_fp64_bilinear_grouped(cuda, N, M, Nrho):
    from dgfs1d_gpu_amd import _lib
    err = _run(cuda, N, M, Nrho, _lib.F64, _lib.VHS, BILINEAR)
    print('fp64 bilinear', N, err)
    assert err < 1e-11


@pytest.mark.parametrize('N,M,Nrho', [(8, 6, 4), (12, 6, 4), (16, 6, 8), (24, 6, 6), (32, 6, 8)])
def test_collision_fp32(cuda, N, M, Nrho):
    from dgfs1d_gpu_amd import _lib
    err = _run(cuda, N, M, Nrho, _lib.F32, _lib.VHS, NODAL)
    print('fp32 nodal', N, err)
    assert err < 2e-5


def test_collision_n48_fp32(cuda):
    """The largest grid (N = 48, FP32 only): nodal f = g tasks and bilinear
    grouped tasks against the FP64 oracle, 3e-5 of max|Q| (the N = 16..32
    FP32 bound grows with the transform length)."""
    from dgfs1d_gpu_amd import _lib
    for name, spec in (('nodal', NODAL[:2]), ('bilinear', BILINEAR[:2])):
        err = _run(cuda, 48, 6, 4, _lib.F32, _lib.VHS, spec)
        print('fp32 N=48', name, err, flush=True)
        assert err < 3e-5, name


@pytest.mark.parametrize('N,M,Nrho', [(8, 6, 4), (16, 6, 8), (32, 6, 4)])
def test_collision_reference_form(cuda, N, M, Nrho):
    """dgfs_collide_reference: FP64 -- equal to the direct Q(f, f) to 1e-12;
    FP32 -- against the FP64 oracle on a state 1e-3 away from the reference
    state, its error is the direct FP32 error scaled down by at least 30x, down
    to the rounding of the FP32 output itself (half an ulp, 6e-8 of |Q|; N = 32
    runs the lane-pair store and delta kernels): the
    computed part scales with |f - f0|/|f|, while the gain and the loss each
    exceed Q several times here."""
    import torch
    from dgfs1d_gpu_amd import _lib
    from dgfs1d_gpu_amd.collision import CollisionPlan
    vm = _mesh(N, M, Nrho)
    gamma = 2.0 * (1.0 - 0.81)
    pre = OC.Precomp(N, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, gamma)
    prefac = 0.37
    K, Ne = 2, 2
    rng = np.random.default_rng(31)
    u0 = np.stack([_field(vm, 40 + s, u=(0.2 - 0.1 * s, 0.05 * s, 0.0), T=0.9 + 0.05 * s)
                   for s in range(K * Ne)])
    u = u0 * (1.0 + 1e-3 * rng.uniform(-1, 1, u0.shape))
    errs = {}
    for dt in (_lib.F64, _lib.F32):
        td = torch.float64 if dt == _lib.F64 else torch.float32
        plan = CollisionPlan(N, M, Nrho, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, gamma, prefac,
                             dt, _lib.VHS, ntask_max=4, device=cuda)
        plan64 = CollisionPlan(N, M, Nrho, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, gamma, prefac,
                               _lib.F64, _lib.VHS, ntask_max=4, device=cuda)
        plan.set_tasks(K, Ne, NODAL)
        plan64.set_tasks(K, Ne, NODAL)
        ud = torch.tensor(u, dtype=td, device=cuda)
        u0d = torch.tensor(u0, dtype=td, device=cuda)
        Q0 = torch.zeros(u.shape, dtype=torch.float64, device=cuda)
        plan64(u0d.double(), Q0)                       # Q(f0, f0) of the stored reference
        plan.reference_set(u0d)
        Qr = torch.zeros_like(ud)
        plan.reference(ud, u0d, Q0, Qr)
        Qd = torch.zeros_like(ud)
        plan(ud, Qd)
        torch.cuda.synchronize()
        ur = ud.double().cpu().numpy()                 # the operands as stored
        ref = np.stack([OC.q_vhs(pre, ur[s], ur[s], prefac) for s in range(K * Ne)])
        scale = np.abs(ref).max()
        errs[dt] = (np.abs(Qr.double().cpu().numpy() - ref).max() / scale,
                    np.abs(Qd.double().cpu().numpy() - ref).max() / scale)
        plan.close()
        plan64.close()
    print('reference form', N, 'fp64 (ref, direct)', errs[_lib.F64], 'fp32', errs[_lib.F32])
    assert errs[_lib.F64][0] < 1e-12
    ref32, dir32 = errs[_lib.F32]
    assert ref32 < 2e-7 and ref32 < dir32 / 30


def test_collision_calls_reproducible(cuda):
    """The FP32 N = 32 lane-pair gain: repeated direct calls, and repeated
    reference-state calls after storing the same reference state again, give
    bitwise equal results (every sum runs in a fixed order; no atomics)."""
    import torch
    from dgfs1d_gpu_amd import _lib
    from dgfs1d_gpu_amd.collision import CollisionPlan
    N, M, Nrho = 32, 6, 4
    vm = _mesh(N, M, Nrho)
    plan = CollisionPlan(N, M, Nrho, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, 0.38, 0.37,
                         _lib.F32, _lib.VHS, ntask_max=4, device=cuda)
    plan.set_tasks(2, 2, NODAL)
    u = torch.tensor(np.stack([_field(vm, 60 + s) for s in range(4)]), dtype=torch.float32,
                     device=cuda)
    u0 = (u * 0.999).contiguous()
    q0 = torch.zeros(u.shape, dtype=torch.float64, device=cuda)
    outs = {'direct': [], 'reference': []}
    for _ in range(3):
        plan.reference_set(u0)
        o = torch.zeros_like(u)
        plan(u, o)
        outs['direct'].append(o)
        o = torch.zeros_like(u)
        plan.reference(u, u0, q0, o)
        outs['reference'].append(o)
    torch.cuda.synchronize()
    plan.close()
    for name, os_ in outs.items():
        assert all(torch.equal(os_[0], x) for x in os_[1:]), name
