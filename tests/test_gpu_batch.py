"""Batched uneven workload: many problems in flat tensors with one collision
launch sequence per velocity grid, one batched transport launch sequence,
HIP-graph capture and the q-group autotuner.

* The batched transport equals each problem's own transport exactly.

* Batched SSP-RK2 steps equal the problems stepped one by one (FP64, 1e-12).
* Every replay of the captured step equals an eager step from the same state
  (several replays, eager work and synchronisations in between).
* Every gain schedule (q-group count, residue split, work-group order) and
  the autotuner's choice keep results unchanged (exact up to summation
  order: FP64 1e-12).
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


def test_batched_transport_equals_per_problem(cuda):
    """dgfs_transport_batch_rhs (three launches for the batch) against each
    problem's own dgfs_transport_rhs on a perturbed state: identical outputs
    and diffuse-wall densities (same arithmetic and summation order)."""
    import torch
    b = _batch(cuda)
    g = torch.Generator(device='cpu').manual_seed(5)
    u = (b.u * (1 + 0.1 * (2 * torch.rand(b.u.numel(), generator=g, dtype=torch.float64)
                           - 1).to(cuda))).contiguous()
    out_b = torch.zeros_like(u)
    b.btransport(u, out_b)
    nw_b = b.btransport.wall_densities()
    out_p = torch.zeros_like(u)
    for k, s in enumerate(b.solvers):
        a, e = int(b.offsets[k]), int(b.offsets[k + 1])
        s.transport(0.0, u[a:e], out_p[a:e])
        nw = s.transport.wall_density()
        assert np.array_equal(nw, nw_b[k]), (k, nw, nw_b[k])
    torch.cuda.synchronize()
    assert torch.equal(out_b, out_p)


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
    # every schedule (residue split, work-group order) gives the same result
    for sm in (1, 2):
        for rm in (1, 0):
            got = torch.zeros_like(b.u)
            for g in b.groups:
                g['plan'].set_schedule(sm, rm)
                g['plan'](b.u[g['lo']:g['hi']], got[g['lo']:g['hi']])
            err = float((got - ref).abs().max() / ref.abs().max())
            print('schedule', sm, rm, err)
            assert err < 1e-12
    res = b.autotune(candidates=(1, 2, 3, 6), reps=2)
    print('autotune', res)
    got = torch.zeros_like(b.u)
    for g in b.groups:
        g['plan'](b.u[g['lo']:g['hi']], got[g['lo']:g['hi']])
    err = float((got - ref).abs().max() / ref.abs().max())
    assert err < 1e-12


# the seeded workload's grids reduced so that the FP64 oracle is cheap
REDUCED_N = {16: 6, 24: 8, 32: 12}


def test_seeded_batch_groups_vs_oracle(cuda):
    """All 64 problems of the seeded uneven workload (make_problems(0, 64):
    Couette / Fourier, Ne in {2, 4, 8}, K in {3, 5, 7}), each on a reduced
    velocity grid (N 16/24/32 -> 6/8/12, N_rho = N/2, M = 12), batched in FP64:
    one RHS of the whole batch from a perturbed state, every problem against
    the oracle's RHS of the same deck (1e-11 of max|RHS|)."""
    import torch
    from dgfs1d_gpu_amd.batch import BatchedStd1D, make_problems, problem_deck
    from oracle.deck import Deck
    specs = [dict(p, N=REDUCED_N[p['N']]) for p in make_problems(0, 64)]
    b = BatchedStd1D(specs, cuda, precision='double', dt=2e-4)
    rng = np.random.default_rng(17)
    u = b.u.cpu().numpy() * (1.0 + 0.05 * rng.uniform(-1, 1, b.total))
    b.u.copy_(torch.tensor(u, dtype=torch.float64, device=cuda))
    out = torch.zeros_like(b.u)
    b.rhs(0.0, b.u, out)
    got = out.cpu().numpy()
    worst = 0.0
    seen = set()
    for k, p in enumerate(b.specs):
        d = Deck(problem_deck(p['kind'], p['Ne'], p['K'], p['N'], p['uw'], p['Tl'], p['Tr'],
                              precision='double', dt=2e-4))
        lo, hi = int(b.offsets[k]), int(b.offsets[k + 1])
        ref = d.op.rhs(0.0, u[lo:hi].reshape(d.u0.shape))
        err = float(np.abs(got[lo:hi] - ref.ravel()).max() / np.abs(ref).max())
        worst = max(worst, err)
        seen.add((p['N'], p['K'], p['Ne']))
    print('64 problems, %d (N, K, Ne) groups, worst RHS error %.3e' % (len(seen), worst))
    assert worst < 1e-11


def test_seeded_batch_full_size(cuda):
    """The seeded workload at its real sizes (27.13 M phase-space points,
    FP32, autotuned plans): a captured step replayed equals an eager step
    from the same state bit for bit, the state stays finite and every
    problem's mass (diffuse walls carry no net mass flux) stays within 1e-5
    over 3 steps."""
    import torch
    from dgfs1d_gpu_amd.batch import BatchedStd1D, make_problems
    b = BatchedStd1D(make_problems(0, 64), cuda, precision='float')
    assert b.points > 2.7e7
    b.autotune(candidates=(1, 2), reps=1)

    def masses():
        out = []
        for k, s in enumerate(b.solvers):
            x = b.solver_state(k).double().cpu().numpy().reshape(s.K, s.Ne, -1).sum(-1)
            jac = np.asarray(s.mesh.jac).reshape(-1)          # nodal-sem-gll: GLL weights
            out.append(float((x * s.basis.w[:, None] * jac[None, :]).sum()))
        return np.array(out)
    m0 = masses()
    b.capture(warmup=1)
    u0 = b.u.clone()
    t0 = b.time
    b.graph.replay()
    torch.cuda.synchronize()
    g = b.u.clone()
    b.u.copy_(u0)
    b.odes.integrate(t0, b.dt, b.u)
    torch.cuda.synchronize()
    assert torch.equal(g, b.u)
    b.time = t0 + b.dt
    b.step()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(b.u).all())
    drift = np.abs(masses() - m0) / np.abs(m0)
    print('points', b.points, 'mass drift over 3 steps', drift.max())
    assert drift.max() < 1e-5
