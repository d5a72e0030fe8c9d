"""Many independent dgfsStd1D problems of uneven size advanced together.

No reference counterpart (the reference runs one deck per process,
dgfs1D/std/std.py:37-423).  Design:

* every problem's state, stage registers and RHS live in ONE flat device
  tensor, problems ordered by velocity grid N, so that
  - the explicit integrator's stage updates are one axnpby launch for the
    whole batch,
  - the collisions of all problems that share a velocity grid (N, M, N_rho,
    L, prefactor) run as ONE batched dgfs_collide launch sequence over a task
    table that spans those problems (tasks may address any slice of the group);
* the transport of every problem runs as one batched launch sequence
  (dgfs_transport_batch_rhs: three launches for the whole batch, each
  problem with its own elements, basis and walls);
* a whole time step can be captured once in a HIP graph and replayed;
* ``autotune`` times, for every collision plan on its real task table, the
  gain kernel's schedules -- q-group count, residue split (the smallest that
  fits LDS or twice it) and work-group order (XCD-aware or dispatch order) --
  and keeps the fastest (dgfs_collision_plan_set_qgroups, _set_schedule).
"""
import time

import numpy as np

from . import _lib
from .collision import CollisionPlan, task_table
from .dictionary import Dictionary
from .std.scattering import _fold_same_slice
from .std.std import DGFSStd1D
from .transport import BatchTransport

R0 = 8.3144598


def problem_deck(kind, Ne, K, N, uw, Tl, Tr, precision='float', dt=2e-4, M=12):
    """Couette (walls at T0*Tl, T0*Tr moving at -/+ uw u0 along y) or Fourier
    (walls at rest at T0*Tl and T0*Tr) deck; VHS argon, Kn = 0.1 as the
    production Couette case, N_rho = N/2, M = 12 (octahedral design)."""
    u0 = np.sqrt(2 * R0 / 39.96e-3 * 273.0)
    uy = uw * u0 if kind == 'couette' else 0.0
    rule = 'octahedral' if M == 12 else 'womersley'
    return f"""[config]
precision = {precision}
[mesh]
xlo = -0.5
xhi = 0.5
Ne = {Ne}
H0 = 1.0
[basis]
kind = nodal-sem-gll
K = {K}
Nqr = 4
[non-dim]
T0 = 273
rho0 = 8.588900796810341e-07
molarMass0 = 39.96e-3
[velocity-mesh]
Nv = {N}
Nrho = {N // 2}
dev = 5
cmax = 0
Tmax = 273
[spherical-design-rule]
ssrule = {rule}
M = {M}
[scattering-model]
type = vhs
omega = 0.81
dRef = 4.17e-10
Tref = 273
muRef = 2.117e-5
[time-integrator]
scheme = ssp-rk2
controller = none
tstart = 0.0
tend = 1.0
dt = {dt}
[soln-ics]
type = maxwellian
rho = 8.588900796810341e-07
T = 273
ux = 0
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-wall-diffuse
T = {273.0 * Tl}
ux = 0
uy = {-uy}
uz = 0
[soln-bcs-xhi]
type = dgfs-wall-diffuse
T = {273.0 * Tr}
ux = 0
uy = {uy}
uz = 0
"""


def make_problems(seed=0, count=64):
    """The seeded uneven workload: Couette or Fourier cases, |u_w| ~ U[0.1, 1]
    with a random sign, wall temperatures ~ U[0.8, 1.5], Ne in {2, 4, 8},
    K in {3, 5, 7}, N in {16, 24, 32}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = 'couette' if rng.uniform() < 0.5 else 'fourier'
        uw = float(rng.uniform(0.1, 1.0) * (1 if rng.uniform() < 0.5 else -1))
        Tl = float(rng.uniform(0.8, 1.5))
        Tr = Tl if kind == 'couette' else float(rng.uniform(0.8, 1.5))
        out.append(dict(kind=kind, Ne=int(rng.choice([2, 4, 8])), K=int(rng.choice([3, 5, 7])),
                        N=int(rng.choice([16, 24, 32])), uw=uw, Tl=Tl, Tr=Tr))
    return out


class BatchedStd1D:
    """Flat-tensor batch of DGFSStd1D problems sharing one dt and scheme."""

    def __init__(self, specs, device, precision='float', dt=2e-4, batched_transport=True):
        import torch
        self.torch = torch
        self.device = device
        specs = sorted(specs, key=lambda p: p['N'])
        self.specs = specs
        self.solvers = []
        for p in specs:
            text = problem_deck(p['kind'], p['Ne'], p['K'], p['N'], p['uw'], p['Tl'], p['Tr'],
                                precision=precision, dt=dt)
            self.solvers.append(DGFSStd1D(Dictionary(text), device=device, writers=False))
        s0 = self.solvers[0]
        self.tdtype = s0.tdtype
        sizes = [s.u.numel() for s in self.solvers]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)])
        self.total = int(self.offsets[-1])
        self.u = torch.empty(self.total, dtype=self.tdtype, device=device)
        for s, a, b in zip(self.solvers, self.offsets[:-1], self.offsets[1:]):
            self.u[a:b].copy_(s.u)
        self.dt = float(dt)
        self.time = 0.0
        self.points = sum(s.K * s.Ne * s.vm.vsize() for s in self.solvers)
        self.btransport = None
        if batched_transport:
            self.btransport = BatchTransport([s.transport for s in self.solvers],
                                             self.offsets[:-1])
        self._groups()
        from .std.integrator import SSPRK2IntegratorStd
        self.odes = SSPRK2IntegratorStd(self.rhs, (1, 1, self.total), self.tdtype, device=device)
        self.graph = None

    def _groups(self):
        """One collision plan per velocity grid, tasks over the group's slices."""
        self.groups = []
        i = 0
        while i < len(self.solvers):
            j = i
            N = self.specs[i]['N']
            while j < len(self.solvers) and self.specs[j]['N'] == N:
                j += 1
            sl = self.solvers[i:j]
            vm = sl[0].vm
            n = vm.vsize()
            base = int(self.offsets[i])
            tasks = []
            for k, s in enumerate(sl):
                off = (int(self.offsets[i + k]) - base) // n
                for o, f, g in s.basis.collision_tasks(s.Ne):
                    tasks.append((o + off, [(a + off, c) for a, c in f],
                                  [(b + off, c) for b, c in g]))
            nsl = (int(self.offsets[j]) - base) // n
            sm = sl[0].sm
            dt = _lib.F64 if self.tdtype == self.torch.float64 else _lib.F32
            plan = CollisionPlan(vm.Nv(), vm.M(), vm.Nrho(), vm.L(), vm.qz(), vm.qw(), vm.sz(),
                                 vm.sw(), 2.0 * (1.0 - sm.omega), sm.prefac, dt, _lib.VHS,
                                 ntask_max=len(tasks), device=self.device)
            plan.set_tasks(1, nsl, _fold_same_slice(tasks))
            self.groups.append(dict(N=N, lo=base, hi=int(self.offsets[j]), plan=plan,
                                    solvers=list(range(i, j)), ntask=len(tasks)))
            i = j

    def rhs(self, t, u_in, u_out):
        if self.btransport is not None:
            self.btransport(u_in, u_out)
        else:
            for k, s in enumerate(self.solvers):
                a, b = int(self.offsets[k]), int(self.offsets[k + 1])
                s.transport(t, u_in[a:b], u_out[a:b])
        for g in self.groups:
            g['plan'](u_in[g['lo']:g['hi']], u_out[g['lo']:g['hi']])

    def step(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self.odes.integrate(self.time, self.dt, self.u)
        self.time += self.dt

    def capture(self, warmup=2):
        """Capture one step (every launch of it) in a HIP graph; later steps
        replay it.  Boundary data are time-independent for these decks."""
        torch = self.torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.odes.integrate(self.time, self.dt, self.u)
                self.time += self.dt
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.odes.integrate(self.time, self.dt, self.u)
        self.graph = g
        self.graph.replay()          # the captured step itself has not run yet
        self.time += self.dt

    def autotune(self, candidates=(1, 2, 3, 4, 6, 12), splits=(1, 2), remaps=(1, 0), reps=3):
        """Per collision plan: time dgfs_collide for every schedule (q-group
        count dividing M, residue-split multiplier, work-group order) and keep
        the fastest.  Returns {N: ((qg, split, remap), {schedule: ms})}."""
        torch = self.torch
        out = {}
        scratch = torch.zeros_like(self.u)
        for g in self.groups:
            plan = g['plan']
            u, Q = self.u[g['lo']:g['hi']], scratch[g['lo']:g['hi']]
            times = {}
            for qg in candidates:
                if plan.M % qg:
                    continue
                for sm in splits:
                    for rm in remaps:
                        plan.set_qgroups(qg)
                        plan.set_schedule(sm, rm)
                        plan(u, Q)
                        torch.cuda.synchronize()
                        t0 = time.perf_counter()
                        for _ in range(reps):
                            plan(u, Q)
                        torch.cuda.synchronize()
                        times[(qg, sm, rm)] = (time.perf_counter() - t0) / reps * 1e3
            best = min(times, key=times.get)
            plan.set_qgroups(best[0])
            plan.set_schedule(best[1], best[2])
            out[g['N']] = (best, times)
        return out

    def solver_state(self, k):
        a, b = int(self.offsets[k]), int(self.offsets[k + 1])
        return self.u[a:b]
