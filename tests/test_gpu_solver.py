"""GPU parity of the full dgfsStd1D operator path against the FP64 oracle.

Covers the fused transport (all boundary kinds), the batched collisions for
all five bases (nodal and modal U/V contractions), invMass, the four explicit
integrators and the FP64 moments.  Decks are small (N = 8) so the oracle
finishes in seconds.  Tolerance: FP64 relative 1e-11 of max|value|.
"""
import numpy as np
import pytest

from oracle import std as OS
from oracle.deck import Deck

pytestmark = pytest.mark.gpu

BASE = """
[config]
precision = double
[mesh]
xlo = -0.5
xhi = 0.5
Ne = 3
H0 = 1.0
[basis]
kind = nodal-sem-gll
K = 3
Nqr = 4
[non-dim]
T0 = 273
rho0 = 8.588900796810341e-07
molarMass0 = 39.96e-3
[velocity-mesh]
Nv = 8
Nrho = 4
dev = 5
cmax = 0
Tmax = 273
[spherical-design-rule]
ssrule = womersley
M = 6
[scattering-model]
type = vhs
omega = 0.81
dRef = 4.17e-10
Tref = 273
[time-integrator]
scheme = euler
tstart = 0.0
tend = 0.01
dt = 0.001
[soln-ics]
type = maxwellian
rho = 8.588900796810341e-07
T = 273
ux = 0
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-wall-diffuse
rho = 8.588900796810341e-07
T = 300
ux = 0
uy = -168.52743267214112
uz = 0
[soln-bcs-xhi]
type = dgfs-wall-diffuse
rho = 8.588900796810341e-07
T = 250
ux = 0
uy = 168.52743267214112
uz = 0
"""


def _ini(over):
    from configparser import ConfigParser
    import io
    cp = ConfigParser(inline_comment_prefixes=[';'])
    cp.optionxform = str
    cp.read_string(BASE)
    for (s, k), v in over.items():
        cp.set(s, k, str(v))
    buf = io.StringIO()
    cp.write(buf)
    return buf.getvalue()


def _pair(cuda, over, seed=3):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    text = _ini(over)
    deck = Deck(text)
    solver = DGFSStd1D(Dictionary(text), device=cuda, writers=False)
    rng = np.random.default_rng(seed)
    u = deck.u0 * (1.0 + 0.1 * rng.uniform(-1, 1, deck.u0.shape))
    solver.u.copy_(_t(u, solver))
    return deck, solver, u


def _t(x, solver):
    import torch
    return torch.tensor(np.ascontiguousarray(x).ravel(), dtype=solver.tdtype,
                        device=solver.device)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize('kind,K', [('nodal-sem-gll', 3), ('nodal-sem-gll', 4), ('nodal-gll', 3),
                                    ('modal-orthonormal-legendre', 3),
                                    ('modal-modified-karniadakis', 3), ('nodal-sem-gll-2', 3),
                                    ('nodal-sem-gll', 5), ('nodal-sem-gll', 7),
                                    ('nodal-gll', 5), ('modal-orthonormal-legendre', 5),
                                    ('modal-modified-karniadakis', 5)])
def test_rhs_bases(cuda, kind, K):
    import torch
    deck, solver, u = _pair(cuda, {('basis', 'kind'): kind, ('basis', 'K'): K})
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print(kind, K, err)
    assert err < 1e-11


@pytest.mark.parametrize('kind', ['nodal-sem-gll', 'modal-modified-karniadakis'])
def test_initial_state(cuda, kind):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    text = _ini({('basis', 'kind'): kind})
    deck = Deck(text)
    solver = DGFSStd1D(Dictionary(text), device=cuda, writers=False)
    err = _rel(solver.u.cpu().numpy().reshape(deck.u0.shape), deck.u0)
    assert err < 1e-13


@pytest.mark.parametrize('lo,hi', [('dgfs-wall-diffuse', 'dgfs-inlet'),
                                   ('dgfs-periodic', 'dgfs-vanish'),
                                   ('dgfs-inlet', 'dgfs-cyclic')])
@pytest.mark.parametrize('model', ['vhs', 'vhs-gll'])
def test_rhs_boundaries(cuda, lo, hi, model):
    import torch
    deck, solver, u = _pair(cuda, {('soln-bcs-xlo', 'type'): lo, ('soln-bcs-xhi', 'type'): hi,
                                   ('scattering-model', 'type'): model})
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print(lo, hi, model, err)
    assert err < 1e-11


@pytest.mark.parametrize('kind', ['nodal-sem-gll', 'nodal-gll'])
@pytest.mark.parametrize('model', ['shakov-direct-gll', 'bgk-direct-gll', 'esbgk-direct-gll'])
def test_rhs_shakov(cuda, kind, model):
    """The relaxation models in the full RHS: 'shakov-direct-gll'
    (std/scattering.py:1636-1833), 'bgk-direct-gll' (:1186-1369; its oracle is
    written from the BGK kernels, a Maxwellian target, and a Pr in the deck is
    ignored) and 'esbgk-direct-gll' (:1372-1632; anisotropic Gaussian target,
    prefactor times Pr)."""
    import torch
    deck, solver, u = _pair(cuda, {('scattering-model', 'type'): model,
                                   ('scattering-model', 'muRef'): 2.117e-5,
                                   ('scattering-model', 'Pr'): 0.6,
                                   ('basis', 'kind'): kind})
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print(model, kind, err)
    assert err < 1e-11


@pytest.mark.parametrize('prec,tol', [('double', 1e-10), ('single', 2e-5)])
@pytest.mark.parametrize('model', ['bgk-gll', 'esbgk-gll'])
def test_rhs_bgk_conservative(cuda, prec, tol, model):
    """'bgk-gll' (std/scattering.py:445-706) and 'esbgk-gll' (:752-1185): the
    GPU's Newton solves for the conserving discrete Maxwellian / Gaussian
    against the oracle's MINPACK solves of the same equations (the ES one in
    the reference's parameters), in the full RHS; the collision output keeps
    each node's discrete mass, momentum and energy."""
    import torch
    deck, solver, u = _pair(cuda, {('scattering-model', 'type'): model,
                                   ('scattering-model', 'muRef'): 2.117e-5,
                                   ('config', 'precision'): prec})
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, solver.u.double().cpu().numpy().reshape(u.shape))
    err = _rel(out.double().cpu().numpy().reshape(ref.shape), ref)
    Q = torch.zeros_like(solver.u)
    solver.sm.collide(solver.u, Q)
    q = Q.double().cpu().numpy().reshape(-1, deck.vm.vsize)
    cv, cw = deck.vm.cv, deck.vm.cw
    phi = np.vstack([np.ones(cv.shape[1]), cv, (cv * cv).sum(0)]) * cw
    mom = np.abs(q @ phi.T).max() / (np.abs(q) @ np.abs(phi.T)).max()
    print(model, prec, err, 'moments of Q', mom, flush=True)
    assert err < tol and mom < (1e-11 if prec == 'double' else 1e-5)


@pytest.mark.parametrize('model', ['shakov-direct-gll', 'esbgk-direct-gll'])
def test_rhs_relaxation_fp32(cuda, model):
    """The FP32 instantiations of the relaxation kernels (FP64 moments and
    tensor): the RHS of the rounded state against the FP64 oracle, 2e-5."""
    import torch
    deck, solver, u = _pair(cuda, {('scattering-model', 'type'): model,
                                   ('scattering-model', 'muRef'): 2.117e-5,
                                   ('config', 'precision'): 'single'})
    assert solver.u.dtype == torch.float32
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, solver.u.double().cpu().numpy().reshape(u.shape))
    err = _rel(out.double().cpu().numpy().reshape(ref.shape), ref)
    print(model, 'fp32', err)
    assert err < 2e-5


@pytest.mark.parametrize('model', ['vhs', 'shakov-direct-gll'])
def test_rhs_projection(cuda, model):
    """projection = conservative: RHS parity and vanishing collision moments."""
    import torch
    over = {('scattering-model', 'type'): model, ('scattering-model', 'muRef'): 2.117e-5,
            ('scattering-model', 'projection'): 'conservative'}
    deck, solver, u = _pair(cuda, over)
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print('projection rhs', model, err)
    assert err < 1e-11
    # collision part alone (nodal basis: invMass is diagonal, apply it back)
    Q = torch.zeros_like(solver.u)
    solver.sm.collide(solver.u, Q)
    q = Q.cpu().numpy().reshape(solver.K * solver.Ne, -1)
    from oracle import relax as OR
    phi = OR.invariants(deck.vm.cv)
    m = np.abs(phi @ q.T) / (np.abs(phi) @ np.abs(q).T)
    print('collision moments after projection', m.max())
    assert m.max() < 1e-12


@pytest.mark.parametrize('scheme', ['euler', 'ssp-rk2', 'ssp-rk3', 'lse-rk45'])
def test_integrators(cuda, scheme):
    deck, solver, u = _pair(cuda, {('time-integrator', 'scheme'): scheme})
    t = 0.0
    for _ in range(3):
        u = deck.step(t, u)
        t += deck.dt
        solver.step()
    err = _rel(solver.u.cpu().numpy().reshape(u.shape), u)
    print(scheme, err)
    assert err < 1e-11


def test_moments(cuda):
    from dgfs1d_gpu_amd.transport import Moments
    deck, solver, u = _pair(cuda, {('velocity-mesh', 'Nv'): 12})
    mom = Moments(solver.basis.interpMat, solver.K, solver.Ne, solver.vm, cuda)
    got = mom(solver.u)
    ref = OS.moments(u, deck.basis.interp, deck.vm)
    scale = np.abs(ref).max(axis=(0, 1))
    err = float((np.abs(got - ref).max(axis=(0, 1)) / scale).max())
    print('moments', err)
    assert err < 1e-12


def test_writers_and_restart(cuda, tmp_path):
    """The moment writer's text table (std/moments.py:153-160, :195-220) at an
    output time holds the positions and the dimensional oracle moments of the
    state in the HDF5 distribution file written at that time
    (std/distribution.py:36-41); a restart from that file (std/std.py:235-247)
    resumes at its time and reproduces the uninterrupted run."""
    import torch
    from dgfs1d_gpu_amd import h5
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.moments import dimensional
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    text = _ini({('time-integrator', 'scheme'): 'ssp-rk2'}) + """
[dgfsresidual]
nsteps = 1
[dgfsmomwriter]
dt-out = 0.002
basedir = %s
basename = mom-{t:.3f}
[dgfsdistwriter]
dt-out = 0.002
basedir = %s
basename = dist-{t:.3f}
""" % (tmp_path, tmp_path)
    deck = Deck(_ini({}))
    s = DGFSStd1D(Dictionary(text), device=cuda)
    s.run(4)
    dist = tmp_path / 'dist-0.002.h5_rank=0'
    data = h5.read(str(dist))
    coeff, attrs = data['coeff']
    assert attrs == {'time': pytest.approx(0.002, abs=1e-15), 'K': 3, 'Ne': 3, 'Nv': 512}
    lines = open(tmp_path / 'mom-0.002.txt').read().splitlines()
    # numpy.savetxt of the reference's header 't={0} \n {1}' with comments '#'
    assert lines[0] == '#t=0.002 '
    assert lines[1] == '# x0, rho, U:x, U:y, T, Q:x, Q:y, P:xx, P:yy, P:xy, p'
    tab = np.loadtxt(tmp_path / 'mom-0.002.txt')
    nd = OS.moments(coeff.reshape(3, 3, 512), deck.basis.interp, deck.vm)   # (Nqr, Ne, 9)
    ref = dimensional(nd, s.vm).swapaxes(0, 1).reshape(-1, 10)
    err = float((np.abs(tab[:, 1:] - ref).max(axis=0) / np.abs(ref).max(axis=0)).max())
    print('moment table', err, flush=True)
    assert err < 1e-12
    # positions: the 4 Gauss-Lobatto points (+-1, +-1/sqrt 5) of each of the 3 elements
    r = np.array([-1.0, -1.0 / np.sqrt(5.0), 1.0 / np.sqrt(5.0), 1.0])
    x = -0.5 + (np.arange(3)[:, None] + 0.5 * (r[None, :] + 1.0)) / 3.0
    assert tab.shape == (4 * 3, 11)
    assert np.abs(tab[:, 0] - x.ravel()).max() < 1e-15
    r = DGFSStd1D(Dictionary(text), device=cuda, restart=str(dist), writers=False)
    assert abs(r.time - 0.002) < 1e-15
    r.run(2)
    assert torch.equal(r.u, s.u)


BKW = """
[config]
precision = double
[mesh]
xlo = 0.0
xhi = 1.0
Ne = 2
H0 = 1.0
[basis]
kind = nodal-sem-gll
K = 3
Nqr = 3
[non-dim]
T0 = 273
rho0 = 2.3e-4
molarMass0 = 39.96e-3
[velocity-mesh]
Nv = 24
Nrho = 12
dev = 6
cmax = 0
Tmax = 273
[spherical-design-rule]
ssrule = womersley
M = 6
[scattering-model]
type = maxwell-gll
eps = %.17g
[time-integrator]
scheme = ssp-rk3
tstart = 0.0
tend = 1.0
dt = 0.05
[soln-ics]
type = bkw
t0 = 6.0
[soln-bcs-xlo]
type = dgfs-periodic
[soln-bcs-xhi]
type = dgfs-periodic
"""


def _bkw_setup(cuda):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    from oracle import collision as OC
    from oracle import velocitymesh as OVM
    vm = OVM.VelocityMesh(1.0, 273.0, 2.3e-4, 39.96e-3, 24, 12, 0.0, 273.0, 6.0, 6)
    eps = 4.0 * np.pi * vm.u0 ** 2 / vm.n0              # prefactor H0 u0^2 / n0 / eps = 1/(4 pi)
    s = DGFSStd1D(Dictionary(BKW % eps), device=cuda, writers=False)
    pre = OC.Precomp(24, vm.L, vm.qz, vm.qw, vm.sz, vm.sw, 0.0)
    return s, vm, pre


def test_maxwell_gll_bkw_collision(cuda):
    """'maxwell-gll' (std/scattering.py:707-743) on the 'bkw' state: the GPU
    collision term of every node equals the oracle's (FP64, 1e-11) and the
    closed-form d f_BKW/dt (spectral accuracy at N = 24, 5e-4)."""
    import torch
    from oracle import collision as OC
    from tests.test_oracle_collision import bkw
    s, vm, pre = _bkw_setup(cuda)
    assert abs(s.sm.prefac * 4.0 * np.pi - 1.0) < 1e-13
    c2 = (vm.cv ** 2).sum(0)
    f, dfdt = bkw(c2, 6.0)
    Q = torch.zeros_like(s.u)
    s.sm.collide(s.u, Q)
    torch.cuda.synchronize()
    q = Q.cpu().numpy().reshape(-1, vm.vsize)
    ref = OC.q_vhs_gll(pre, f, 1.0 / (4.0 * np.pi))
    e_or = float(np.abs(q - ref).max() / np.abs(ref).max())
    e_cf = float(np.abs(q - dfdt).max() / np.abs(dfdt).max())
    print('bkw collision: oracle %.2e closed form %.2e' % (e_or, e_cf), flush=True)
    assert e_or < 1e-11 and e_cf < 5e-4


def test_maxwell_gll_bkw_relaxation(cuda):
    """A space-homogeneous run (periodic walls) from f_BKW(6) over t = 1
    (20 SSP-RK3 steps, the state changes by 37 %): every node stays uniform,
    matches the oracle's SSP-RK3 trajectory (1e-11) and the closed form
    f_BKW(7) (1e-4; the oracle's own error is 1.7e-5)."""
    from oracle import collision as OC
    from tests.test_oracle_collision import bkw
    s, vm, pre = _bkw_setup(cuda)
    s.run()
    assert s.nacptsteps == 20 and abs(s.time - 1.0) < 1e-12
    u = s.u.cpu().numpy().reshape(-1, vm.vsize)
    c2 = (vm.cv ** 2).sum(0)
    f, _ = bkw(c2, 6.0)
    Q = lambda g: OC.q_vhs_gll(pre, g, 1.0 / (4.0 * np.pi))
    for _ in range(20):
        u1 = f + 0.05 * Q(f)
        u2 = 0.75 * f + 0.25 * (u1 + 0.05 * Q(u1))
        f = f / 3.0 + 2.0 / 3.0 * (u2 + 0.05 * Q(u2))
    ref, _ = bkw(c2, 7.0)
    e_or = float(np.abs(u - f).max() / np.abs(f).max())
    e_cf = float(np.abs(u - ref).max() / np.abs(ref).max())
    print('bkw run: oracle %.2e closed form %.2e' % (e_or, e_cf), flush=True)
    assert e_or < 1e-11 and e_cf < 1e-4


EXPR_WALLS = {
    ('soln-bcs-xlo', 'type'): 'dgfs-wall-expr-diffuse',
    ('soln-bcs-xlo', 'ux'): '0',
    ('soln-bcs-xlo', 'uy'): '-168.5*sin(2*pi*t/0.004)',
    ('soln-bcs-xlo', 'uz'): '20*t/0.004',
    ('soln-bcs-xlo', 'T'): '273*(1 + 0.2*t/0.004)',
    ('soln-bcs-xhi', 'type'): 'dgfs-inlet-nondim',
    ('soln-bcs-xhi', 'rho'): 1.1,
    ('soln-bcs-xhi', 'ux'): -0.3,
    ('soln-bcs-xhi', 'T'): 1.2,
    # the inlet refuses a Maxwellian whose grid sum misses rho by more than
    # 1e-5 (std/bc.py:300-302): N = 12 resolves this one
    ('velocity-mesh', 'Nv'): 12,
}


@pytest.mark.parametrize('t', [0.0, 0.0013, 0.0031])
@pytest.mark.parametrize('model', ['vhs', 'vhs-gll'])
def test_rhs_expr_wall_and_nondim_inlet(cuda, t, model):
    """'dgfs-wall-expr-diffuse' (std/bc.py:136-198: wall velocity and
    temperature as expressions of t, evaluated at the stage time) on the left
    and 'dgfs-inlet-nondim' (:269-324: a non-dimensional Maxwellian inflow
    moving at ux) on the right, in the full RHS at three times, against the
    oracle."""
    import torch
    over = dict(EXPR_WALLS)
    over[('scattering-model', 'type')] = model
    deck, solver, u = _pair(cuda, over)
    out = torch.zeros_like(solver.u)
    solver.rhs(t, solver.u, out)
    ref = deck.op.rhs(t, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print('expr wall / nondim inlet', model, t, err)
    assert err < 1e-11


def test_expr_wall_steps(cuda):
    """Three SSP-RK3 steps with the time-dependent wall: the stage times reach
    the boundary (t, t + dt, t + dt/2), so the trajectory matches the oracle's
    only if every stage evaluates the expressions at its own time."""
    over = dict(EXPR_WALLS)
    over[('time-integrator', 'scheme')] = 'ssp-rk3'
    deck, solver, u = _pair(cuda, over)
    t = 0.0
    for _ in range(3):
        u = OS.integrate(deck.op, 'ssp-rk3', t, deck.dt, u)
        t += deck.dt
        solver.step()
    err = _rel(solver.u.cpu().numpy().reshape(u.shape), u)
    print('expr wall ssp-rk3 3 steps', err)
    assert err < 1e-11
