"""GPU parity of the dgfsBi1D operator path against the FP64 oracle.

Covers the per-species transport with mass-dependent wall Maxwellians, the
four pair collisions with the U/V contractions, invMass and the four explicit
mixture integrators.  Decks are small (N = 8, Mpre = 32) so the oracle
(oracle/bideck.py, oracle/bivss.py) finishes in seconds.  Tolerance: FP64
relative 1e-11 of max|value|.
"""
import numpy as np
import pytest

from oracle.bideck import BiDeck

pytestmark = pytest.mark.gpu

BASE = """
[config]
precision = double
[mesh]
xlo = 0.
xhi = 1e-3
Ne = 2
H0 = 1e-3
[basis]
kind = nodal-sem-gll
K = 3
Nqr = 4
[non-dim]
T0 = 273
rho0 = 0.00011140244750324529
molarMass0 = 39.96e-3
[velocity-mesh]
nspcs = 2
Nv = 8
Nrho = 4
dev = 5
cmax = 0
Tmax = 273
molarMass1 = 39.96e-3
molarMass2 = 83.77e-3
[spherical-design-rule]
ssrule = womersley
M = 6
[scattering-model]
type = vss
ssrulepre = womersley
Mpre = 32
alpha00 = 1.4
omega00 = 0.81
dRef00 = 4.11e-10
Tref00 = 273
alpha01 = 1.36
omega01 = 0.805
dRef01 = 4.405e-10
Tref01 = 273
alpha10 = 1.36
omega10 = 0.805
dRef10 = 4.405e-10
Tref10 = 273
alpha11 = 1.32
omega11 = 0.80
dRef11 = 4.7e-10
Tref11 = 273
[time-integrator]
scheme = euler
tstart = 0.0
tend = 0.03
dt = 0.01
[soln-ics]
type = maxwellian
nden1 = 1.6802782428845444e+21
nden2 = 8.008802839916987e+20
T = 273
ux = 0
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-wall-diffuse
T = 263
ux = 0
uy = 0
uz = 0
[soln-bcs-xhi]
type = dgfs-wall-diffuse
T = 283
ux = 0
uy = 0
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
    import torch
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    text = _ini(over)
    deck = BiDeck(text)
    solver = DGFSBi1D(Dictionary(text), device=cuda, writers=False)
    rng = np.random.default_rng(seed)
    us = [u * (1.0 + 0.1 * rng.uniform(-1, 1, u.shape)) for u in deck.u0]
    for p in range(2):
        solver.us[p].copy_(torch.tensor(us[p].ravel(), dtype=solver.tdtype, device=cuda))
    return deck, solver, us


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_bi_initial_state_matches(cuda):
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    text = _ini({})
    deck = BiDeck(text)
    s = DGFSBi1D(Dictionary(text), device=cuda, writers=False)
    for p in range(2):
        err = _rel(s.us[p].cpu().numpy().reshape(deck.u0[p].shape), deck.u0[p])
        print('bi ic', p, err, flush=True)
        assert err < 1e-13


@pytest.mark.parametrize('over', [
    {},
    {('soln-bcs-xlo', 'uy'): -50.0, ('soln-bcs-xhi', 'uy'): 50.0, ('soln-bcs-xlo', 'T'): 273,
     ('soln-bcs-xhi', 'T'): 273},
    {('basis', 'kind'): 'modal-orthonormal-legendre'},
    {('soln-bcs-xlo', 'type'): 'dgfs-periodic', ('soln-bcs-xhi', 'type'): 'dgfs-periodic'},
], ids=['fourier', 'couette', 'modal', 'periodic'])
def test_bi_rhs(cuda, over):
    import torch
    deck, s, us = _pair(cuda, over)
    for p in range(2):
        out = torch.zeros_like(s.us[p])
        s.rhs(p, 0.0, s.us, out)
        torch.cuda.synchronize()
        ref = deck.op.rhs(p, 0.0, us)
        err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
        print('bi rhs', over, p, err, flush=True)
        assert err < 1e-11


@pytest.mark.parametrize('scheme', ['euler', 'ssp-rk2', 'ssp-rk3', 'lse-rk45'])
def test_bi_steps(cuda, scheme):
    deck, s, us = _pair(cuda, {('time-integrator', 'scheme'): scheme})
    t = 0.0
    start = [u.copy() for u in us]
    for k in range(2):
        s.step()
        us = deck.step(t, us)
        t += deck.dt
        for p in range(2):
            err = _rel(s.us[p].cpu().numpy().reshape(us[p].shape), us[p])
            print('bi step', scheme, k, p, err, flush=True)
            assert err < 1e-11
    for p in range(2):        # two steps move each species' state by about 2% of its size
        assert np.abs(us[p] - start[p]).max() > 1e-3 * np.abs(start[p]).max()


def test_bi_writers_and_restart(cuda, tmp_path):
    """Moments, residual and distribution writers run; a restart from the
    written distribution reproduces the state."""
    import torch
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    over = {('time-integrator', 'scheme'): 'ssp-rk2'}
    text = _ini(over) + """
[dgfsresidual]
nsteps = 1
[dgfsmomwriter]
dt-out = 0.01
basedir = %s
basename = mom-{t:.2f}
[dgfsdistwriter]
dt-out = 0.02
basedir = %s
basename = dist-{t:.2f}
""" % (tmp_path, tmp_path)
    s = DGFSBi1D(Dictionary(text), device=cuda)
    s.run()
    assert s.nacptsteps == 3
    assert len(s.residual.history) == 3 and all(np.all(r > 0) for _, _, r in s.residual.history)
    mom = np.loadtxt(tmp_path / 'mom-0.03.txt')
    assert mom.shape == (2 * 4, 1 + 14)
    dist = tmp_path / 'dist-0.02.h5_rank=0'         # HDF5, the reference's name and layout
    assert dist.exists()
    from dgfs1d_gpu_amd import h5
    assert sorted(h5.read(str(dist))) == ['coeff0', 'coeff1']
    r = DGFSBi1D(Dictionary(text), device=cuda, restart=str(dist),
                 writers=False)
    assert abs(r.time - 0.02) < 1e-12
    s2 = DGFSBi1D(Dictionary(text), device=cuda, writers=False)
    s2.run(2)
    for p in range(2):
        assert torch.equal(r.us[p], s2.us[p])


def test_bi_mixture_projection_entry(cuda):
    """dgfs_project_conservative_mixture against oracle.relax.project_mixture on
    random slices (FP64 and FP32 data, FP64 moments and solve)."""
    import torch
    from dgfs1d_gpu_amd.bi.velocitymesh import DGFSVelocityMeshBi
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.relax import MixtureProjection
    from oracle.relax import project_mixture
    for prec, tol in (('double', 1e-12), ('single', 2e-5)):
        cfg = Dictionary(_ini({('config', 'precision'): prec}))
        vm = DGFSVelocityMeshBi(cfg, device=cuda)
        td = torch.float64 if prec == 'double' else torch.float32
        m0, m1 = vm.masses()
        rng = np.random.default_rng(11)
        ns, n = 5, vm.vsize()
        Q0 = rng.standard_normal((ns, n))
        Q1 = rng.standard_normal((ns, n))
        Q0 = Q0.astype(np.float32).astype(np.float64) if prec == 'single' else Q0
        Q1 = Q1.astype(np.float32).astype(np.float64) if prec == 'single' else Q1
        P = MixtureProjection(vm.cv(), vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), m0, m1, td, cuda)
        d0 = torch.tensor(Q0, dtype=td, device=cuda)
        d1 = torch.tensor(Q1, dtype=td, device=cuda)
        o0, o1 = torch.zeros_like(d0), torch.zeros_like(d1)
        P(d0, d1, o0, o1, ns, accumulate=False)
        torch.cuda.synchronize()
        cvd = np.asarray(vm.cv(), dtype=np.float64 if prec == 'double' else np.float32)
        for s in range(ns):
            r0, r1 = project_mixture(Q0[s], Q1[s], cvd.astype(np.float64), m0, m1)
            e = max(np.abs(o0[s].double().cpu().numpy() - r0).max(),
                    np.abs(o1[s].double().cpu().numpy() - r1).max()) / np.abs(Q0[s]).max()
            print('mixture projection', prec, s, e, flush=True)
            assert e < tol


@pytest.mark.parametrize('scheme', ['euler', 'ssp-rk2'])
def test_bi_steps_projected(cuda, scheme):
    """The solver with [scattering-model] projection = conservative against the
    oracle with the mixture projection of every node's collision terms."""
    over = {('time-integrator', 'scheme'): scheme, ('scattering-model', 'projection'):
            'conservative'}
    deck, s, us = _pair(cuda, over)
    t = 0.0
    for k in range(2):
        s.step()
        us = deck.step(t, us)
        t += deck.dt
        for p in range(2):
            err = _rel(s.us[p].cpu().numpy().reshape(us[p].shape), us[p])
            print('bi projected step', scheme, k, p, err, flush=True)
            assert err < 1e-11


@pytest.mark.parametrize('over,tol', [
    ({('config', 'precision'): 'single'}, 2e-5),
    ({('soln-bcs-xlo', 'type'): 'dgfs-inlet', ('soln-bcs-xlo', 'nden1'): 1.7e21,
      ('soln-bcs-xlo', 'nden2'): 7.9e20, ('soln-bcs-xlo', 'ux'): 30.0}, 1e-11),
], ids=['fp32', 'inlet'])
def test_bi_rhs_more(cuda, over, tol):
    """FP32 state and collisions (config 5's precision) against the FP64
    oracle, and the mixture inlet (no velocity shift in its fac)."""
    import torch
    deck, s, us = _pair(cuda, over)
    for p in range(2):
        out = torch.zeros_like(s.us[p])
        s.rhs(p, 0.0, s.us, out)
        torch.cuda.synchronize()
        ref = deck.op.rhs(p, 0.0, [s.us[q].double().cpu().numpy().reshape(us[q].shape)
                                    for q in range(2)])
        err = _rel(out.double().cpu().numpy().reshape(ref.shape), ref)
        print('bi rhs', list(over.values())[0], p, err, flush=True)
        assert err < tol


def test_bi_expr_wall_matches_diffuse_wall(cuda):
    """dgfs-wall-expr-diffuse with constant expressions equals
    dgfs-wall-diffuse with the same wall state (the expression wall's
    exp(-m |c-u|^2/T) enters the kernel as T/m)."""
    import torch
    over = {('soln-bcs-xlo', 'T'): 263, ('soln-bcs-xlo', 'uy'): -40.0}
    _, s1, us = _pair(cuda, over)
    over2 = dict(over)
    over2[('soln-bcs-xlo', 'type')] = 'dgfs-wall-expr-diffuse'
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    s2 = DGFSBi1D(Dictionary(_ini(over2)), device=cuda, writers=False)
    for p in range(2):
        s2.us[p].copy_(s1.us[p])
    for p in range(2):
        o1, o2 = torch.zeros_like(s1.us[p]), torch.zeros_like(s2.us[p])
        s1.rhs(p, 0.0, s1.us, o1)
        s2.rhs(p, 0.0, s2.us, o2)
        torch.cuda.synchronize()
        err = float((o1 - o2).abs().max() / o1.abs().max())
        print('bi expr wall', p, err, flush=True)
        assert err < 1e-12
