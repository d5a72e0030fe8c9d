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
                                    ('modal-modified-karniadakis', 3), ('nodal-sem-gll-2', 3)])
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
def test_rhs_shakov(cuda, kind):
    """'shakov-direct-gll' plugin in the full RHS (std/scattering.py:1636-1833)."""
    import torch
    deck, solver, u = _pair(cuda, {('scattering-model', 'type'): 'shakov-direct-gll',
                                   ('scattering-model', 'muRef'): 2.117e-5,
                                   ('basis', 'kind'): kind})
    out = torch.zeros_like(solver.u)
    solver.rhs(0.0, solver.u, out)
    ref = deck.op.rhs(0.0, u)
    err = _rel(out.cpu().numpy().reshape(ref.shape), ref)
    print('shakov', kind, err)
    assert err < 1e-11


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
