"""Host-side logic without a GPU: INI dictionary semantics, CLI overrides,
restricted expressions, plugin registry, spherical rules, velocity mesh."""
import os

import numpy as np
import pytest

from dgfs1d_gpu_amd.dictionary import Dictionary
from dgfs1d_gpu_amd.initialize import overrides_to_dict
from dgfs1d_gpu_amd.nputil import npeval, subclass_where


def test_dictionary_and_overrides():
    ini = '[config]\nprecision = float\n[mesh]\nNe = 2 ; comment\nxlo = -0.5\n'
    cfg = Dictionary(ini, defaults=overrides_to_dict([('mesh::Ne', '4'), ('basis::K', '5')]))
    assert cfg.lookupint('mesh', 'Ne') == 4 and cfg.lookupint('basis', 'K') == 5
    assert cfg.dtype == np.float32 and cfg.dtypename == 'float'
    assert cfg.lookupfloat('mesh', 'xlo') == np.float32(-0.5)
    assert cfg.lookupordefault('mesh', 'missing', 7) == 7
    assert cfg.lookupexpr('mesh', 'xlo') == '(-0.5)'


def test_npeval_restrictions():
    assert npeval('pow(x, 2) + sin(0)', {'x': 3.0}) == 9.0
    for bad in ('x**2', 'x.__class__', 'x; 1'):
        with pytest.raises(ValueError):
            npeval(bad, {'x': 1.0})
    with pytest.raises(SyntaxError):          # statements are not expressions
        npeval('import os', {'x': 1.0})


def test_registries():
    from dgfs1d_gpu_amd.basis import Basis
    from dgfs1d_gpu_amd.std.bc import DGFSBCStd
    from dgfs1d_gpu_amd.std.integrator import DGFSIntegratorStd
    from dgfs1d_gpu_amd.std.scattering import DGFSScatteringModelStd
    assert subclass_where(Basis, basis_kind='nodal-sem-gll').__name__ == 'NodalSemGLL'
    for k in ('euler', 'ssp-rk2', 'ssp-rk3', 'lse-rk45', 'lesrk-45'):
        subclass_where(DGFSIntegratorStd, intg_kind=k)
    for k in ('dgfs-wall-diffuse', 'dgfs-inlet', 'dgfs-periodic', 'dgfs-cyclic', 'dgfs-vanish',
              'dgfs-wall-expr-diffuse', 'dgfs-inlet-nondim'):
        subclass_where(DGFSBCStd, type=k)
    for k in ('vhs', 'vhs-gll', 'maxwell-gll', 'shakov-direct-gll', 'bgk-direct-gll',
              'esbgk-direct-gll', 'bgk-gll', 'esbgk-gll'):
        subclass_where(DGFSScatteringModelStd, scattering_model=k)
    from dgfs1d_gpu_amd.astd.scattering import DGFSScatteringModelAstd
    for k in ('shakov-direct-gll', 'bgk-direct-gll', 'esbgk-direct-gll'):
        subclass_where(DGFSScatteringModelAstd, scattering_model=k)
    with pytest.raises(KeyError):
        subclass_where(Basis, basis_kind='nope')


@pytest.mark.parametrize('npts', [2, 6, 12, 32, 48, 70, 94, 120, 156, 192])
def test_stored_designs_antipodal(npts):
    from dgfs1d_gpu_amd.sphericaldesign import get_sphquadrule
    p = get_sphquadrule('symmetric', 'womersley', npts)
    M = npts // 2
    assert p.shape == (npts, 3)
    assert np.abs(np.linalg.norm(p, axis=1) - 1).max() < 1e-12
    assert np.abs(p[M:] + p[:M]).max() < 1e-12


def test_octahedral_design():
    """24-point antipodal set, spherical 5-design (monomial averages exact)."""
    from math import gamma
    from itertools import product
    from dgfs1d_gpu_amd.sphericaldesign import get_sphquadrule
    from oracle.deck import octahedral24_oracle
    p = get_sphquadrule('symmetric', 'octahedral', 24)
    assert np.abs(p - octahedral24_oracle()).max() < 1e-15
    assert np.abs(p[12:] + p[:12]).max() == 0.0

    def avg(i, j, k):
        if i % 2 or j % 2 or k % 2:
            return 0.0
        return 2 * gamma((i + 1) / 2) * gamma((j + 1) / 2) * gamma((k + 1) / 2) / \
            gamma((i + j + k + 3) / 2) / (4 * np.pi)
    for i, j, k in product(range(6), repeat=3):
        if i + j + k <= 5:
            assert abs((p[:, 0] ** i * p[:, 1] ** j * p[:, 2] ** k).mean() - avg(i, j, k)) < 1e-14


def test_tabulated_rule():
    from dgfs1d_gpu_amd.sphericaldesign import get_sphquadrule
    p = get_sphquadrule('symmetric', '(0, 0, 1) (0, 0, -1)', 2)
    assert p.shape == (2, 3)


def test_velocity_mesh_matches_oracle():
    from dgfs1d_gpu_amd.std.velocitymesh import DGFSVelocityMeshStd
    from oracle.velocitymesh import VelocityMesh
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cfg = Dictionary.load(os.path.join(root, 'dgfs1d_gpu_amd', 'configs', 'couette_small.ini'))
    p = DGFSVelocityMeshStd(cfg)
    o = VelocityMesh(1.0, 273.0, 8.588900796810341e-07, 39.96e-3, 16, 8, 0.0, 273.0, 5.0, 6)
    assert abs(p.L() - o.L) < 1e-14 and abs(p.cw() - o.cw) < 1e-14
    assert np.abs(p.cv() - o.cv).max() < 1e-14
    assert np.abs(p.qz() - o.qz).max() < 1e-13 and np.abs(p.qw() - o.qw).max() < 1e-13
    assert np.abs(p.sz() - o.sz).max() == 0.0


def test_tensor_grid_detection():
    """The Newton solves' tensor layout (relax.tensor_grid_n): the velocity
    mesh's points are the tensor grid of its 1D nodes, in FP64 and rounded to
    FP32; a permuted, perturbed or non-cubic point set is not."""
    from dgfs1d_gpu_amd.std.velocitymesh import DGFSVelocityMeshStd
    from dgfs1d_gpu_amd.relax import tensor_grid_n
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cfg = Dictionary.load(os.path.join(root, 'dgfs1d_gpu_amd', 'configs', 'couette_small.ini'))
    cv = DGFSVelocityMeshStd(cfg).cv()
    n = cv.shape[1]
    N = int(round(n ** (1 / 3)))
    assert tensor_grid_n(cv[0], cv[1], cv[2], n) == N
    c32 = cv.astype(np.float32)
    assert tensor_grid_n(c32[0], c32[1], c32[2], n) == N
    assert tensor_grid_n(cv[1], cv[0], cv[2], n) == 0          # x and y swapped
    bad = cv[2].copy()
    bad[n // 2] += 1e-12
    assert tensor_grid_n(cv[0], cv[1], bad, n) == 0
    assert tensor_grid_n(cv[0][:-1], cv[1][:-1], cv[2][:-1], n - 1) == 0


def test_mesh_types():
    from dgfs1d_gpu_amd.mesh import Mesh
    base = '[mesh]\nxlo=0\nxhi=2\nNe=4\nH0=2\n'
    m = Mesh(Dictionary(base))
    assert np.allclose(m.xmesh, [0, 0.25, 0.5, 0.75, 1.0]) and np.allclose(m.invjac, 8.0)
    m = Mesh(Dictionary(base + 'type=bump\nmu=1.0\n'))
    assert m.Ne == 4 and np.isclose(m.xmesh[0], 0) and np.isclose(m.xmesh[-1], 1)


def test_dictionary_set_precision():
    from dgfs1d_gpu_amd.dictionary import Dictionary
    d = Dictionary('[config]\nprecision = float\n')
    assert d.dtype == np.float32
    d.set('config', 'precision', 'double')
    assert d.dtype == np.float64 and d.dtypename == 'double'


def test_batch_problem_generator():
    from dgfs1d_gpu_amd.batch import make_problems, problem_deck
    from dgfs1d_gpu_amd.dictionary import Dictionary
    a, b = make_problems(7), make_problems(7)
    assert a == b and len(a) == 64
    for p in a:
        assert p['Ne'] in (2, 4, 8) and p['K'] in (3, 5, 7) and p['N'] in (16, 24, 32)
        assert 0.1 <= abs(p['uw']) <= 1.0 and 0.8 <= p['Tl'] <= 1.5 and 0.8 <= p['Tr'] <= 1.5
        if p['kind'] == 'couette':
            assert p['Tl'] == p['Tr']
    d = Dictionary(problem_deck(**{k: v for k, v in a[0].items()}))
    assert d.lookupint('velocity-mesh', 'Nrho') == a[0]['N'] // 2
    assert {p['kind'] for p in a} == {'couette', 'fourier'}


def test_rankine_hugoniot_limits():
    from dgfs1d_gpu_amd.closure_train import rankine_hugoniot
    r, u, T = rankine_hugoniot(1.0)
    assert abs(r - 1) < 1e-14 and abs(u - 1) < 1e-14 and abs(T - 1) < 1e-14
    r, u, T = rankine_hugoniot(1e4)                    # strong-shock limit (g+1)/(g-1) = 4
    assert abs(r - 4.0) < 1e-6
    # mass flux and momentum conservation across the shock for M = 2
    g = 5.0 / 3.0
    r, u, T = rankine_hugoniot(2.0)
    p = r * T                                          # p2/p1
    assert abs(r * u - 1.0) < 1e-14
    m2 = 4.0
    assert abs((1 + g * m2) - (p + g * m2 * r * u * u)) < 1e-12


def _dg_file(path, fun, K, Ne, Nv=3, L=1.0):
    """Nodal DG coefficients (GLL nodes) of fun(x, v) on Ne equal elements of [0, L]."""
    from dgfs1d_gpu_amd.quadratures import zwglj
    r, _ = zwglj(K, 0.0, 0.0)
    h = L / Ne
    x = (np.arange(Ne)[None, :] + 0.5 * (r[:, None] + 1.0)) * h          # (K, Ne)
    coeff = np.stack([fun(x, v) for v in range(Nv)], axis=-1)           # (K, Ne, Nv)
    np.savez(path, coeff=coeff.ravel(), time=0.0, K=K, Ne=Ne, Nv=Nv)


def test_error_tool_polynomial_is_exact(tmp_path):
    """A polynomial of degree < K is represented exactly on both meshes."""
    from dgfs1d_gpu_amd.error import compute_error
    fun = lambda x, v: 1.0 + (v + 1) * x - 0.7 * x ** 2
    _dg_file(tmp_path / 'c.npz', fun, 3, 4)
    _dg_file(tmp_path / 'f.npz', fun, 3, 8)
    assert compute_error(tmp_path / 'c.npz', tmp_path / 'f.npz') < 1e-14


def test_error_tool_converges_and_measures_l1(tmp_path):
    """For a smooth function the coarse/fine difference falls like h^K; for
    fine data shifted by a constant it equals that constant (mean L1 over
    the reference element's weights, which sum to 2, times 0.5)."""
    from dgfs1d_gpu_amd.error import compute_error
    fun = lambda x, v: np.sin(2 * np.pi * x + v)
    errs = []
    for Ne in (4, 8):
        _dg_file(tmp_path / 'c.npz', fun, 3, Ne)
        _dg_file(tmp_path / 'f.npz', fun, 3, 2 * Ne)
        errs.append(compute_error(tmp_path / 'c.npz', tmp_path / 'f.npz'))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.5, (errs, rate)
    _dg_file(tmp_path / 'c.npz', fun, 3, 4)
    _dg_file(tmp_path / 'f.npz', lambda x, v: fun(x, v) + 0.25, 3, 8)
    e = compute_error(tmp_path / 'c.npz', tmp_path / 'f.npz')
    assert abs(e - 0.25) < 1e-3


def test_error_tool_rejects_mismatched_meshes(tmp_path):
    from dgfs1d_gpu_amd.error import compute_error
    fun = lambda x, v: x
    _dg_file(tmp_path / 'c.npz', fun, 3, 4)
    _dg_file(tmp_path / 'f.npz', fun, 3, 6)
    with pytest.raises(ValueError):
        compute_error(tmp_path / 'c.npz', tmp_path / 'f.npz')


def test_bi_mesh_and_initial_state_match_oracle():
    """dgfsBi1D host set-up (velocity mesh, spherical rules, Maxwellian IC and
    diffuse-wall f0 per species) against oracle/bideck.py."""
    from dgfs1d_gpu_amd.bi.initcond import DGFSMaxwellianInitConditionBi
    from dgfs1d_gpu_amd.bi.velocitymesh import DGFSVelocityMeshBi
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from oracle.bideck import BiDeck
    from tests.test_gpu_bi_solver import _ini
    text = _ini({})
    deck = BiDeck(text)
    cfg = Dictionary(text)
    vm = DGFSVelocityMeshBi(cfg)
    ov = deck.vm
    assert vm.L() == ov.L and vm.masses() == ov.masses
    assert np.allclose(vm.qz(), ov.qz, rtol=1e-15) and np.allclose(vm.qw(), ov.qw, rtol=1e-15)
    assert np.array_equal(vm.sz(), ov.sz) and vm.sw() == ov.sw
    ic = DGFSMaxwellianInitConditionBi(cfg, vm, 'soln-ics')
    for p in range(2):
        fb = deck.basis.fwd
        u0 = np.einsum('kq,qen->ken', fb, np.broadcast_to(ic.get_init_vals()[p],
                                                            (fb.shape[1], deck.Ne, vm.vsize())))
        assert np.allclose(u0, deck.u0[p], rtol=1e-14, atol=0)
    wall = DGFSMaxwellianInitConditionBi(cfg, vm, 'soln-bcs-xlo', wall=True)
    for p in range(2):
        f0 = deck.op.tr[p].wl.f0
        assert np.allclose(wall.get_init_vals()[p], f0, rtol=1e-14, atol=0)


def test_bi_registries():
    from dgfs1d_gpu_amd.bi.bc import DGFSBCBi
    from dgfs1d_gpu_amd.bi.initcond import DGFSInitConditionBi
    from dgfs1d_gpu_amd.bi.integrator import DGFSIntegratorBi
    from dgfs1d_gpu_amd.bi.scattering import DGFSScatteringModelBi
    for k in ('euler', 'ssp-rk2', 'ssp-rk3', 'lse-rk45'):
        subclass_where(DGFSIntegratorBi, intg_kind=k)
    for k in ('dgfs-wall-diffuse', 'dgfs-inlet', 'dgfs-periodic', 'dgfs-wall-expr-diffuse'):
        subclass_where(DGFSBCBi, type=k)
    for k in ('maxwellian', 'maxwellian-expr'):
        subclass_where(DGFSInitConditionBi, model=k)
    subclass_where(DGFSScatteringModelBi, scattering_model='vss')


def test_bi_moments_of_a_maxwellian_state():
    """The mixture moment writer returns the deck's number densities,
    temperature and zero velocity / heat flux for the Maxwellian initial
    state (well resolved: Nv = 24, dev = 5)."""
    import torch
    from dgfs1d_gpu_amd.bi.initcond import DGFSMaxwellianInitConditionBi
    from dgfs1d_gpu_amd.bi.moments import bi_moments
    from dgfs1d_gpu_amd.bi.velocitymesh import DGFSVelocityMeshBi
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from tests.test_gpu_bi_solver import _ini
    cfg = Dictionary(_ini({('velocity-mesh', 'Nv'): 24}))
    vm = DGFSVelocityMeshBi(cfg)
    ic = DGFSMaxwellianInitConditionBi(cfg, vm, 'soln-ics')
    K, Ne, Nqr = 2, 3, 2
    im = np.array([[1.0, 1.0], [0.0, 0.0]])          # reconstruction reads mode 0
    coeffs = []
    for p in range(2):
        u = np.zeros((K, Ne, vm.vsize()))
        u[0] = ic.get_init_vals()[p]
        coeffs.append(torch.tensor(u.ravel()))
    out = bi_moments(coeffs, im, vm)
    nden = [1.6802782428845444e+21, 8.008802839916987e+20]
    for p in range(2):
        o = 7 * p
        assert np.allclose(out[..., o + 6], nden[p], rtol=1e-8)          # nden
        assert np.allclose(out[..., o + 3], 273.0, rtol=1e-6)            # T
        assert np.abs(out[..., o + 1]).max() < 1e-9                      # U:x
        assert np.abs(out[..., o + 4]).max() < 1e-9 * 0.5 * vm.rho0() * vm.u0() ** 3  # Q:x
        # the reference writes p_p = m_p R0/molarMass0 rho_p T_p (bi/moments.py:106-109),
        # which is m_p^2 times n_p k T (k = R0/NA); the writer keeps that column
        m = vm.masses()[p]
        assert np.allclose(out[..., o + 5], m * m * nden[p] * 8.3144598 / 6.0221409e+23 * 273.0,
                           rtol=1e-6)


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with a message (one-GPU build)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert 'one GPU' in p.stderr


def test_bi_maxwellian_expr_with_constants_equals_maxwellian():
    """The mixture's maxwellian-expr initial condition with constant
    expressions gives the maxwellian one at every solution point."""
    from dgfs1d_gpu_amd.bi.initcond import (DGFSMaxwellianExprInitConditionBi,
                                            DGFSMaxwellianInitConditionBi)
    from dgfs1d_gpu_amd.bi.velocitymesh import DGFSVelocityMeshBi
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from tests.test_gpu_bi_solver import _ini
    cfg = Dictionary(_ini({('soln-ics', 'ux'): 20.0, ('velocity-mesh', 'Nv'): 24}))
    vm = DGFSVelocityMeshBi(cfg)
    a = DGFSMaxwellianInitConditionBi(cfg, vm, 'soln-ics')
    cfg2 = Dictionary(_ini({('soln-ics', 'type'): 'maxwellian-expr', ('soln-ics', 'ux'): 20.0,
                            ('velocity-mesh', 'Nv'): 24}))
    b = DGFSMaxwellianExprInitConditionBi(cfg2, vm, 'soln-ics')
    xsol = np.linspace(0, 1, 6).reshape(3, 2)
    for p in range(2):
        u1 = np.zeros((3, 2, vm.vsize()))
        u2 = np.zeros_like(u1)
        a.apply_init_vals(p, u1, 3, 2, xsol)
        b.apply_init_vals(p, u2, 3, 2, xsol)
        assert np.allclose(u1, u2, rtol=1e-13, atol=0)


def test_maxwell_gll_parameters():
    """'maxwell-gll' (std/scattering.py:707-743): gamma 0, omega 1, prefactor
    H0 u0^2 / n0 / eps (here 1/(4 pi)), BGK and ES-BGK prefactors equal."""
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.scattering import DGFSMaxwellGLLScatteringModelStd
    from dgfs1d_gpu_amd.std.velocitymesh import DGFSVelocityMeshStd
    from tests.test_gpu_solver import BKW
    R0, NA = 8.3144598, 6.0221409e+23
    u0 = np.sqrt(2 * R0 / 39.96e-3 * 273.0)
    n0 = 2.3e-4 / 39.96e-3 * NA
    cfg = Dictionary(BKW % (4 * np.pi * u0 ** 2 / n0))
    m = DGFSMaxwellGLLScatteringModelStd.__new__(DGFSMaxwellGLLScatteringModelStd)
    m.cfg, m.vm = cfg, DGFSVelocityMeshStd(cfg, device='cpu')
    m.load_parameters()
    assert m._gamma == 0.0 and m.omega == 1.0
    assert m.prefac == pytest.approx(1 / (4 * np.pi), rel=1e-12)
    assert m.prefacBGK == m.prefacESBGK == m.prefac
