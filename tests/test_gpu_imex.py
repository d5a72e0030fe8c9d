"""GPU parity of the IMEX path (dgfsAstd1D) against the FP64 oracle.

The 14-moment Shakov penalisation kernels (moments, target, collide) and
full IMEX steps -- ARS tableaux and linear-multistep schemes with their
ARS-443 starting values -- on a small Couette deck (N = 8).  Tolerance: FP64
relative 1e-11 of max|value|.
"""
import numpy as np
import pytest

from oracle import imex as OI
from oracle.deck import Deck

pytestmark = pytest.mark.gpu

EXTRA = """
[penalized-scattering-model]
type = shakov-direct-gll
"""


def _text(scheme, model='vhs', penal='shakov-direct-gll'):
    from tests.test_gpu_solver import _ini
    return _ini({('time-integrator', 'scheme'): scheme, ('scattering-model', 'type'): model,
                 ('scattering-model', 'muRef'): 2.117e-5}) + EXTRA.replace('shakov-direct-gll',
                                                                          penal)


def _oracle(deck, Pr=2.0 / 3.0, model='shakov'):
    b, op = deck.basis, deck.op
    K, Ne = deck.K, deck.Ne
    n = deck.vm.vsize
    from oracle import relax as OR
    cp = deck.cp
    prefac = OR.shakov_prefactor(float(cp.get('mesh', 'H0')), deck.vm.u0, deck.vm.T0,
                                 deck.vm.n0, float(cp.get('scattering-model', 'muRef')),
                                 float(cp.get('scattering-model', 'Tref')),
                                 float(cp.get('scattering-model', 'omega')))
    if model == 'esbgk':
        prefac *= Pr                          # t0 Pr p0 / mu(T0)
    pm = OI.Penal(deck.vm.cv, deck.vm.cw, prefac, float(cp.get('scattering-model', 'omega')), Pr,
                  model=model)

    def explicit(t, f):
        u = f.reshape(K, Ne, n)
        ux = op.transport(t, u)
        return (b.invmass @ ux.reshape(K, -1)).reshape(K * Ne, n)

    def boltz(t, f):
        from oracle import collision as OC
        out = np.zeros_like(f)
        for s in range(K * Ne):
            if deck.model == 'vhs':
                out[s] = OC.q_vhs(deck.pre, f[s], f[s], deck.prefac)
            else:
                out[s] = OC.q_vhs_gll(deck.pre, f[s], deck.prefac)
        return out
    return pm, explicit, boltz


def _solver(cuda, text):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.astd.abstd import DGFSAstd1D
    return DGFSAstd1D(Dictionary(text), device=cuda, writers=False)


def test_penalisation_kernels(cuda):
    import torch
    text = _text('imex-lm-bdf1')
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, _, _ = _oracle(deck)
    rng = np.random.default_rng(2)
    f = deck.u0.reshape(s.K * s.Ne, -1) * (1.0 + 0.1 * rng.uniform(-1, 1, (s.K * s.Ne, 1)))
    f = f * (1.0 + 0.05 * rng.uniform(-1, 1, f.shape))
    fd = torch.tensor(f.ravel(), dtype=torch.float64, device=cuda)
    U = s.psm.new_moments(s.K * s.Ne)
    s.psm.moment(0.0, fd, U)
    Ur = pm.moment(f)
    eU = np.abs(U.cpu().numpy().reshape(Ur.shape) - Ur).max() / np.abs(Ur).max()
    M = torch.zeros_like(fd)
    s.psm.constructMaxwellian(0.0, U, M)
    Mr = pm.target(Ur)
    eM = np.abs(M.cpu().numpy().reshape(Mr.shape) - Mr).max() / np.abs(Mr).max()
    P = torch.zeros_like(fd)
    s.psm.collide(0.0, U, M, fd, P)
    Pr_ = pm.penal(Ur, f)
    eP = np.abs(P.cpu().numpy().reshape(Pr_.shape) - Pr_).max() / np.abs(Pr_).max()
    print('penalisation', eU, eM, eP)
    assert max(eU, eM, eP) < 1e-12


def test_esbgk_penalisation_kernels(cuda):
    """ES-BGK penalisation kernels on strongly anisotropic slices: the target
    (dgfs_esbgk_target) against oracle.imex.es_target, and the LM moment
    update with the stresses relaxing at nu/Pr against oracle.imex.lm_moments
    (srate = 1/Pr; the Shakov rate srate = 1 misses by far more)."""
    import torch
    text = _text('imex-lm-bdf2', penal='esbgk-direct-gll')
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, _, _ = _oracle(deck, model='esbgk')
    cv, ns = deck.vm.cv, s.K * s.Ne
    rng = np.random.default_rng(6)
    f = np.empty((ns, cv.shape[1]))
    for i in range(ns):
        Tx, Ty, Tz = rng.uniform(0.6, 1.4, 3)
        C = cv - rng.uniform(-0.3, 0.3, 3)[:, None]
        f[i] = rng.uniform(0.5, 1.5) / (np.pi ** 1.5 * np.sqrt(Tx * Ty * Tz)) * np.exp(
            -C[0] ** 2 / Tx - C[1] ** 2 / Ty - C[2] ** 2 / Tz - 0.2 * C[0] * C[1])
    fd = torch.tensor(f.ravel(), dtype=torch.float64, device=cuda)
    U = s.psm.new_moments(ns)
    s.psm.moment(0.0, fd, U)
    Ur = pm.moment(f)
    M = torch.zeros_like(fd)
    s.psm.constructMaxwellian(0.0, U, M)
    Mr = pm.target(Ur)
    eM = np.abs(M.cpu().numpy().reshape(Mr.shape) - Mr).max() / np.abs(Mr).max()
    A, B, Cc = OI.LM['imex-lm-bdf2']
    dt = 0.02
    U1 = Ur * (1.0 + 0.01 * rng.uniform(-1, 1, Ur.shape))
    LU = [0.1 * Ur * rng.uniform(-1, 1, Ur.shape) for _ in range(2)]
    LM = [0.1 * Ur * rng.uniform(-1, 1, Ur.shape) for _ in range(2)]
    T = lambda x: torch.tensor(np.ascontiguousarray(x).ravel(), dtype=torch.float64, device=cuda)
    Ud = [T(Ur), T(U1), s.psm.new_moments(ns)]
    s.psm.updateMomentLM(dt, A, B, Cc, [T(x) for x in LU], [T(x) for x in LM], Ud, 2)
    torch.cuda.synchronize()
    got = Ud[2].cpu().numpy().reshape(Ur.shape)
    ref = OI.lm_moments(dt, A, B, Cc, LU, LM, [Ur, U1], 2, pm.prefac, pm.omega, pm.Pr, pm.srate)
    slip = OI.lm_moments(dt, A, B, Cc, LU, LM, [Ur, U1], 2, pm.prefac, pm.omega, pm.Pr, 1.0)
    eU = np.abs(got - ref).max() / np.abs(ref).max()
    eS = np.abs(slip - ref)[:, 5:11].max() / np.abs(ref).max()
    print('esbgk kernels: target %.2e moments %.2e (srate slip %.2e)' % (eM, eU, eS), flush=True)
    assert eM < 1e-12 and eU < 1e-12 and eS > 1e3 * max(eU, 1e-16)


@pytest.mark.parametrize('scheme', ['ars-111', 'ars-222', 'ars-443'])
def test_ars_steps(cuda, scheme):
    text = _text(scheme)
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, explicit, _ = _oracle(deck)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    t = 0.0
    for k in range(3):
        f = OI.ars_step(scheme, pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    print(scheme, err)
    assert err < 1e-11


@pytest.mark.parametrize('scheme', ['ars-222', 'imex-lm-bdf2'])
def test_bgk_penalisation_steps(cuda, scheme):
    """[penalized-scattering-model] type = bgk-direct-gll
    (astd/scattering.py:44-286): the Maxwellian target, i.e. the oracle's
    Shakov penalisation at Pr = 1 (its heat-flux term vanishes)."""
    text = _text(scheme, penal='bgk-direct-gll')
    deck = Deck(text)
    s = _solver(cuda, text)
    assert s.psm._Pr == 1.0
    pm, explicit, boltz = _oracle(deck, Pr=1.0)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    it = OI.LMIntegrator(scheme, pm, explicit, boltz, 0.0, deck.dt, f.copy()) \
        if scheme.startswith('imex-lm') else None
    t = 0.0
    for k in range(3):
        f = it.step(t, deck.dt, k, f) if it else OI.ars_step(scheme, pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    print('bgk penalisation', scheme, err)
    assert err < 1e-11


@pytest.mark.parametrize('scheme', ['ars-443', 'imex-lm-bdf2'])
def test_esbgk_penalisation_steps(cuda, scheme):
    """[penalized-scattering-model] type = esbgk-direct-gll
    (astd/scattering.py:287-354): the ES-BGK target from the moments
    (dgfs_esbgk_target) and the stresses relaxing at nu/Pr (srate = 1/Pr in
    the moment updates), against the oracle's esbgk penalisation."""
    text = _text(scheme, penal='esbgk-direct-gll')
    deck = Deck(text)
    s = _solver(cuda, text)
    assert abs(s.psm.srate * 2.0 / 3.0 - 1.0) < 1e-15
    pm, explicit, boltz = _oracle(deck, model='esbgk')
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    it = OI.LMIntegrator(scheme, pm, explicit, boltz, 0.0, deck.dt, f.copy()) \
        if scheme.startswith('imex-lm') else None
    t = 0.0
    for k in range(3):
        f = it.step(t, deck.dt, k, f) if it else OI.ars_step(scheme, pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    print('esbgk penalisation', scheme, err)
    assert err < 1e-11


@pytest.mark.parametrize('scheme,model', [('imex-lm-bdf1', 'vhs'), ('imex-lm-bdf2', 'vhs'),
                                          ('imex-lm-bdf3', 'vhs-gll'), ('imex-lm-cn2', 'vhs')])
def test_lm_steps(cuda, scheme, model):
    text = _text(scheme, model)
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, explicit, boltz = _oracle(deck)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    it = OI.LMIntegrator(scheme, pm, explicit, boltz, 0.0, deck.dt, f.copy())
    t = 0.0
    for k in range(4):
        f = it.step(t, deck.dt, k, f)
        s.step()
        t += deck.dt
        err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
        print(scheme, model, k, err)
        assert err < 1e-11


@pytest.mark.parametrize('scheme', ['sspl-222', 'sspl-443'])
def test_sspl_steps(cuda, scheme):
    text = _text(scheme)
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, explicit, _ = _oracle(deck)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    t = 0.0
    for k in range(3):
        f = OI.sspl_step(scheme, pm, explicit, t, deck.dt, f)
        s.step()
        t += deck.dt
    err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
    moved = np.abs(f - deck.u0.reshape(f.shape)).max() / np.abs(f).max()
    print(scheme, err, moved)
    assert err < 1e-11 and moved > 1e-6


@pytest.mark.parametrize('scheme', ['bdf-111', 'bdf-222', 'bdf-333'])
def test_bdf_steps(cuda, scheme):
    text = _text(scheme)
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, explicit, _ = _oracle(deck)
    it = OI.BDFIntegrator(scheme, pm, explicit)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    t = 0.0
    for k in range(4):
        f = it.step(t, deck.dt, k, f)
        s.step()
        t += deck.dt
        err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
        moved = np.abs(f - deck.u0.reshape(f.shape)).max() / np.abs(f).max()
        print(scheme, k, err, moved)
        assert err < 1e-11 and moved > 1e-6


def test_boltz_ars111_steps(cuda):
    text = _text('boltz-ars-111')
    deck = Deck(text)
    s = _solver(cuda, text)
    pm, explicit, boltz = _oracle(deck)
    it = OI.BDFIntegrator('bdf-111', pm, explicit, boltz)
    f = deck.u0.reshape(s.K * s.Ne, -1).copy()
    t = 0.0
    for k in range(3):
        f = it.step(t, deck.dt, k, f)
        s.step()
        t += deck.dt
        err = np.abs(s.u.cpu().numpy().reshape(f.shape) - f).max() / np.abs(f).max()
        moved = np.abs(f - deck.u0.reshape(f.shape)).max() / np.abs(f).max()
        print('boltz-ars-111', k, err, moved)
        assert err < 1e-11 and moved > 1e-6
