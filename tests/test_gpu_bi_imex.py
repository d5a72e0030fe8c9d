"""The binary mixture on the IMEX path (dgfs1d_gpu_amd.bi.abi: the
single-species IMEX integrators with each species penalised by its own Shakov
model) against the FP64 oracle (oracle.bideck.BiDeck.imex_parts with
oracle.imex.LMIntegrator / ars_step), and config 5's grid at full size.

* N = 8 decks (Mpre = 32), FP64: IMEX-BDF2 on a nodal and on a modal basis
  (with the mixture conservation projection), and ARS-222; three steps from a
  perturbed state, 1e-11 relative to max|f|.
* The N = 48 deck (configs/bi_fourier_n48.ini: Ne = 8, modal K = 5, M = 24,
  N_rho = 24, Mpre = 192, FP32, IMEX-BDF2): two steps stay finite, each
  species' number of particles changes by less than 1e-5 (diffuse walls and
  the projection keep it; the FP32 update rounds), and the step moves the
  state.
* The same physical deck at N = 24 (N_rho = 8): FP32 against FP64 after three
  IMEX-BDF2 steps, 1e-5 relative to max|f|.
"""
import os

import numpy as np
import pytest

from oracle import imex as OI
from oracle.bideck import BiDeck

from test_gpu_bi_solver import _ini

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IMEX = {('penalized-scattering-model', 'type'): 'shakov-direct-gll',
        ('penalized-scattering-model', 'Pr'): 2.0 / 3.0,
        ('time-integrator', 'dt'): 0.002}


def _with_penalty(over):
    from configparser import ConfigParser
    import io
    text = _ini({k: v for k, v in over.items() if k[0] != 'penalized-scattering-model'})
    cp = ConfigParser(inline_comment_prefixes=[';'])
    cp.optionxform = str
    cp.read_string(text)
    cp.add_section('penalized-scattering-model')
    for (s, k), v in over.items():
        if s == 'penalized-scattering-model':
            cp.set(s, k, str(v))
    buf = io.StringIO()
    cp.write(buf)
    return buf.getvalue()


CASES = {
    'nodal-bdf2': {('time-integrator', 'scheme'): 'imex-lm-bdf2'},
    'modal-bdf2-projection': {('time-integrator', 'scheme'): 'imex-lm-bdf2',
                              ('basis', 'kind'): 'modal-orthonormal-legendre',
                              ('basis', 'Nq'): 4, ('scattering-model', 'projection'):
                                  'conservative'},
    'nodal-ars222': {('time-integrator', 'scheme'): 'ars-222'},
}


@pytest.mark.parametrize('case', sorted(CASES))
def test_bi_imex_against_oracle(cuda, case):
    import torch
    from dgfs1d_gpu_amd.bi.abi import DGFSBiAstd1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    over = dict(IMEX)
    over.update(CASES[case])
    text = _with_penalty(over)
    deck = BiDeck(text)
    s = DGFSBiAstd1D(Dictionary(text), device=cuda, writers=False)
    rng = np.random.default_rng(7)
    F0 = np.concatenate([(u * (1.0 + 0.1 * rng.uniform(-1, 1, u.shape))).reshape(-1, u.shape[-1])
                         for u in deck.u0])
    s.u.copy_(torch.tensor(F0.ravel(), dtype=s.tdtype, device=cuda))
    s.odes = s._make_integrator(s.cfg, (s.K, s.Ne, s.Nv))      # history from this state
    explicit, boltz, pm, to_points, to_coeffs = deck.imex_parts()
    kind, dt = s.cfg.lookup('time-integrator', 'scheme'), s.dt
    W0 = to_points(F0)
    if kind.startswith('imex-lm'):
        it = OI.LMIntegrator(kind, pm, explicit, boltz, 0.0, dt, W0.copy())
        step = lambda k, w: it.step(k * dt, dt, k, w)
    else:
        # ARS: explicit part transport - P (the Boltzmann part is replaced by
        # the penalisation, as in the single-species driver)
        step = lambda k, w: OI.ars_step(kind, pm, explicit, k * dt, dt, w)
    w = W0.copy()
    for k in range(3):
        print(case, 'step', k, flush=True)
        s.step()
        w = step(k, w)
        f = to_coeffs(w)
        got = s.u.cpu().numpy().reshape(f.shape)
        err = float(np.abs(got - f).max() / np.abs(f).max())
        print(case, 'step', k, 'rel err %.3e' % err, flush=True)
        assert err < 1e-11, (case, k, err)
    assert np.abs(f - F0).max() > 1e-4 * np.abs(F0).max()          # the state moved


def _n48_deck(over=None):
    from dgfs1d_gpu_amd.dictionary import Dictionary
    text = open(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'bi_fourier_n48.ini')).read()
    cfg = Dictionary(text)
    for (sect, key), v in (over or {}).items():
        cfg.set(sect, key, str(v))
    return cfg


def _particles(s):
    """Each species' number of particles up to a common constant: the sum of
    the modal-orthonormal mode-0 coefficients (the element means; the
    elements are equal) over elements and the velocity grid."""
    K, Ne, n = s.K, s.Ne, s.Nv
    return [float(u.view(K, Ne, n)[0].double().sum()) for u in s.us]


def test_bi_imex_config5_grid(cuda):
    import torch
    from dgfs1d_gpu_amd.bi.abi import DGFSBiAstd1D
    cfg = _n48_deck()
    s = DGFSBiAstd1D(cfg, device=cuda, writers=False)
    assert s.Nv == 48 ** 3 and s.K == 5 and s.Ne == 8 and s.vm.M() == 24 and s.vm.Nrho() == 24
    u0 = s.u.clone()
    n0 = _particles(s)
    for k in range(2):
        print('config-5 grid IMEX-BDF2 step', k, flush=True)
        s.step()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(s.u).all())
    n1 = _particles(s)
    drift = [abs(a - b) / abs(b) for a, b in zip(n1, n0)]
    moved = float((s.u - u0).abs().max() / u0.abs().max())
    print('particle drift per species', drift, 'state change', moved)
    assert max(drift) < 1e-5
    assert moved > 1e-6


def test_bi_imex_fp32_against_fp64(cuda):
    from dgfs1d_gpu_amd.bi.abi import DGFSBiAstd1D
    runs = {}
    for prec in ('double', 'single'):
        cfg = _n48_deck({('velocity-mesh', 'Nv'): 24, ('velocity-mesh', 'Nrho'): 8,
                         ('config', 'precision'): prec})
        s = DGFSBiAstd1D(cfg, device=cuda, writers=False)
        for k in range(3):
            print(prec, 'step', k, flush=True)
            s.step()
        runs[prec] = s.u.double().cpu().numpy()
    err = float(np.abs(runs['single'] - runs['double']).max() / np.abs(runs['double']).max())
    print('IMEX-BDF2 N=24 FP32 vs FP64: %.3e' % err)
    assert err < 1e-5
