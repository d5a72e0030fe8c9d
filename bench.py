"""Benchmark: phase-space updates/s of the production Couette case on one MI355X.

Workload (dgfs1d_gpu_amd/configs/couette_prod.ini): single-species VHS
Couette, Ne=4, K=3 (nodal-sem-gll), N=32^3, M=12, N_rho=16, SSP-RK2,
dt=5e-4, FP32 spectral arithmetic, FP64 wall reductions and FP64
conservation projection of the collision output.  One step = one
SSP-RK2 time step (2 RHS evaluations, each with 12 batched collision
integrals).  Inputs are resident in HBM before timing starts.

    python bench.py [--gpus 1] [--steps 52] [--warmup 3] [--no-cpu-baseline]

Prints one JSON line.  ``accuracy`` compares the moments of the final FP32
state with an FP64 run of the same steps (after the timed region).
``roofline`` is for the dominant kernel (the gain
kernel of the collision operator), timed with HIP events recorded around it
on its launch stream over the timed region.  ``cpu_baseline`` is the FP64
NumPy/SciPy oracle on the host cores (the job's CPU share), timed over 3 full
SSP-RK2 steps of the same deck after one warm-up step (median step time).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3      # MI355X published dense FP32
PEAK_HBM_GBS = 8000.0

# HBM bytes per gain dispatch (k_gain_pair at N=32 FP32), from two rocprofv3 counter passes over this
# script (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what
# the same counter reads for a device-to-device copy of 512 MiB (--copy-calibration):
# FETCH x 2.000, WRITE x 1.000.  Keyed by workload; re-measured whenever the
# kernel changes, null for workloads not measured.
GAIN_TRAFFIC_KIB = {
    'couette Ne=4 K=3 N=32^3 M=12 Nrho=16 ssp-rk2 dt=0.0005': (30957.8, 2.000, 52224.0, 1.000),
    'couette Ne=4 K=3 N=32^3 M=12 Nrho=16 ssp-rk2 dt=0.0005 reference-state R=40':
        (629353.4, 2.000, 70656.0, 1.000),
}


def fft_flops(n):
    return 5.0 * n * math.log2(n)


def host_cpu():
    """CPU model, logical CPUs of the host (nproc) and the CPUs this process may
    run on (its affinity mask: the job's share of the host)."""
    model = None
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    return model, os.cpu_count() or 1, share


def cpu_baseline(deck_text, steps=3):
    """The FP64 oracle (the reference's algorithm restated in NumPy/SciPy, FFTs
    through scipy.fft with one worker per CPU of this job) advancing the same
    deck by full time steps: one untimed step, then ``steps`` timed steps;
    the median step time gives the rate.  Threads: the CPUs in this process'
    affinity mask, capped by OMP_NUM_THREADS when that is set (the job's
    share of a shared host)."""
    import numpy as np
    from oracle import collision as OC
    from oracle.deck import Deck
    model, nproc, share = host_cpu()
    cores = share
    omp = os.environ.get('OMP_NUM_THREADS')
    if omp and omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    OC.set_fft_workers(cores)
    try:
        d = Deck(deck_text)
        u = d.u0
        t = d.t0
        u = d.step(t, u)                              # untimed warm-up step
        t += d.dt
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            u = d.step(t, u)
            times.append(time.perf_counter() - t0)
            t += d.dt
    finally:
        OC.set_fft_workers(None)
    step_s = float(np.median(times))
    pts = d.Ne * d.K * d.vm.vsize
    return dict(value=pts / step_s, unit='phase-space point-steps/s', cores=cores,
                kind='port', cpu_model=model, nproc=nproc, affinity_cpus=share,
                step_s=times,
                sample='%d full %s steps of the same deck by the FP64 oracle after one '
                       'warm-up step (%d collision integrals per step, scipy.fft with %d '
                       'workers), median %.2f s per step'
                       % (steps, d.scheme, 2 * d.K * d.Ne, cores, step_s))


MOMENT_NAMES = ('rho', 'ux', 'uy', 'T', 'qx', 'qy', 'Pxx', 'Pyy', 'Pxy')


def fp64_accuracy(text, nsteps, s32, dev):
    """Moments of the FP32 run against the FP64 path after the same steps, and
    the invariant moments of the collision output of both final states."""
    import numpy as np
    import torch
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    from dgfs1d_gpu_amd.transport import Moments
    cfg = Dictionary(text)
    cfg.set('config', 'precision', 'double')
    s64 = DGFSStd1D(cfg, device=dev, writers=False)
    for _ in range(nsteps):
        s64.step()
    mom = Moments(s64.basis.interpMat, s64.K, s64.Ne, s64.vm, dev)
    m64 = mom(s64.state_fp64()).reshape(-1, 9)
    m32 = mom(s32.state_fp64()).reshape(-1, 9)      # FP32 state (hi + lo), FP64 moments
    err = {}
    speed = float(np.abs(m64[:, 1:3]).max())
    for i, name in enumerate(MOMENT_NAMES):
        # u_x stays near zero in this flow: relative to the largest flow speed
        scale = speed if name == 'ux' else float(np.abs(m64[:, i]).max())
        err[name] = {'max_abs': float(np.abs(m32[:, i] - m64[:, i]).max()), 'field_max': scale}
        err[name]['rel'] = err[name]['max_abs'] / scale if scale > 0 else None
    cv = s64.vm.cv()
    phi = np.vstack([np.ones(cv.shape[1]), cv, (cv * cv).sum(0)])
    cons = {}
    for tag, s in (('f64', s64), ('f32', s32)):
        Q = torch.zeros_like(s.u)
        s.sm.collide(s.u, Q)
        q = Q.double().cpu().numpy().reshape(s.K * s.Ne, -1)
        cons[tag] = float((np.abs(phi @ q.T) / (np.abs(phi) @ np.abs(q).T)).max())
    return {'steps': nsteps, 'moments_fp32_vs_fp64': err,
            'collision_invariant_moments_rel': cons}


def direct_form(text, steps, warmup, dev, accuracy=True):
    """The same deck with the direct FP32 collision (reference-refresh = 0):
    ms per step, phase-space updates/s and, unless accuracy is False, the
    moments against the FP64 path after the same steps."""
    import torch
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    cfg = Dictionary(text)
    cfg.set('scattering-model', 'reference-refresh', '0')
    s = DGFSStd1D(cfg, device=dev, writers=False)
    for _ in range(warmup):
        s.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {'collision_form': 'direct (FP32 Q(f, f))', 'ms_per_step': wall / steps * 1e3,
           'value': s.Ne * s.K * s.vm.vsize() * steps / wall}
    if accuracy:
        acc = fp64_accuracy(text, warmup + steps, s, dev)
        out['moments_fp32_vs_fp64'] = {k: v['rel'] for k, v in acc['moments_fp32_vs_fp64'].items()}
    del s
    return out


def batched_workload(seed, count, steps, warmup, dev):
    """The seeded uneven batch (64 Couette/Fourier problems, Ne in {2,4,8},
    K in {3,5,7}, N in {16,24,32}, M=12, N_rho=N/2, FP32): autotune the
    collision plans, capture one step in a HIP graph, time graph replays."""
    import torch
    from dgfs1d_gpu_amd.batch import BatchedStd1D, make_problems
    specs = make_problems(seed, count)
    t0 = time.perf_counter()
    b = BatchedStd1D(specs, dev, precision='float')
    setup = time.perf_counter() - t0
    tune = b.autotune()
    b.capture(warmup=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {'problems': count, 'seed': seed, 'points': b.points,
            'groups': [{'N': g['N'], 'problems': len(g['solvers']), 'collisions_per_rhs':
                        g['ntask'], 'schedule': dict(zip(('qgroups', 'split_mult', 'xcd_remap'),
                                                         tune[g['N']][0]))}
                       for g in b.groups],
            'transport': 'batched (3 launches per RHS for the 64 problems)',
            'steps': steps, 'ms_per_step': wall / steps * 1e3,
            'value': b.points * steps / wall, 'unit': 'phase-space point-steps/s',
            'graph': True, 'setup_s': setup,
            'autotune_ms': {str(k): {'qg%d_split%d_remap%d' % sc: ms for sc, ms in v[1].items()}
                            for k, v in tune.items()}}


def closure_forward(nodes, G, reps, dev):
    """Learned-closure network forward over `nodes` velocity slices of G^3
    (random weights and Maxwellian-mixture inputs), HIP-event timed."""
    import numpy as np
    import torch
    from dgfs1d_gpu_amd.learned import ClosureEngine, ClosureNet
    net = ClosureNet(dev, seed=0)
    eng = ClosureEngine(net, G, nodes)
    c0 = np.linspace(-5 + 5 / G, 5 - 5 / G, G)
    Z, Y, X = np.meshgrid(c0, c0, c0, indexing='ij')
    rng = np.random.default_rng(0)
    f = np.stack([np.exp(-((X - a) ** 2 + Y ** 2 + Z ** 2) / T) / (np.pi * T) ** 1.5
                  for a, T in zip(rng.uniform(-1, 1, nodes), rng.uniform(0.7, 1.5, nodes))])
    u = torch.tensor(f.reshape(-1), dtype=torch.float32, device=dev)
    src = torch.arange(nodes, dtype=torch.int32, device=dev)
    sc = torch.ones(nodes, dtype=torch.float32, device=dev)
    Q = torch.zeros_like(u)
    osc = torch.ones(nodes, dtype=torch.float32, device=dev)
    flop = 2.0 * G ** 3 * 27 * (32 + 32 * 32 + 32 * 32 + 32) * nodes
    out = {'nodes': nodes, 'G': G, 'peak_bf16_dense_tflops': 2500.0}
    # mode 0: training forward (stores every activation); 1: inference (Q += out)
    for tag, kw in (('train', dict(mode=0)), ('infer', dict(mode=1, dst=src, oscale=osc, Q=Q))):
        for _ in range(3):
            eng.forward(u, src, sc, **kw)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            eng.forward(u, src, sc, **kw)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[tag] = {'ms_per_call': ms, 'us_per_node': ms / nodes * 1e3,
                    'tflops': flop / (ms * 1e-3) / 1e12}
    return out


def steady_state(dev, tau=1e-6, tau2=1e-5, max_steps=6000):
    """Second part of the metric: wall seconds until the reference residual
    (checked every 10 steps) falls below tau (and, from the same runs, tau2)
    on the Mach-2 hard-sphere normal
    shock (Ne=32, K=3, N=32^3, M=12, N_rho=16, FP32, SSP-RK2 dt=5e-4), plain
    spectral versus the learned closure (100-call spectral warm start, checks
    every 10 calls on 25 % of the nodes, tolerance 0.05, the network
    re-evaluated every 4th RHS call and its output reused in between), and
    the same closure schedule without the network (Q_Shakov + the lagged
    spectral correction: closure-use-net = false), the method the network is
    measured against."""
    from dgfs1d_gpu_amd.closure_train import shock_deck
    from dgfs1d_gpu_amd.steady import run_to_steady
    weights = os.path.join(ROOT, 'dgfs1d_gpu_amd', 'data', 'closure_hs_n32.npz')
    extra = ('closure = learned\nclosure-weights = %s\nclosure-check-every = 10\n'
             'closure-tol = 0.05\nclosure-sample = 0.25\nclosure-start = 100\n'
             'closure-net-every = 4\n' % weights)
    out = {'workload': 'Mach-2 hard-sphere normal shock Ne=32 K=3 N=32^3 M=12 Nrho=16 FP32 '
                       'SSP-RK2 dt=5e-4', 'tau': tau, 'tau2': tau2}
    for tag, text in (('spectral', shock_deck(mach=2.0)),
                      ('learned', shock_deck(mach=2.0, extra=extra)),
                      ('no_network', shock_deck(mach=2.0,
                                                extra=extra + 'closure-use-net = false\n'))):
        s, rep, _ = run_to_steady(text, [tau, tau2], max_steps, 10, dev)
        hit = rep['reached'].get('%g' % tau)
        hit2 = rep['reached'].get('%g' % tau2)
        out[tag] = {'wall_s': hit['wall_s'] if hit else None,
                    'steps': hit['steps'] if hit else None,
                    'wall_s_tau2': hit2['wall_s'] if hit2 else None,
                    'steps_tau2': hit2['steps'] if hit2 else None,
                    'ms_per_step': rep['ms_per_step']}
        if tag != 'spectral':
            cl = s.sm.closure
            out[tag].update(nodes_learned=int(cl.learned.sum()), nodes=int(cl.ns),
                            fallbacks=int(cl.fallbacks), readmissions=int(cl.readmissions),
                            net_every=int(cl.net_every))
        del s
    if out['spectral']['wall_s'] and out['learned']['wall_s']:
        out['speedup'] = out['spectral']['wall_s'] / out['learned']['wall_s']
    if out['spectral']['wall_s_tau2'] and out['learned']['wall_s_tau2']:
        out['speedup_tau2'] = out['spectral']['wall_s_tau2'] / out['learned']['wall_s_tau2']
    if out['spectral']['wall_s'] and out['no_network']['wall_s']:
        out['speedup_no_network'] = out['spectral']['wall_s'] / out['no_network']['wall_s']
    return out


def relax_models(dev, reps, config):
    """One collision call of each relaxation model on the Couette workload's
    grid (Ne=4, K=3, N=32^3, FP32: 12 node slices), timed with HIP events on
    the current stream over reps calls: ms per call and the rate of its
    algorithmic bytes, f read twice and Q read and written (4 x 12 x N^3 x 4 B;
    the Newton models re-read the velocity grid on every iteration)."""
    import configparser
    import io
    import torch
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D
    out = {'workload': 'Couette grid Ne=4 K=3 N=32^3 FP32, one collision call (12 slices)'}
    for model in ('shakov-direct-gll', 'bgk-direct-gll', 'esbgk-direct-gll', 'bgk-gll',
                  'esbgk-gll'):
        cp = configparser.ConfigParser(inline_comment_prefixes=[';'])
        cp.optionxform = str
        cp.read(config)
        cp.set('scattering-model', 'type', model)
        if not cp.has_option('scattering-model', 'muRef'):
            cp.set('scattering-model', 'muRef', '2.117e-5')
        buf = io.StringIO()
        cp.write(buf)
        s = DGFSStd1D(Dictionary(buf.getvalue()), device=dev, writers=False)
        Q = torch.zeros_like(s.u)
        for _ in range(3):
            s.sm.collide(s.u, Q)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            s.sm.collide(s.u, Q)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 4 * s.u.numel() * s.u.element_size()
        out[model] = {'ms_per_call': ms, 'GB_per_s': nbytes / (ms * 1e-3) / 1e9}
        del s
    return out


# the reference's published single-GPU times (README.md:104-115): binary VSS
# Couette, FP64, M=12, N_rho=8, first 52 steps on one Tesla P100
BI_PUBLISHED = {(72, 3, 20): 47.58, (72, 3, 32): 126.601, (72, 3, 48): 391.943,
                (72, 6, 20): 94.682, (72, 6, 32): 253.016, (72, 6, 48): 782.343,
                (216, 3, 20): 141.754, (216, 3, 32): 378.956, (216, 3, 48): 1172.907,
                (216, 6, 20): 283.091, (216, 6, 32): 759.149, (216, 6, 48): 2347.099}


def bi_couette_published(dev, cases=((72, 3, 20),), steps=52):
    """The reference's only published GPU benchmark: the binary-mixture VSS
    Couette case (walls at -50/+50 m/s, 273 K; Ar-Kr mass ratio, Mpre=192,
    nodal-sem-gll, SSP-RK2 dt=0.01, FP64, M=12, N_rho=8), wall time of the
    first 52 steps after 2 untimed warm-up steps, against the published P100
    time of the same phase space Ne/K/N^3."""
    from configparser import ConfigParser
    import io
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    import torch
    base = open(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'bi_fourier.ini')).read()
    out = []
    for Ne, K, N in cases:
        cp = ConfigParser(inline_comment_prefixes=[';'])
        cp.optionxform = str
        cp.read_string(base)
        for sect in ('dgfsresidual', 'dgfsdistwriter', 'dgfsmomwriter'):
            cp.remove_section(sect)
        for (sect, key), v in {('mesh', 'Ne'): Ne, ('basis', 'K'): K, ('basis', 'Nq'): K + 1,
                               ('velocity-mesh', 'Nv'): N, ('velocity-mesh', 'Nrho'): 8,
                               ('time-integrator', 'scheme'): 'ssp-rk2',
                               ('soln-bcs-xlo', 'T'): 273, ('soln-bcs-xhi', 'T'): 273,
                               ('soln-bcs-xlo', 'uy'): -50, ('soln-bcs-xhi', 'uy'): 50}.items():
            cp.set(sect, key, str(v))
        buf = io.StringIO()
        cp.write(buf)
        s = DGFSBi1D(Dictionary(buf.getvalue()), device=dev, writers=False)
        s.run(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        pub = BI_PUBLISHED.get((Ne, K, N))
        out.append({'phase_space': '%d/%d/%d^3' % (Ne, K, N), 'steps': steps, 'wall_s': wall,
                    'published_p100_s': pub, 'speedup_vs_published': pub / wall if pub else None,
                    'point_steps_per_s': Ne * K * N ** 3 * steps / wall})
        del s
    return out


def bi_fourier_n48(dev, steps=3, warmup=1, scheme=None):
    """The largest mixture grid (configs/bi_fourier_n48.ini: Ne=8, modal K=5,
    N=48^3, M=24, N_rho=24, Mpre=192, FP32 collisions with the FP64 mixture
    projection, IMEX-BDF2 with species-wise Shakov penalties; scheme = an
    explicit kind such as ssp-rk2 runs the explicit driver instead): set-up
    time (including the IMEX start-up history) and ms per step."""
    from dgfs1d_gpu_amd.bi.abi import DGFSBiAstd1D
    from dgfs1d_gpu_amd.bi.bi import DGFSBi1D
    from dgfs1d_gpu_amd.dictionary import Dictionary
    import torch
    text = open(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'bi_fourier_n48.ini')).read()
    cfg = Dictionary(text)
    if scheme:
        cfg.set('time-integrator', 'scheme', scheme)
    imex = cfg.lookup('time-integrator', 'scheme').startswith(('imex', 'ars', 'sspl', 'bdf',
                                                                'boltz'))
    t0 = time.perf_counter()
    s = (DGFSBiAstd1D if imex else DGFSBi1D)(cfg, device=dev, writers=False)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    s.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.run(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    K, Ne, n = s.K, s.Ne, s.Nv
    out = {'workload': 'binary VSS Fourier Ne=8 modal K=5 N=48^3 M=24 Nrho=24 Mpre=192 FP32 '
                       '+ FP64 mixture projection, %s' % cfg.lookup('time-integrator', 'scheme'),
           'tasks_per_pair': s.sm.tasks['ntask'], 'setup_s': setup, 'steps': steps,
           'ms_per_step': wall / steps * 1e3,
           'point_steps_per_s': Ne * K * n * steps / wall,
           'chunk': s.sm.plans[0, 0].query()['chunk']}
    del s
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=52)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--config', default=os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs',
                                                     'couette_prod.ini'))
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-accuracy', action='store_true',
                    help='skip the FP64 comparison run after the timed region')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the seeded 64-problem uneven batch (autotuned, graph-captured)')
    ap.add_argument('--batch-seed', type=int, default=0)
    ap.add_argument('--no-steady', action='store_true',
                    help='skip the time-to-steady-state comparison (learned closure vs '
                         'spectral, Mach-2 hard-sphere shock, residual 1e-5)')
    ap.add_argument('--no-bi', action='store_true',
                    help='skip the binary-mixture Couette run against the published P100 time')
    ap.add_argument('--bi-only', default=None, metavar='Ne/K/N[,Ne/K/N...]',
                    help='only run the binary-mixture Couette benchmark at these phase spaces')
    ap.add_argument('--bi-n48', type=int, default=0, metavar='STEPS',
                    help='only time STEPS steps of the N=48 mixture deck (configs/bi_fourier_n48.ini)')
    ap.add_argument('--bi-n48-scheme', default=None,
                    help='time scheme of --bi-n48 (default: the deck\'s imex-lm-bdf2)')
    ap.add_argument('--closure-forward', type=int, default=0, metavar='NODES',
                    help='only time the learned-closure network forward on NODES slices of 32^3')
    ap.add_argument('--relax-models', type=int, default=0, metavar='REPS',
                    help='only time REPS collision calls of each relaxation model on the '
                         'Couette grid')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='only run 5 device-to-device copies of MIB MiB (the known-size '
                         'kernel that scales the FETCH_SIZE / WRITE_SIZE counters)')
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        return 2

    import torch
    if args.bi_n48:
        print(json.dumps(bi_fourier_n48(torch.device('cuda:0'), args.bi_n48,
                                        scheme=args.bi_n48_scheme)))
        return 0
    if args.bi_only:
        cases = [tuple(int(x) for x in c.split('/')) for c in args.bi_only.split(',')]
        print(json.dumps(bi_couette_published(torch.device('cuda:0'), cases)))
        return 0
    if args.relax_models:
        print(json.dumps(relax_models(torch.device('cuda:0'), args.relax_models, args.config)))
        return 0
    if args.closure_forward:
        print(json.dumps(closure_forward(args.closure_forward, 32, 10, torch.device('cuda:0'))))
        return 0
    if args.copy_calibration:
        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda:0')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_calibration_bytes': nbytes, 'launches': 5}))
        return 0
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.std import DGFSStd1D

    text = open(args.config).read()
    cfg = Dictionary(text)
    dev = torch.device('cuda:0')
    s = DGFSStd1D(cfg, device=dev, writers=False)
    plan = s.sm.plan
    for _ in range(args.warmup):
        s.step()
    torch.cuda.synchronize()

    nrhs = {'euler': 1, 'ssp-rk2': 2, 'ssp-rk3': 3, 'lse-rk45': 5, 'lesrk-45': 5}[
        cfg.lookup('time-integrator', 'scheme')]
    plan.timing(args.steps * nrhs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    stage_ms, ncalls = plan.timing_read()
    plan.timing(0)

    N, Ne, K = s.vm.Nv(), s.Ne, s.K
    n = N ** 3
    B = s.vm.M() * s.vm.Nrho()
    ntask = s.ntasks
    ref = getattr(s.sm, '_ref', None)
    fgs = plan.tasks['fg_same'] or ref is not None
    # k_gain, algorithmic count (the survey's, of the work the kernel replaces):
    # per (task, p, q) the reference's two inverse 3D transforms (cos and sin
    # parts) per function (f, and g when f != g) and its pointwise count, 14
    # per (p, q, k) for f = g, 26 for the bilinear form
    gain_flops = ntask * B * ((2 if fgs else 4) * fft_flops(n) + (14 if fgs else 26) * n)
    # executed count: one packed inverse transform per (p, q) and function (h
    # alone in the reference-state form), and 22 pointwise flops there (f = g's
    # 14 and one more complex product and sum with the stored transforms of f0)
    if ref is not None:
        nvar, pw = 1, 22
    else:
        nvar, pw = (1, 14) if plan.tasks['fg_same'] else (2, 26)
    gain_exec = ntask * B * (nvar * fft_flops(n) + pw * n)
    # HBM bytes the reference-state gain streams per launch: u+ and u- of f0
    # (two complex values) per (task, p, q, point)
    gain_stream = ntask * B * n * 4 * (4 if cfg.dtype.__name__ == 'float32' else 8) \
        if ref is not None else 0
    nc = max(ncalls, 1)
    stages = dict(zip(plan.STAGES, (m / nc for m in stage_ms)))
    gain_ms = stages['gain']
    achieved = gain_flops / (gain_ms * 1e-3) / 1e12
    executed = gain_exec / (gain_ms * 1e-3) / 1e12
    coll_ms = sum(stage_ms) / nc
    q_survey = (4 + 5 * B) * fft_flops(n) + 26 * B * n   # per Q, reference-structured count
    value = Ne * K * n * args.steps / wall
    q = plan.query()
    workload = 'couette Ne=%d K=%d N=%d^3 M=%d Nrho=%d %s dt=%g%s' % (
        Ne, K, N, s.vm.M(), s.vm.Nrho(), cfg.lookup('time-integrator', 'scheme'), s.dt,
        '' if ref is None else ' reference-state R=%d' % ref['R'])
    tr = GAIN_TRAFFIC_KIB.get(workload)
    traffic = None if tr is None else (tr[0] * tr[1] + tr[2] * tr[3]) * 1024.0
    out = {
        'metric': 'phase-space updates/sec at Ne/K/N^3',
        'value': value,
        'unit': 'phase-space point-steps/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32' if cfg.dtype.__name__ == 'float32' else 'f64',
        'data': 'synthetic (deterministic Maxwellian initial state, diffuse walls)',
        'config': {'workload': workload,
            'Ne': Ne, 'K': K, 'N': N, 'M': s.vm.M(), 'Nrho': s.vm.Nrho(),
            'collisions_per_rhs': ntask, 'split_S': q['split'], 'qgroups': q['qgroups'],
            'pgroups': q['pgroups'],
            'collision_form': 'direct' if ref is None else
            'reference-state, FP64 Q(f0,f0) refreshed every %d calls' % ref['R']},
        'roofline': {'bound': 'fp32-valu', 'kernel': 'k_gain_pair', 'achieved': achieved,
                     'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / PEAK_FP32_TFLOPS, 'traffic': traffic,
                     'flops_per_launch': gain_flops, 'avg_launch_ms': gain_ms,
                     'executed_flops_per_launch': gain_exec, 'executed_tflops': executed,
                     'executed_frac': executed / PEAK_FP32_TFLOPS,
                     'peak_without_packed_fp32': PEAK_FP32_TFLOPS / 2,
                     'streamed_bytes_per_launch': gain_stream,
                     'streamed_GBs': gain_stream / (gain_ms * 1e-3) / 1e9,
                     'streamed_frac_of_hbm': gain_stream / (gain_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
        'collision_stage_ms': stages,
        'collision_ms_per_rhs': coll_ms,
        'collision_integrals_per_s': ntask * ncalls / (sum(stage_ms) * 1e-3),
        'reference_equivalent_tflops': ntask * q_survey / (coll_ms * 1e-3) / 1e12,
    }
    if not args.no_accuracy:
        out['accuracy'] = fp64_accuracy(text, args.warmup + args.steps, s, dev)
    if ref is not None:
        out['direct_form'] = direct_form(text, args.steps, args.warmup, dev,
                                         accuracy=not args.no_accuracy)
    if not args.no_batch or not args.no_steady or not args.no_bi:
        del s
    if not args.no_batch:
        out['batched'] = batched_workload(args.batch_seed, 64, 20, args.warmup, dev)
    if not args.no_steady:
        out['time_to_steady_state'] = steady_state(dev)
    if not args.no_bi:
        out['bi_couette_published'] = bi_couette_published(dev, ((72, 3, 20), (72, 3, 32)))
    if not args.no_cpu_baseline:
        out['cpu_baseline'] = cpu_baseline(text)
        out['gpu_over_cpu'] = value / out['cpu_baseline']['value']
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
