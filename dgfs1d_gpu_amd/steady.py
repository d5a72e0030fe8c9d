"""Time to steady state: the plain FP32 spectral path versus the learned closure.

    python -m dgfs1d_gpu_amd.steady --weights w.npz [--mach 2.0] [--tau 1e-5]

Runs the hard-sphere normal shock (closure_train.shock_deck: Ne=32, K=3,
N=32^3, M=12, N_rho=16, FP32, SSP-RK2 dt=5e-4) from the same initial
profile twice -- spectral collision, then Q_Shakov + network with the
spectral fallback -- until the reference residual
sqrt(sum (f^{n+1} - f^n)^2 / sum (f^n)^2) (dgfs1D/std/residual.py:21-29),
checked every ``nres`` steps, falls below each threshold.  Reports wall
seconds and steps per threshold, the closure's fallback statistics and the
moment differences of the two final states.
"""
import argparse
import json
import sys
import time

import numpy as np


def run_to_steady(text, taus, max_steps, nres, device, log=None):
    import torch
    from .dictionary import Dictionary
    from .std.std import DGFSStd1D
    from .transport import Moments, Residual
    s = DGFSStd1D(Dictionary(text), device=device, writers=False)
    res = Residual(device)
    prev = s.u.clone()
    taus = sorted(taus, reverse=True)
    hit = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step = 0
    r = float('inf')
    while step < max_steps and len(hit) < len(taus):
        check = (step + 1) % nres == 0
        if check:
            prev.copy_(s.u)                 # f^n of the step that is checked
        s.step()
        step += 1
        if check:
            r = res(s.u, prev)
            now = time.perf_counter() - t0
            for tau in taus:
                if tau not in hit and r < tau:
                    hit[tau] = dict(wall_s=now, steps=step)
            if log and step % (nres * 20) == 0:
                log('step %d residual %.3e (%.1f s)' % (step, r, now))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    mom = Moments(s.basis.interpMat, s.K, s.Ne, s.vm, device)
    m = mom(s.u).reshape(-1, 9)
    return s, dict(steps=step, wall_s=wall, ms_per_step=wall / step * 1e3, last_residual=r,
                   reached={('%g' % k): v for k, v in hit.items()}), m


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--weights', required=True)
    ap.add_argument('--mach', type=float, default=2.0)
    ap.add_argument('--taus', default='1e-4,1e-5,1e-6')
    ap.add_argument('--max-steps', type=int, default=4000)
    ap.add_argument('--nres', type=int, default=10)
    ap.add_argument('--tol', type=float, default=0.05)
    ap.add_argument('--check-every', type=int, default=10)
    ap.add_argument('--sample', type=float, default=0.25)
    ap.add_argument('--start', default='100',
                    help='RHS calls on the spectral operator before the closure takes over '
                         '(comma-separated: one learned run per value)')
    ap.add_argument('--sparse', default=None, metavar='CHECK_EVERY:SAMPLE',
                    help='also run the closure with and without the network at this sparser '
                         'check schedule (first --start value)')
    ap.add_argument('--net-every', default='1',
                    help='closure-net-every of the learned runs (comma-separated: one learned '
                         'run per value, first --start value)')
    ap.add_argument('--out', default=None)
    args = ap.parse_args(argv)
    import torch
    from .closure_train import shock_deck
    dev = torch.device('cuda:0')
    log = lambda m: print(m, flush=True)
    taus = [float(x) for x in args.taus.split(',')]
    base = shock_deck(mach=args.mach)
    _, spec, m_spec = run_to_steady(base, taus, args.max_steps, args.nres, dev, log)
    log('spectral: ' + json.dumps(spec))
    runs = {}
    starts = [int(x) for x in str(args.start).split(',')]
    ce, sa = args.check_every, args.sample
    plan = [('learned' if i == 0 else 'learned_start%d' % st, True, st, ce, sa)
            for i, st in enumerate(starts)] + [('shakov_corrected', False, starts[0], ce, sa)]
    nets = [int(x) for x in str(args.net_every).split(',')]
    plan = [p + (nets[0],) for p in plan] + [('learned_net%d' % k, True, starts[0], ce, sa, k)
                                            for k in nets[1:]]
    if args.sparse:
        ce2, sa2 = args.sparse.split(':')
        plan += [('learned_sparse', True, starts[0], int(ce2), float(sa2)),
                 ('shakov_corrected_sparse', False, starts[0], int(ce2), float(sa2))]
        plan = [p if len(p) == 6 else p + (nets[0],) for p in plan]
    for tag, use_net, start, ce_, sa_, ne_ in plan:
        extra = ('closure = learned\nclosure-weights = %s\nclosure-check-every = %d\n'
                 'closure-tol = %g\nclosure-sample = %g\nclosure-use-net = %s\n'
                 'closure-start = %d\nclosure-net-every = %d\n'
                 % (args.weights, ce_, args.tol, sa_, 'true' if use_net else 'false', start, ne_))
        s, rep, m = run_to_steady(shock_deck(mach=args.mach, extra=extra), taus, args.max_steps,
                                  args.nres, dev, log)
        cl = s.sm.closure
        rep['nodes_learned_at_end'] = int(cl.learned.sum())
        rep['fallbacks'] = int(cl.fallbacks)
        rep['readmissions'] = int(cl.readmissions)
        rep['nodes_total'] = int(cl.ns)
        rep['checks'] = len(cl.checks)
        errs = np.array([e for _, d in cl.checks for e in d.values()])
        rep['check_err_median'] = float(np.median(errs)) if len(errs) else None
        rep['check_err_p90'] = float(np.percentile(errs, 90)) if len(errs) else None
        rep['start'] = start
        rep['check_every'] = ce_
        rep['sample'] = sa_
        rep['net_every'] = ne_
        rep['speedup_to_tau'] = {k: spec['reached'][k]['wall_s'] / v['wall_s']
                                 for k, v in rep['reached'].items() if k in spec['reached']}
        log('%s: %s' % (tag, json.dumps(rep)))
        runs[tag] = (rep, m)
        del s
    lrn, m_lrn = runs['learned']
    names = ('rho', 'ux', 'uy', 'T', 'qx', 'qy', 'Pxx', 'Pyy', 'Pxy')
    dm = {nm: float(np.abs(m_lrn[:, i] - m_spec[:, i]).max() /
                    max(np.abs(m_spec[:, i]).max(), 1e-300)) for i, nm in enumerate(names)}
    speed = {}
    for k in lrn['reached']:
        if k in spec['reached']:
            speed[k] = spec['reached'][k]['wall_s'] / lrn['reached'][k]['wall_s']
    scale = np.abs(m_spec[:, :4]).max(axis=0)          # rho, ux, uy, T scales
    ref = {'rho': scale[0], 'ux': max(scale[1], scale[2]), 'uy': max(scale[1], scale[2]),
           'T': scale[3]}
    # q and P are compared relative to rho T^1.5 and rho T (their natural sizes)
    rT = float(np.abs(m_spec[:, 0] * m_spec[:, 3]).max())
    rT15 = float(np.abs(m_spec[:, 0] * np.abs(m_spec[:, 3]) ** 1.5).max())
    for nm in ('qx', 'qy'):
        ref[nm] = rT15
    for nm in ('Pxx', 'Pyy', 'Pxy'):
        ref[nm] = rT
    dm_scaled = {nm: float(np.abs(m_lrn[:, i] - m_spec[:, i]).max() / ref[nm])
                 for i, nm in enumerate(names)}
    out = dict(mach=args.mach, spectral=spec, learned=lrn,
               other_runs={k: v[0] for k, v in runs.items()
                           if k not in ('learned', 'shakov_corrected')},
               moment_diff_scaled_learned_vs_spectral=dm_scaled,
               shakov_corrected_without_network=runs['shakov_corrected'][0],
               moment_rel_diff_learned_vs_spectral=dm, speedup_to_tau=speed,
               per_step_speedup=spec['ms_per_step'] / lrn['ms_per_step'])
    print(json.dumps(out))
    if args.out:
        with open(args.out, 'w') as fh:
            json.dump(out, fh)
    return 0


if __name__ == '__main__':
    sys.exit(main())
