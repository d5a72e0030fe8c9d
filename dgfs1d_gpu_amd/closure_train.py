"""Training data and training for the learned collision closure.

    python -m dgfs1d_gpu_amd.closure_train --out weights.npz [--shocks 32]
        [--couette 12] [--fourier 12] [--epochs 10] [--report r.json]

Data: (f, R) pairs with R = Q_spectral(f) - Q_Shakov(f) taken from the FP32
spectral solver (hard-sphere gas, 32^3 velocity grid) on normal-shock decks
with Rankine-Hugoniot states at random upstream Mach numbers in [1.2, 3] and
temperatures in [0.8, 2] T0 (``shock_deck``), and on Couette (random wall
speeds, equal wall temperatures) and Fourier (resting walls at two random
temperatures) decks (``wall_deck``); every ``every`` steps every node of the
run is one pair.  The pairs live in preallocated device buffers.  Scaling:
f / rho and R / (prefac rho^2) (see learned.py).
"""
import argparse
import json
import math
import sys
import time

import numpy as np

R0 = 8.3144598
HELIUM = 4.0047236985e-3


def rankine_hugoniot(mach, gamma=5.0 / 3.0):
    """Downstream/upstream ratios (rho, u, T) of a normal shock."""
    m2 = mach * mach
    r = (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    p = (2.0 * gamma * m2 - (gamma - 1.0)) / (gamma + 1.0)
    return r, 1.0 / r, p / r


def shock_deck(mach=2.0, T1=223.0, Ne=32, K=3, N=32, M=12, Nrho=16, precision='float',
               dt=5e-4, tend=5.0, rho1=1.91607e-5, dev=7.0, scheme='ssp-rk2', extra=''):
    """Hard-sphere (omega = 0.5) normal shock between Rankine-Hugoniot states,
    after the reference's examples/std/normalshock/dgfs1D.ini layout: x in
    [-15, 15] mm, linear initial profile, inlet boundaries."""
    r, ur, tr = rankine_hugoniot(mach)
    u1 = mach * math.sqrt(5.0 / 3.0 * R0 * T1 / HELIUM)
    rho2, u2, T2 = rho1 * r, u1 * ur, T1 * tr
    rule = 'octahedral' if M == 12 else 'womersley'
    return f"""[config]
precision = {precision}
[mesh]
xlo = -15e-3
xhi = 15e-3
Ne = {Ne}
H0 = 30e-3
[basis]
kind = nodal-sem-gll
K = {K}
Nqr = 4
[non-dim]
T0 = {T1}
rho0 = {rho1}
molarMass0 = {HELIUM}
[velocity-mesh]
Nv = {N}
Nrho = {Nrho}
dev = {dev}
cmax = 0
Tmax = {T1}
[spherical-design-rule]
ssrule = {rule}
M = {M}
[scattering-model]
type = vhs
omega = 0.5
dRef = 2.17e-10
Tref = 273
muRef = 1.865e-5
{extra}
[time-integrator]
scheme = {scheme}
controller = none
tstart = 0.0
tend = {tend}
dt = {dt}
[soln-ics]
type = maxwellian-expr
rho = {rho1} + ({rho2}-{rho1})*(x+0.5)
T = {T1} + ({T2}-{T1})*(x+0.5)
ux = {u1} + ({u2}-{u1})*(x+0.5)
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-inlet
rho = {rho1}
T = {T1}
ux = {u1}
uy = 0
uz = 0
[soln-bcs-xhi]
type = dgfs-inlet
rho = {rho2}
T = {T2}
ux = {u2}
uy = 0
uz = 0
"""


def wall_deck(uy_lo=-0.5, uy_hi=0.5, T_lo=1.0, T_hi=1.0, N=32, Ne=4, K=3, M=12, Nrho=16,
              precision='float', dt=5e-4):
    """Diffuse walls at x = -/+ 0.5 (H0 = 1 m) with wall speeds uy u0 and
    wall temperatures T T0 (T0 = 273 K), gas initially at rest at T0
    (hard-sphere gas, the closure's gas model)."""
    u0 = math.sqrt(2.0 * R0 / HELIUM * 273.0)
    return shock_deck(N=N, Ne=Ne, K=K, M=M, Nrho=Nrho, precision=precision)\
        .split('[soln-ics]')[0].replace('xlo = -15e-3', 'xlo = -0.5').replace(
            'xhi = 15e-3', 'xhi = 0.5').replace('H0 = 30e-3', 'H0 = 1.0').replace(
            'T0 = 223.0', 'T0 = 273').replace('Tmax = 223.0', 'Tmax = %g' % (
                273 * max(1.0, T_lo, T_hi))).replace('dt = 0.0005', 'dt = %g' % dt) + f"""[soln-ics]
type = maxwellian
rho = 1.91607e-5
T = 273
ux = 0
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-wall-diffuse
T = {273 * T_lo}
ux = 0
uy = {uy_lo * u0}
uz = 0
[soln-bcs-xhi]
type = dgfs-wall-diffuse
T = {273 * T_hi}
ux = 0
uy = {uy_hi * u0}
uz = 0
"""


def couette_deck(uw=0.5, Tw=1.0, **kw):
    """Couette deck: walls moving at -/+ uw u0, both at Tw T0."""
    return wall_deck(-uw, uw, Tw, Tw, **kw)


def fourier_deck(T_lo=0.9, T_hi=1.1, **kw):
    """Fourier heat-transfer deck: resting walls at T_lo T0 and T_hi T0."""
    return wall_deck(0.0, 0.0, T_lo, T_hi, **kw)


def deck_nodes(text):
    """K * Ne of a deck."""
    import configparser
    cp = configparser.ConfigParser(inline_comment_prefixes=[';'])
    cp.read_string(text)
    return int(cp['basis']['K']) * int(cp['mesh']['Ne'])


def collect(text, steps, every, device, buf, off, log=None):
    """Run the spectral FP32 solver; every ``every`` steps write (f, R, Q)
    of all nodes, scaled, into the preallocated buffers buf['f'], buf['t'],
    buf['qs'] at rows off, off + 1, ...; returns the next free row."""
    import torch
    from .dictionary import Dictionary
    from .relax import Shakov
    from .std.std import DGFSStd1D
    s = DGFSStd1D(Dictionary(text), device=device, writers=False)
    sm, vm = s.sm, s.vm
    ns, n = s.K * s.Ne, vm.vsize()
    cfg = s.cfg
    omega = sm.omega
    t0 = vm.H0() / vm.u0()
    visc = float(cfg.lookupfloat('scattering-model', 'muRef')) * \
        (vm.T0() / float(cfg.lookupfloat('scattering-model', 'Tref'))) ** omega
    p0 = vm.n0() * vm.R0 / vm.NA * vm.T0()
    sk = Shakov(n, vm.d_cvx(), vm.d_cvy(), vm.d_cvz(), vm.cw(), t0 * p0 / visc, omega,
                float(cfg.lookupordefault('scattering-model', 'Pr', 2. / 3.)), s.tdtype, device)
    sk.set_tasks(1, ns, [(i, [(i, 1.0)], [(i, 1.0)]) for i in range(ns)])
    from .collision import task_table
    tab = task_table(1, ns, [(i, [(i, 1.0)], [(i, 1.0)]) for i in range(ns)], s.tdtype, device)
    Qs, Qk = torch.zeros_like(s.u), torch.zeros_like(s.u)
    for k in range(steps + 1):
        if k % every == 0:
            Qs.zero_()
            Qk.zero_()
            sm.plan(s.u, Qs, tasks=tab)
            sk(s.u, Qk)
            f = s.u.view(ns, n).double()
            rho = f.sum(1) * vm.cw()
            den = (sm.prefac * rho * rho)[:, None]
            buf['f'][off:off + ns].copy_(f / rho[:, None])
            buf['t'][off:off + ns].copy_((Qs - Qk).view(ns, n).double() / den)
            buf['qs'][off:off + ns].copy_(Qs.view(ns, n).double() / den)
            off += ns
        if k < steps:
            s.step()
    if log:
        log('collected %d snapshots x %d nodes (%d pairs so far)' % (steps // every + 1, ns, off))
    return off


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    ap.add_argument('--shocks', type=int, default=32,
                    help='normal-shock runs at random Mach in [1.2, 3] and T in [0.8, 2] T0')
    ap.add_argument('--couette', type=int, default=12, help='random Couette runs')
    ap.add_argument('--fourier', type=int, default=12, help='random Fourier runs')
    ap.add_argument('--steps', type=int, default=1200, help='steps per shock run')
    ap.add_argument('--every', type=int, default=24, help='snapshot interval (shock runs)')
    ap.add_argument('--wall-ne', type=int, default=16, help='elements of the wall runs')
    ap.add_argument('--wall-steps', type=int, default=780)
    ap.add_argument('--wall-every', type=int, default=20)
    ap.add_argument('--epochs', type=int, default=10)
    ap.add_argument('--batch', type=int, default=16)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--report', default=None, help='write a JSON summary here')
    args = ap.parse_args(argv)
    import torch
    from .learned import ClosureNet, train
    dev = torch.device('cuda:0')
    log = lambda m: print(m, flush=True)
    rng = np.random.default_rng(args.seed)
    runs = []
    for _ in range(args.shocks):
        mach, temp = float(rng.uniform(1.2, 3.0)), float(rng.uniform(0.8, 2.0))
        runs.append(('shock M=%.2f T=%.2f' % (mach, temp), shock_deck(mach=mach, T1=223.0 * temp),
                     args.steps, args.every))
    for _ in range(args.couette):
        uw, tw = float(rng.uniform(0.1, 1.0)), float(rng.uniform(0.8, 2.0))
        runs.append(('couette uw=%.2f Tw=%.2f' % (uw, tw),
                     couette_deck(uw=uw, Tw=tw, Ne=args.wall_ne), args.wall_steps,
                     args.wall_every))
    for _ in range(args.fourier):
        tl, th = sorted(float(x) for x in rng.uniform(0.8, 2.0, 2))
        runs.append(('fourier Tl=%.2f Th=%.2f' % (tl, th),
                     fourier_deck(T_lo=tl, T_hi=th, Ne=args.wall_ne), args.wall_steps,
                     args.wall_every))
    S = sum((st // ev + 1) * deck_nodes(text) for _, text, st, ev in runs)
    G = 32
    n = G ** 3
    log('%d runs, %d pairs, %.1f GB of device buffers' % (len(runs), S, 3 * S * n * 4 / 1e9))
    buf = {k: torch.empty((S, n), dtype=torch.float32, device=dev) for k in ('f', 't', 'qs')}
    t0 = time.perf_counter()
    off = 0
    for name, text, st, ev in runs:
        log(name)
        off = collect(text, st, ev, dev, buf, off, log)
    assert off == S
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    f, tgt, qs = buf['f'], buf['t'], buf['qs']
    perm = rng.permutation(S)
    nval = max(1, S // 10)
    val, tr = np.sort(perm[:nval]), perm[nval:]
    log('%d pairs (G=%d), %d for validation; data %.1f s' % (S, G, nval, t_data))
    # output scale: the RMS per-voxel magnitude of the scaled collision
    # integrals; |Q| floor: 10 % of their median norm (both in scaled units)
    qn = torch.cat([qs[i:i + 4096].double().norm(dim=1) for i in range(0, S, 4096)])
    sigma = float(qn.pow(2).mean().sqrt()) / math.sqrt(n)
    qfloor = 0.1 * float(qn.median()) / sigma
    net = ClosureNet(dev, seed=args.seed, meta=dict(out_scale=sigma, qfloor=qfloor))
    tgt.div_(sigma)
    qs.div_(sigma)
    t1 = time.perf_counter()
    hist = train(net, f, tgt, G, batch=args.batch, epochs=args.epochs, lr=args.lr,
                 seed=args.seed, log=log, ref=qs, floor2=qfloor ** 2, rows=tr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    net.save(args.out)
    vi = torch.tensor(val, device=dev)
    rep = evaluate(net, f.index_select(0, vi), tgt.index_select(0, vi), qs.index_select(0, vi),
                   G, qfloor)
    rep.update(pairs=S, train_pairs=int(len(tr)), grid=G, runs=[r[0] for r in runs],
               data_s=t_data, train_s=t_train, epochs=args.epochs, loss_history=hist,
               out_scale=sigma, qfloor=qfloor, inference=time_inference(net, f, G))
    log(json.dumps({k: v for k, v in rep.items() if k != 'runs'}))
    if args.report:
        with open(args.report, 'w') as fh:
            json.dump(rep, fh)
    return 0


def time_inference(net, f, G, nb=96, reps=10):
    """Network forward over nb nodes (mode 0), HIP-event timed: ms per call,
    ms per node and the algorithmic TFLOP/s (3.74 GFLOP per node at G = 32:
    2 G^3 27 (32 + 32*32 + 32*32 + 32) multiply-adds)."""
    import torch
    from .learned import ClosureEngine
    nb = min(nb, f.shape[0])
    eng = ClosureEngine(net, G, nb)
    src = torch.arange(nb, dtype=torch.int32, device=net.device)
    ones = torch.ones(nb, dtype=torch.float32, device=net.device)
    fb = f[:nb].contiguous()
    for _ in range(3):
        eng.forward(fb, src, ones, mode=0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        eng.forward(fb, src, ones, mode=0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flop = 2.0 * G ** 3 * 27 * (32 + 32 * 32 + 32 * 32 + 32) * nb
    return dict(nodes=nb, ms_per_call=ms, ms_per_node=ms / nb, tflops=flop / (ms * 1e-3) / 1e12)


def evaluate(net, f, tgt, qs, G, qfloor=0.0, batch=16):
    """Relative L2 error of Q on held-out pairs, Shakov alone (R = 0) versus
    Shakov + network, relative to |Q| and to max(|Q|, qfloor) (scaled units,
    targets already divided by the output scale)."""
    import torch
    from .learned import ClosureEngine
    S = f.shape[0]
    nb = min(batch, S)
    eng = ClosureEngine(net, G, nb)
    src = torch.arange(nb, dtype=torch.int32, device=net.device)
    ones = torch.ones(nb, dtype=torch.float32, device=net.device)
    out = dict(sk=[], net=[], skf=[], netf=[])
    for i in range(0, S - nb + 1, nb):
        fb = f[i:i + nb].contiguous()
        pred = eng.forward(fb, src, ones, mode=0).view(nb, -1).double()
        t = tgt[i:i + nb].double()
        qn = qs[i:i + nb].double().norm(dim=1)
        qf = qn.clamp_min(qfloor)
        for key, num in (('sk', t.norm(dim=1)), ('net', (t - pred).norm(dim=1))):
            out[key].append((num / qn).cpu().numpy())
            out[key + 'f'].append((num / qf).cpu().numpy())
    r = {k: np.concatenate(v) for k, v in out.items()}
    rep = dict(val_pairs=int(len(r['sk'])))
    for k, name in (('sk', 'shakov'), ('net', 'learned'), ('skf', 'shakov_floored'),
                    ('netf', 'learned_floored')):
        rep['q_rel_l2_%s_median' % name] = float(np.median(r[k]))
        rep['q_rel_l2_%s_p90' % name] = float(np.percentile(r[k], 90))
    return rep


if __name__ == '__main__':
    sys.exit(main())
