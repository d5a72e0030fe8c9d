"""Training data and training for the learned collision closure.

    python -m dgfs1d_gpu_amd.closure_train --out weights.npz [--machs 1.2,2,3]
        [--temps 0.8,1.4,2.0] [--steps 120] [--every 10] [--epochs 8]

Data: (f, R) pairs with R = Q_spectral(f) - Q_Shakov(f) taken from the FP32
spectral solver on hard-sphere normal-shock decks (Rankine-Hugoniot states at
the given upstream Mach numbers and temperatures, ``shock_deck``) and on
Couette decks with random wall speeds; every ``every`` steps every node of the
run is one pair.  Scaling: f / rho and R / (prefac rho^2) (see learned.py).
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


def couette_deck(uw=0.5, Tw=1.0, N=32, Ne=4, K=3, M=12, Nrho=16, precision='float', dt=5e-4):
    """Couette deck of the production case with wall speed uw u0 and wall
    temperature Tw T0 (hard-sphere gas so the closure sees one gas model)."""
    u0 = math.sqrt(2.0 * R0 / HELIUM * 273.0)
    return shock_deck(N=N, Ne=Ne, K=K, M=M, Nrho=Nrho, precision=precision)\
        .split('[soln-ics]')[0].replace('xlo = -15e-3', 'xlo = -0.5').replace(
            'xhi = 15e-3', 'xhi = 0.5').replace('H0 = 30e-3', 'H0 = 1.0').replace(
            'T0 = 223.0', 'T0 = 273').replace('Tmax = 223.0', 'Tmax = 273').replace(
            'dt = 0.0005', 'dt = %g' % dt) + f"""[soln-ics]
type = maxwellian
rho = 1.91607e-5
T = 273
ux = 0
uy = 0
uz = 0
[soln-bcs-xlo]
type = dgfs-wall-diffuse
T = {273 * Tw}
ux = 0
uy = {-uw * u0}
uz = 0
[soln-bcs-xhi]
type = dgfs-wall-diffuse
T = {273 * Tw}
ux = 0
uy = {uw * u0}
uz = 0
"""


def collect(text, steps, every, device, store, log=None):
    """Run the spectral FP32 solver; every ``every`` steps append (f, R)
    of all nodes, scaled, to store (lists of device tensors)."""
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
            f = s.u.view(ns, n).float()
            rho = f.double().sum(1) * vm.cw()
            store['f'].append((f.double() / rho[:, None]).float())
            store['t'].append(((Qs - Qk).view(ns, n).double() /
                               (sm.prefac * rho * rho)[:, None]).float())
            store['qs'].append((Qs.view(ns, n).double() / (sm.prefac * rho * rho)[:, None]).float())
        if k < steps:
            s.step()
    if log:
        log('collected %d snapshots x %d nodes' % (steps // every + 1, ns))
    return s


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    ap.add_argument('--machs', default='1.2,1.6,2.0,2.5,3.0')
    ap.add_argument('--temps', default='0.8,1.4,2.0')
    ap.add_argument('--couette', type=int, default=4, help='random Couette decks')
    ap.add_argument('--steps', type=int, default=120)
    ap.add_argument('--every', type=int, default=10)
    ap.add_argument('--epochs', type=int, default=8)
    ap.add_argument('--batch', type=int, default=16)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--report', default=None, help='write a JSON summary here')
    args = ap.parse_args(argv)
    import torch
    from .learned import ClosureNet, train
    dev = torch.device('cuda:0')
    log = lambda m: print(m, flush=True)
    store = dict(f=[], t=[], qs=[])
    t0 = time.perf_counter()
    rng = np.random.default_rng(args.seed)
    for mach in [float(x) for x in args.machs.split(',')]:
        for temp in [float(x) for x in args.temps.split(',')]:
            collect(shock_deck(mach=mach, T1=223.0 * temp), args.steps, args.every, dev, store,
                    log)
    for _ in range(args.couette):
        collect(couette_deck(uw=float(rng.uniform(0.1, 1.0)), Tw=float(rng.uniform(0.8, 1.5))),
                args.steps, args.every, dev, store, log)
    f = torch.cat(store['f'])
    tgt = torch.cat(store['t'])
    qs = torch.cat(store['qs'])
    t_data = time.perf_counter() - t0
    S = f.shape[0]
    G = round(f.shape[1] ** (1.0 / 3.0))
    perm = torch.tensor(rng.permutation(S), device=dev)
    nval = max(1, S // 10)
    val, tr = perm[:nval], perm[nval:]
    log('%d pairs (G=%d), %d for validation; data %.1f s' % (S, G, nval, t_data))
    # output scale: the RMS per-voxel magnitude of the scaled collision
    # integrals; |Q| floor: 10 % of their median norm (both in scaled units)
    qn = qs.double().norm(dim=1)
    sigma = float(qn.pow(2).mean().sqrt()) / math.sqrt(qs.shape[1])
    qfloor = 0.1 * float(qn.median()) / sigma
    net = ClosureNet(dev, seed=args.seed, meta=dict(out_scale=sigma, qfloor=qfloor))
    tn = (tgt / sigma).contiguous()
    qsn = (qs / sigma).contiguous()
    t1 = time.perf_counter()
    hist = train(net, f.index_select(0, tr), tn.index_select(0, tr), G, batch=args.batch,
                 epochs=args.epochs, lr=args.lr, seed=args.seed, log=log,
                 ref=qsn.index_select(0, tr), floor2=qfloor ** 2)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    net.save(args.out)
    rep = evaluate(net, f.index_select(0, val), tn.index_select(0, val),
                   qsn.index_select(0, val), G, qfloor)
    rep.update(pairs=S, grid=G, data_s=t_data, train_s=t_train, epochs=args.epochs,
               loss_history=hist, out_scale=sigma, qfloor=qfloor,
               inference=time_inference(net, f, G))
    log(json.dumps(rep))
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
