"""Learned collision closure: network, training and the runtime fallback policy.

No reference counterpart.  The network (see ``ClosureNet``) predicts the
residual R(f) = Q(f, f) - Q_Shakov(f) of the spectral operator over the
Shakov model (dgfs1D/std/scattering.py:1636-1833) on the velocity grid.
Scaling: the input is f / rho (unit density, rho = cw sum f) and the target
R / (prefac rho^2), both exact consequences of Q and Q_Shakov being
quadratic in f and proportional to their prefactors (whose ratio depends
only on the gas model).  All arithmetic runs in the HIP kernels of
libdgfs_hip.so (closure.hip): bf16 MFMA for the 32-channel layers, FP32
accumulation, FP32 Adam on master weights.

``LearnedClosure`` is the runtime: Q = Q_Shakov + R_net for every node,
except nodes handed back to the spectral operator.  Every ``check_every``
calls the spectral operator evaluates a sample of the learned nodes; a node
whose relative L2 error (relative to max(|Q|, the network's |Q| floor), so
that near-equilibrium nodes with Q -> 0 are judged on an absolute scale)
exceeds ``tol`` is switched to the spectral operator for the rest of the run.
"""
import numpy as np

from . import _lib

C = 32


class ClosureNet:
    """FP32 master weights on the device plus bf16 packed copies."""

    SHAPES = dict(W1=(27, C), b1=(C,), W2=(27, C, C), b2=(C,), W3=(27, C, C), b3=(C,),
                  W4=(27, C), b4=(1,))

    def __init__(self, device, seed=0, params=None, meta=None):
        """meta: out_scale (the network predicts target / out_scale) and
        qfloor (floor of |Q| in scaled units for relative errors)."""
        import torch
        self.torch = torch
        self.device = device
        self.meta = dict(out_scale=1.0, qfloor=0.0)
        self.meta.update(meta or {})
        if params is None:
            params = self.init_params(seed)
        self.p = {k: torch.tensor(np.asarray(params[k], np.float32).reshape(self.SHAPES[k]),
                                  device=device) for k in self.SHAPES}
        self.packed = {k: torch.empty(27 * C * C, dtype=torch.bfloat16, device=device)
                       for k in ('W2f', 'W2t', 'W3f', 'W3t')}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step_count = 0
        self.pack()

    @classmethod
    def init_params(cls, seed):
        """He-uniform fan-in initialisation (fan-in 27 for layer 1, 864 after)."""
        rng = np.random.default_rng(seed)
        out = {}
        for k, shp in cls.SHAPES.items():
            if k.startswith('b'):
                out[k] = np.zeros(shp)
            else:
                fan = 27 if k == 'W1' else 27 * C
                lim = np.sqrt(6.0 / fan)
                out[k] = rng.uniform(-lim, lim, shp)
        out['W4'] *= 0.1
        return out

    def params_numpy(self):
        return {k: v.cpu().numpy().astype(np.float64) for k, v in self.p.items()}

    def save(self, path):
        np.savez(path, meta=np.array([self.meta['out_scale'], self.meta['qfloor']]),
                 **{k: v.cpu().numpy() for k, v in self.p.items()})

    @classmethod
    def load(cls, path, device):
        with np.load(path) as z:
            meta = dict(out_scale=float(z['meta'][0]), qfloor=float(z['meta'][1])) \
                if 'meta' in z.files else None
            return cls(device, params={k: z[k] for k in cls.SHAPES}, meta=meta)

    def pack(self, stream=None):
        P, K = self.p, self.packed
        _lib.check(_lib.lib().dgfs_closure_pack(
            _lib.ptr(P['W2']), _lib.ptr(P['W3']), _lib.ptr(K['W2f']), _lib.ptr(K['W2t']),
            _lib.ptr(K['W3f']), _lib.ptr(K['W3t']), _lib.stream_ptr(stream)))

    def cstruct(self):
        P, K = self.p, self.packed
        s = _lib.ClosureNet()
        s.W1, s.b1, s.b2, s.b3 = (P['W1'].data_ptr(), P['b1'].data_ptr(), P['b2'].data_ptr(),
                                  P['b3'].data_ptr())
        s.W4, s.b4 = P['W4'].data_ptr(), P['b4'].data_ptr()
        s.W2f, s.W2t, s.W3f, s.W3t = (K['W2f'].data_ptr(), K['W2t'].data_ptr(),
                                      K['W3f'].data_ptr(), K['W3t'].data_ptr())
        return s

    def gstruct(self):
        s = _lib.ClosureGrad()
        for k in self.SHAPES:
            setattr(s, k, self.g[k].data_ptr())
        return s


class ClosureEngine:
    """Buffers for one batch shape (nb slices of G^3) and the kernel calls."""

    def __init__(self, net, G, nb, chunk=4096):
        import torch
        self.torch, self.net, self.G, self.nb = torch, net, int(G), int(nb)
        dev = net.device
        nvox = self.nb * self.G ** 3
        self.chunk = max(int(chunk), (nvox + 65534) // 65535)
        self.act = torch.empty(6 * nvox * C, dtype=torch.bfloat16, device=dev)
        self.gact = None
        self.fin = None
        self.work = None
        self.pred = torch.empty(nvox, dtype=torch.float32, device=dev)
        self.dout = torch.empty(nvox, dtype=torch.float32, device=dev)
        self.loss = torch.empty(self.nb, dtype=torch.float64, device=dev)
        nlw = _lib.lib().dgfs_closure_loss_work_bytes(self.G, self.nb)
        self.lwork = torch.empty(nlw // 8 + 1, dtype=torch.float64, device=dev)

    def _train_buffers(self):
        t = self.torch
        if self.gact is None:
            dev = self.net.device
            nvox = self.nb * self.G ** 3
            self.gact = t.empty(2 * nvox * C, dtype=t.bfloat16, device=dev)
            self.fin = t.empty(nvox, dtype=t.float32, device=dev)
            nw = _lib.lib().dgfs_closure_work_bytes(self.G, self.nb, self.chunk)
            self.work = t.empty(nw // 4 + 1, dtype=t.float32, device=dev)

    def forward(self, u, src, scale, mode=0, dst=None, oscale=None, Q=None, stream=None):
        """mode 0: self.pred = net(scale * u[src]); mode 1: Q[dst] += oscale * net(...)."""
        net = self.net.cstruct()
        _lib.check(_lib.lib().dgfs_closure_forward(
            _lib.C.byref(net), _lib.dtype_code(u.dtype), self.G, self.nb, _lib.ptr(u),
            _lib.ptr(src), _lib.ptr(scale), _lib.ptr(self.act), int(mode),
            _lib.ptr(self.pred if mode == 0 else None), _lib.ptr(dst), _lib.ptr(oscale),
            _lib.ptr(Q), _lib.stream_ptr(stream)))
        return self.pred

    def backward(self, u, src, scale, dout, stream=None):
        self._train_buffers()
        net, gr = self.net.cstruct(), self.net.gstruct()
        _lib.check(_lib.lib().dgfs_closure_backward(
            _lib.C.byref(net), _lib.C.byref(gr), _lib.dtype_code(u.dtype), self.G, self.nb,
            _lib.ptr(u), _lib.ptr(src), _lib.ptr(scale), _lib.ptr(self.act), _lib.ptr(dout),
            _lib.ptr(self.gact), _lib.ptr(self.fin), self.chunk, _lib.ptr(self.work),
            _lib.stream_ptr(stream)))
        return self.net.g

    def loss_grad(self, target, ref=None, floor2=0.0, stream=None):
        """loss_b = |pred_b - target_b|^2 / max(|ref_b|^2, floor2) (ref = target when None)."""
        _lib.check(_lib.lib().dgfs_closure_loss(
            self.G, self.nb, _lib.ptr(self.pred), _lib.ptr(target), _lib.ptr(ref),
            float(floor2), _lib.ptr(self.dout), _lib.ptr(self.loss), _lib.ptr(self.lwork),
            _lib.stream_ptr(stream)))
        return self.loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        net = self.net
        net.step_count += 1
        t = net.step_count
        lr_t = lr * np.sqrt(1.0 - beta2 ** t) / (1.0 - beta1 ** t)
        for k in net.SHAPES:
            _lib.check(_lib.lib().dgfs_adam(
                net.p[k].numel(), _lib.ptr(net.p[k]), _lib.ptr(net.g[k]), _lib.ptr(net.m[k]),
                _lib.ptr(net.v[k]), float(lr_t), float(beta1), float(beta2), float(eps),
                _lib.stream_ptr(stream)))
        net.pack(stream)


def train(net, f, target, G, batch=16, epochs=10, lr=1e-3, seed=0, log=None, ref=None,
          floor2=0.0, rows=None):
    """Adam on the relative-L2 loss over (f, target) pairs: f (S, G^3) FP32
    already scaled to unit density, target (S, G^3) FP32 scaled residuals;
    the error of each pair is measured relative to ref (S, G^3) when given
    (the scaled collision integral), else to the target.  rows: the pair
    indices to train on (default all; no copy of the data is made).  Returns
    the per-epoch mean training loss."""
    import torch
    rows = np.arange(f.shape[0]) if rows is None else np.asarray(rows)
    S = len(rows)
    nb = min(batch, S)
    eng = ClosureEngine(net, G, nb)
    dev = net.device
    src = torch.arange(nb, dtype=torch.int32, device=dev)
    ones = torch.ones(nb, dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    fb = torch.empty((nb, G ** 3), dtype=torch.float32, device=dev)
    tb = torch.empty((nb, G ** 3), dtype=torch.float32, device=dev)
    rb = torch.empty((nb, G ** 3), dtype=torch.float32, device=dev) if ref is not None else None
    hist = []
    for ep in range(epochs):
        perm = torch.tensor(rows[rng.permutation(S)], device=dev)
        tot = torch.zeros((), dtype=torch.float64, device=dev)
        cnt = 0
        for i in range(0, S - nb + 1, nb):
            idx = perm[i:i + nb]
            fb.copy_(f.index_select(0, idx))
            tb.copy_(target.index_select(0, idx))
            if rb is not None:
                rb.copy_(ref.index_select(0, idx))
            eng.forward(fb, src, ones, mode=0)
            tot += eng.loss_grad(tb, rb, floor2).sum()
            eng.backward(fb, src, ones, eng.dout)
            eng.adam(lr)
            cnt += nb
        hist.append(float(tot) / max(cnt, 1))
        if log:
            log('epoch %d loss %.4e' % (ep, hist[-1]))
    return hist


class LearnedClosure:
    """Q = Q_Shakov + R_net + C per node, with the spectral operator as judge
    and fallback.

    C is each node's correction at its last spectral evaluation,
    C = Q_spectral - (Q_Shakov + R_net) (zero for spectral nodes); it is set for
    every node on the first call and refreshed, round robin, for a fraction
    ``sample`` of the learned nodes every ``check_every`` calls.  Before a
    node's correction is refreshed its prediction is judged against the
    spectral result: relative L2 error above ``tol`` (relative to
    max(|Q|, the network's |Q| floor)) hands the node to the spectral operator.
    The spectral nodes are sampled the same way (their own round robin) and
    keep a candidate correction from their previous sample: a node whose
    model plus candidate correction is within ``tol * readmit`` of the
    spectral result returns to the closure with a fresh correction.
    ``use_net=False`` keeps the same schedule without the network
    (Q_Shakov + C), the method the network is measured against.
    ``net_every = k > 1`` evaluates the network on the learned nodes every
    k-th call and reuses its last output R_net in between (Q_Shakov is
    evaluated every call; checks always use a fresh evaluation).  The first
    ``start`` calls use the spectral operator for every node (the initial
    transient, where the residual changes fastest).

    nodes: velocity slices s with tasks (s -> s) of a nodal basis.  shakov:
    relax.Shakov bound to the grid; spectral: collision.CollisionPlan;
    prefac: the spectral operator's prefactor (scales R)."""

    def __init__(self, net, G, nslice, shakov, spectral, prefac, cw, device, check_every=10,
                 sample=0.25, tol=0.05, seed=0, use_net=True, start=0, readmit=0.5,
                 net_every=1):
        import torch
        from .collision import task_table
        self.torch, self.net, self.G, self.ns = torch, net, int(G), int(nslice)
        self.shakov, self.spectral, self.prefac, self.cw = shakov, spectral, float(prefac), cw
        self.device, self.check_every, self.sample, self.tol = device, check_every, sample, tol
        self.use_net = use_net
        self.start = int(start)              # calls served by the spectral operator first
        self.task_table = task_table
        self.readmit = float(readmit)
        self.net_every = max(1, int(net_every))
        self.R = None                        # last network output on the learned nodes
        self._rstale = True
        self.learned = np.ones(self.ns, bool)
        self.calls = 0
        self.cursor = 0
        self.cursor_s = 0
        self.fallbacks = 0
        self.readmissions = 0
        self._qbuf = None
        self.checks = []                     # (call, {node: error}) history
        self._eng = {}
        self.C = None
        self.Cc = None                       # candidate corrections of spectral nodes
        self._tabs = {}
        self._rebuild()

    def _tab(self, nodes, dtype):
        key = (tuple(int(s) for s in nodes), dtype)
        if key not in self._tabs:
            self._tabs[key] = self.task_table(
                1, self.ns, [(int(s), [(int(s), 1.0)], [(int(s), 1.0)]) for s in nodes], dtype,
                self.device)
        return self._tabs[key]

    def _rebuild(self):
        idx = np.nonzero(self.learned)[0].astype(np.int32)
        spec = np.nonzero(~self.learned)[0].astype(np.int32)
        self.idx_np = idx
        self.idx = self.torch.tensor(idx, device=self.device)
        self.sh_tab = self._tab(idx, self.shakov.tdtype) if len(idx) else None
        self.sp_tab = self._tab(spec, self.spectral.tdtype) if len(spec) else None
        self.nl = len(idx)
        self._rstale = True

    def _engine(self, nb):
        if nb not in self._eng:
            self._eng[nb] = ClosureEngine(self.net, self.G, nb)
        return self._eng[nb]

    def _scales(self, u, idx):
        """1/rho and out_scale prefac rho^2 per listed slice (rho = cw sum f)."""
        n = self.G ** 3
        rho = u.view(self.ns, n).sum(1, dtype=self.torch.float64).index_select(0, idx) * self.cw
        osc = self.net.meta['out_scale'] * self.prefac * rho * rho
        return (1.0 / rho).float(), osc.float()

    def _model(self, u, Q, nodes, idx, tab):
        """Q += Q_Shakov + R_net on the listed nodes (no correction)."""
        self.shakov(u, Q, tasks=tab)
        if self.use_net:
            sc, osc = self._scales(u, idx)
            self._engine(len(nodes)).forward(u, idx, sc, mode=1, dst=idx, oscale=osc, Q=Q)

    def _spectral(self, u, Q, nodes):
        self.spectral(u, Q, tasks=self._tab(nodes, self.spectral.tdtype))

    def __call__(self, u, Q):
        """Q += collision term of every slice of u (shape (nslice * G^3,))."""
        t = self.torch
        self.calls += 1
        if self.calls <= self.start:
            self._spectral(u, Q, np.arange(self.ns, dtype=np.int32))
            return
        if self.C is None:                   # first call: corrections of every node
            self.C = t.zeros_like(u)
            allnodes = np.arange(self.ns, dtype=np.int32)
            qs = t.zeros_like(u)
            self._spectral(u, qs, allnodes)
            self._model(u, self.C, allnodes, t.tensor(allnodes, device=self.device),
                        self._tab(allnodes, self.shakov.tdtype))
            self.C.neg_().add_(qs)
        if self.nl and self.use_net and self.net_every > 1:
            self.shakov(u, Q, tasks=self.sh_tab)
            if self._rstale or (self.calls - self.start) % self.net_every == 0:
                if self.R is None:
                    self.R = t.zeros_like(u)
                self.R.zero_()
                sc, osc = self._scales(u, self.idx)
                self._engine(self.nl).forward(u, self.idx, sc, mode=1, dst=self.idx, oscale=osc,
                                              Q=self.R)
                self._rstale = False
            Q.add_(self.R)
        elif self.nl:
            self._model(u, Q, self.idx_np, self.idx, self.sh_tab)
        Q.add_(self.C)
        if self.sp_tab is not None:
            self.spectral(u, Q, tasks=self.sp_tab)
        if self.check_every and (self.calls - self.start) % self.check_every == 0 and self.nl:
            self.check(u)

    def _buffers(self, u):
        if self._qbuf is None or self._qbuf[0].numel() != u.numel():
            self._qbuf = (self.torch.empty_like(u), self.torch.empty_like(u))
        ql, qs = self._qbuf
        ql.zero_()
        qs.zero_()
        return ql, qs

    def _sample(self, cand, cursor):
        k = min(len(cand), max(1, int(round(self.sample * len(cand)))))
        return cand[(cursor + np.arange(k)) % len(cand)], (cursor + k) % max(1, len(cand))

    def check(self, u):
        """Judge and refresh a round-robin sample of the learned nodes; offer
        a round-robin sample of the spectral nodes their way back."""
        t = self.torch
        pl, ps = np.zeros(0, np.int64), np.zeros(0, np.int64)
        if self.nl:
            pl, self.cursor = self._sample(self.idx_np, self.cursor)
        spec = np.nonzero(~self.learned)[0]
        if len(spec) and self.readmit > 0:
            ps, self.cursor_s = self._sample(spec, self.cursor_s)
        pick = np.sort(np.concatenate([pl, ps])).astype(np.int32)
        if not len(pick):
            return np.zeros(0)
        idx = t.tensor(pick, device=self.device)
        n = self.G ** 3
        ql, qs = self._buffers(u)
        self._model(u, ql, pick, idx, self._tab(pick, self.shakov.tdtype))
        self._spectral(u, qs, pick)
        if self.Cc is None:
            self.Cc = t.zeros_like(self.C)
        Cv, Ccv = self.C.view(self.ns, n), self.Cc.view(self.ns, n)
        was = self.learned[pick]
        a = ql.view(self.ns, n).index_select(0, idx).double()
        b = qs.view(self.ns, n).index_select(0, idx).double()
        w = t.tensor(was, device=self.device).view(-1, 1)
        c = t.where(w, Cv.index_select(0, idx), Ccv.index_select(0, idx)).double()
        rho = u.view(self.ns, n).sum(1, dtype=t.float64).index_select(0, idx) * self.cw
        floor = self.net.meta['qfloor'] * self.net.meta['out_scale'] * self.prefac * rho * rho
        den = t.maximum(b.norm(dim=1), floor).clamp_min(1e-300)
        err = ((a + c - b).norm(dim=1) / den).cpu().numpy()
        self.checks.append((self.calls, dict(zip(pick.tolist(), err.tolist()))))
        bad = pick[was & (err > self.tol)]
        back = pick[~was & (err < self.tol * self.readmit)]
        fresh = (b - a).to(self.C.dtype)
        Cv.index_copy_(0, idx.long(), fresh)               # refreshed corrections
        Ccv.index_copy_(0, idx.long(), fresh)              # candidates for the next sample
        off = pick[~was & ~(err < self.tol * self.readmit)]
        off = np.concatenate([bad, off])
        if len(off):
            Cv.index_fill_(0, t.tensor(off, device=self.device).long(), 0.0)
        if len(bad) or len(back):
            self.learned[bad] = False
            self.learned[back] = True
            self.fallbacks += len(bad)
            self.readmissions += len(back)
            self._rebuild()
        return err
