"""Host wrappers of the transport-side kernels (libdgfs_hip.so).

All functions take contiguous torch tensors on the HIP device and enqueue on
the current torch stream; none of them synchronises the host except
``Residual.__call__`` and ``Moments.__call__``, which return host values."""
import numpy as np

from . import _lib


def _dev_f64(x, device):
    import torch
    return torch.tensor(np.ascontiguousarray(x, dtype=np.float64).ravel(), dtype=torch.float64,
                        device=device)


class MatOp:
    """out = A @ in over the leading axis (GiMMiK replacement, util.py:20-25)."""

    def __init__(self, A, device):
        A = np.atleast_2d(np.asarray(A, dtype=np.float64))
        self.Kout, self.Kin = A.shape
        self.is_identity = self.Kout == self.Kin and np.array_equal(A, np.eye(self.Kin))
        self.A = _dev_f64(A, device)

    def __call__(self, x, out, accumulate=False, stream=None):
        ncol = x.numel() // self.Kin
        if x.numel() != ncol * self.Kin or out.numel() != ncol * self.Kout:
            raise ValueError('matop: operand sizes do not match the operator')
        _lib.check(_lib.lib().dgfs_matop(
            _lib.dtype_code(x.dtype), self.Kout, self.Kin, ncol, _lib.ptr(self.A), _lib.ptr(x),
            _lib.ptr(out), int(accumulate), _lib.stream_ptr(stream)))


def axnpby(x0, terms, a0, stream=None):
    """x0 = a0 x0 + sum a_i x_i for (a_i, x_i) in terms (<= 3 terms)."""
    if len(terms) > 3:
        raise ValueError('axnpby takes at most 3 extra operands')
    xs = [_lib.ptr(x) for _, x in terms] + [_lib.C.c_void_p(None)] * (3 - len(terms))
    ks = [float(a) for a, _ in terms] + [0.0] * (3 - len(terms))
    for _, x in terms:
        if x.numel() != x0.numel() or x.dtype != x0.dtype:
            raise ValueError('axnpby: operand mismatch')
    _lib.check(_lib.lib().dgfs_axnpby(
        _lib.dtype_code(x0.dtype), x0.numel(), 1 + len(terms), _lib.ptr(x0), xs[0], xs[1], xs[2],
        float(a0), ks[0], ks[1], ks[2], _lib.stream_ptr(stream)))


def axnpby_compensated(x0, x0_lo, terms, a0, stream=None):
    """(x0, x0_lo) = a0 (x0 + x0_lo) + sum a_i (x_i + lo_i) for (a_i, x_i, lo_i)
    in terms (<= 3 terms, FP32, lo_i / x0_lo may be None)."""
    import torch
    if len(terms) > 3:
        raise ValueError('axnpby takes at most 3 extra operands')
    for t in [x0, x0_lo] + [x for _, x, _ in terms] + [lo for _, _, lo in terms]:
        if t is not None and (t.dtype != torch.float32 or t.numel() != x0.numel()):
            raise ValueError('axnpby_compensated: FP32 operands of equal size required')
    null = _lib.C.c_void_p(None)
    xs, ls = [], []
    for _, x, lo in terms:
        xs.append(_lib.ptr(x))
        ls.append(_lib.ptr(lo))
    xs += [null] * (3 - len(terms))
    ls += [null] * (3 - len(terms))
    ks = [float(a) for a, _, _ in terms] + [0.0] * (3 - len(terms))
    _lib.check(_lib.lib().dgfs_axnpby_compensated(
        x0.numel(), 1 + len(terms), _lib.ptr(x0), _lib.ptr(x0_lo), xs[0], ls[0], xs[1], ls[1],
        xs[2], ls[2], float(a0), ks[0], ks[1], ks[2], _lib.stream_ptr(stream)))


class Transport:
    """Fused DG transport operator (std.py:283-348) bound to one basis/mesh."""

    def __init__(self, basis, invjac, vm, left, right, dtype, device, upwind=None,
                 entropy=False):
        """upwind = (Dp, Dm), the upwind summation-by-parts operators in the
        'ux = A u' orientation: selects the split form of the upwind driver;
        entropy=True selects the entropy-stable volume term (basis derivMat)."""
        import torch
        self.torch = torch
        self.K, self.Ne, self.Nv = basis.K, len(invjac), vm.vsize()
        self.device = device
        self.dtype = dtype
        D = basis.mat_deriv if upwind is None else upwind[0]
        keep = dict(D=_dev_f64(D, device), BL=_dev_f64(basis.BL, device),
                    BR=_dev_f64(basis.BR, device), gLD=_dev_f64(basis.gLD, device),
                    gRD=_dev_f64(basis.gRD, device), invjac=_dev_f64(invjac, device))
        nbytes = _lib.lib().dgfs_transport_work_bytes(self.Nv)
        keep['work'] = torch.zeros(nbytes // 8 + 1, dtype=torch.float64, device=device)
        keep['nw'] = torch.zeros(2, dtype=torch.float64, device=device)
        self.keep = keep
        self.cx, self.cy, self.cz = vm.d_cvx(), vm.d_cvy(), vm.d_cvz()
        self.cw = vm.cw()
        self.left, self.right = left, right
        tp = _lib.Transport()
        tp.dtype = _lib.dtype_code(dtype)
        tp.K, tp.Ne, tp.Nv = self.K, self.Ne, self.Nv
        if upwind is not None:
            keep['Dm'] = _dev_f64(upwind[1], device)
        for name in ('D', 'BL', 'BR', 'gLD', 'gRD', 'invjac', 'work', 'nw'):
            setattr(tp, name, keep[name].data_ptr())
        tp.Dm = keep['Dm'].data_ptr() if upwind is not None else None
        if entropy:
            keep['De'] = _dev_f64(basis.derivMat, device)
        tp.De = keep['De'].data_ptr() if entropy else None
        tp.cx, tp.cy, tp.cz = self.cx.data_ptr(), self.cy.data_ptr(), self.cz.data_ptr()
        tp.cw = float(self.cw)
        self.tp = tp

    @staticmethod
    def _fill(w, p):
        w.kind, w.nl, w.ushift = p['kind'], p['nl'], p.get('ushift', 0.0)
        f0 = p.get('f0')
        w.f0 = f0.data_ptr() if f0 is not None else None
        w.ux, w.uy, w.uz, w.T = p.get('ux', 0.0), p.get('uy', 0.0), p.get('uz', 0.0), p.get('T', 1.0)

    def __call__(self, t, u, ux, stream=None):
        self._fill(self.tp.left, self.left.wall_params(t))
        self._fill(self.tp.right, self.right.wall_params(t))
        _lib.check(_lib.lib().dgfs_transport_rhs(_lib.C.byref(self.tp), _lib.ptr(u), _lib.ptr(ux),
                                                 float(t), _lib.stream_ptr(stream)))

    def wall_density(self):
        """n_w of the last call (left, right) -- host sync."""
        return self.keep['nw'].cpu().numpy()


class BatchTransport:
    """The transport of many independent problems of one dtype in three
    launches (dgfs_transport_batch_*): ``transports`` are the problems'
    Transport objects (their walls evaluated at t = 0; expression walls are
    refused), ``offsets`` the start of each problem's state in the flat
    arrays the call receives."""

    def __init__(self, transports, offsets):
        n = len(transports)
        if n == 0 or len(offsets) != n:
            raise ValueError('one offset per transport')
        arr = (_lib.Transport * n)()
        for i, tr in enumerate(transports):
            tr._fill(tr.tp.left, tr.left.wall_params(0.0))
            tr._fill(tr.tp.right, tr.right.wall_params(0.0))
            arr[i] = tr.tp
        off = (_lib.C.c_long * n)(*[int(o) for o in offsets])
        self._keep = transports                  # their device tables must outlive the batch
        self._h = _lib.C.c_void_p()
        _lib.check(_lib.lib().dgfs_transport_batch_create(arr, n, off, _lib.C.byref(self._h)))
        self.n = n

    def __call__(self, u, ux, stream=None):
        _lib.check(_lib.lib().dgfs_transport_batch_rhs(self._h, _lib.ptr(u), _lib.ptr(ux),
                                                       _lib.stream_ptr(stream)))

    def wall_densities(self):
        """(nprob, 2) n_w of the last call -- host sync."""
        out = (_lib.C.c_double * (2 * self.n))()
        _lib.check(_lib.lib().dgfs_transport_batch_nw(
            self._h, _lib.C.cast(out, _lib.C.POINTER(_lib.C.c_double))))
        return np.array(out[:]).reshape(self.n, 2)

    def close(self):
        if self._h:
            _lib.lib().dgfs_transport_batch_destroy(self._h)
            self._h = _lib.C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Limiter:
    """Linear minmod limiter (dgfs_limit_linear; reference abstd.py:402-416):
    cell means and the linear slope by the basis' operators, then the
    per-element limiting."""

    def __init__(self, basis, xsol, invjac, Nv, dtype, device):
        import torch
        A, DrLin = basis.limiter_ops()
        self.avg, self.drlin = MatOp(A, device), MatOp(DrLin, device)
        K, Ne = basis.K, len(invjac)
        self.K, self.Ne, self.Nv = K, Ne, int(Nv)
        invjac = np.asarray(invjac, dtype=np.float64).ravel()
        h = 2.0 / invjac                                    # limiters.mako:37
        xsol = np.asarray(xsol, dtype=np.float64)           # (K, Ne) node coordinates
        x0 = xsol[0] + h / 2                                # :38
        # node offsets from the cell centre and element widths: FP64 device arrays
        self.xrel = _dev_f64((xsol - x0[None, :]).ravel(), device)
        self.h = _dev_f64(h, device)
        n = K * Ne * self.Nv
        self.uavg = torch.zeros(n, dtype=dtype, device=device)
        self.ulx = torch.zeros(n, dtype=dtype, device=device)
        self.invjac = torch.tensor(invjac, dtype=dtype, device=device).view(1, Ne, 1)

    def __call__(self, u_in, u_out, stream=None):
        self.avg(u_in, self.uavg, stream=stream)
        self.drlin(u_in, self.ulx, stream=stream)
        self.ulx.view(self.K, self.Ne, self.Nv).mul_(self.invjac)     # mul_by_invjac
        _lib.check(_lib.lib().dgfs_limit_linear(
            _lib.dtype_code(u_in.dtype), self.K, self.Ne, self.Nv, _lib.ptr(self.xrel),
            _lib.ptr(self.h),
            _lib.ptr(u_in), _lib.ptr(self.uavg), _lib.ptr(self.ulx), _lib.ptr(u_out),
            _lib.stream_ptr(stream)))


class PositivityLimiter:
    """Positivity limiter of the upwind driver (dgfs_limit_positivity;
    reference abstdu.py:432-442): cell means by the basis' operator."""

    def __init__(self, basis, Ne, Nv, dtype, device):
        import torch
        A, _ = basis.limiter_ops()
        self.avg = MatOp(A, device)
        self.K, self.Ne, self.Nv = basis.K, int(Ne), int(Nv)
        self.uavg = torch.zeros(self.K * self.Ne * self.Nv, dtype=dtype, device=device)

    def __call__(self, u_in, u_out, stream=None):
        self.avg(u_in, self.uavg, stream=stream)
        _lib.check(_lib.lib().dgfs_limit_positivity(
            _lib.dtype_code(u_in.dtype), self.K, self.Ne, self.Nv, _lib.ptr(u_in),
            _lib.ptr(self.uavg), _lib.ptr(u_out), _lib.stream_ptr(stream)))


class Residual:
    def __init__(self, device):
        import torch
        self.work = torch.zeros(_lib.lib().dgfs_residual_work_bytes() // 8 + 1,
                                dtype=torch.float64, device=device)
        self.out = torch.zeros(2, dtype=torch.float64, device=device)

    def sums(self, cur, prev, stream=None):
        _lib.check(_lib.lib().dgfs_residual(
            _lib.dtype_code(cur.dtype), cur.numel(), _lib.ptr(cur), _lib.ptr(prev),
            _lib.ptr(self.work), _lib.ptr(self.out), _lib.stream_ptr(stream)))
        return self.out

    def __call__(self, cur, prev):
        s = self.sums(cur, prev).cpu().numpy()
        return float(np.sqrt(s[0] / s[1]))


class Moments:
    """FP64 moments at the reconstruction points (moments.py:11-62)."""

    def __init__(self, interp, K, Ne, vm, device):
        import torch
        self.K, self.Ne, self.Nv = K, Ne, vm.vsize()
        self.Nqr = interp.shape[1]
        self.im = _dev_f64(interp, device)
        self.nnode = self.Nqr * Ne
        nb = _lib.lib().dgfs_moments_work_bytes(self.Nv, self.nnode)
        self.work = torch.zeros(nb // 8 + 1, dtype=torch.float64, device=device)
        self.out = torch.zeros(self.nnode * 9, dtype=torch.float64, device=device)
        self.vm = vm

    def compute(self, u, stream=None):
        vm = self.vm
        if u.dtype != vm.d_cvx().dtype:
            raise TypeError('moments: state dtype %s differs from the velocity grid dtype %s'
                            % (u.dtype, vm.d_cvx().dtype))
        _lib.check(_lib.lib().dgfs_moments(
            _lib.dtype_code(u.dtype), self.K, self.Ne, self.Nqr, self.Nv, _lib.ptr(self.im),
            _lib.ptr(u), _lib.ptr(vm.d_cvx()), _lib.ptr(vm.d_cvy()), _lib.ptr(vm.d_cvz()),
            float(vm.cw()), _lib.ptr(self.work), _lib.ptr(self.out), _lib.stream_ptr(stream)))
        return self.out

    def __call__(self, u):
        """(Nqr, Ne, 9) non-dimensional {rho, ux, uy, T, qx, qy, Pxx, Pyy, Pxy}."""
        return self.compute(u).cpu().numpy().reshape(self.Nqr, self.Ne, 9)
