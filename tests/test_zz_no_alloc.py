"""dgfs_project_conservative_noalloc (the projection on the library's own
fixed device workspace) against dgfs_project_conservative (caller workspace),
called through the C ABI: bit-identical outputs on two calls in a row, the
second smaller than the first (fewer slices, smaller grid), for FP64 and FP32
and both accumulate modes; the projected FP64 moments vanish; a request
larger than the fixed workspace is refused with status 2 before any device
work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(N, L=5.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    I, J, K = np.meshgrid(c0, c0, c0, indexing='ij')
    return np.stack([I.ravel(), J.ravel(), K.ravel()])


def _case(N, ns, td, dev, seed):
    import torch
    cv = _grid(N)
    np_dt = np.float64 if td == torch.float64 else np.float32
    cvd = cv.astype(np_dt).astype(np.float64)
    phi = np.vstack([np.ones(cv.shape[1]), cvd, (cvd * cvd).sum(0)])
    gram_inv = np.ascontiguousarray(np.linalg.inv(phi @ phi.T))
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((ns, N ** 3)) * np.exp(-(cv * cv).sum(0) / 1.5)[None, :]
    Q += 0.2 * rng.standard_normal((ns, 1))
    base = rng.standard_normal((ns, N ** 3))
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=td, device=dev)
    return dict(cx=t(cv[0]), cy=t(cv[1]), cz=t(cv[2]), gram_inv=gram_inv, Q=t(Q),
                base=t(base), phi=phi, N=N, ns=ns)


def _call(c, out, accumulate, noalloc):
    import torch
    from dgfs1d_gpu_amd import _lib
    L = _lib.lib()
    code = _lib.dtype_code(c['Q'].dtype)
    gp = c['gram_inv'].ctypes.data_as(C.POINTER(C.c_double))
    args = (code, c['ns'], c['N'] ** 3, _lib.ptr(c['cx']), _lib.ptr(c['cy']),
            _lib.ptr(c['cz']), gp, _lib.ptr(c['Q']), _lib.ptr(out), int(accumulate))
    stream = _lib.stream_ptr(None)
    if noalloc:
        st = L.dgfs_project_conservative_noalloc(*args, stream)
    else:
        nb = L.dgfs_projection_work_bytes(c['ns'], c['N'] ** 3)
        work = torch.empty(max(1, nb // 8), dtype=torch.float64, device=out.device)
        st = L.dgfs_project_conservative(*args, _lib.ptr(work), stream)
    assert st == 0, L.dgfs_last_error()
    torch.cuda.synchronize()


@pytest.mark.parametrize('dtype', ['f64', 'f32'])
def test_noalloc_projection_matches_workspace_entry_point(cuda, dtype):
    import torch
    td = torch.float64 if dtype == 'f64' else torch.float32
    # two calls in a row, the second smaller: 12 slices of 32^3, then 3 of 16^3
    for N, ns, seed in ((32, 12, 1), (16, 3, 2)):
        c = _case(N, ns, td, cuda, seed)
        print('noalloc projection', dtype, 'N', N, 'nslice', ns, flush=True)
        for accumulate in (1, 0):
            ref = c['base'].clone() if accumulate else torch.empty_like(c['Q'])
            got = c['base'].clone() if accumulate else torch.empty_like(c['Q'])
            _call(c, ref, accumulate, noalloc=False)
            _call(c, got, accumulate, noalloc=True)
            assert torch.equal(ref, got), (dtype, N, accumulate)
        if dtype == 'f64':
            qc = got.cpu().numpy().reshape(ns, -1)
            m = np.abs(c['phi'] @ qc.T) / (np.abs(c['phi']) @ np.abs(qc).T)
            assert m.max() < 1e-12


def test_noalloc_projection_refuses_oversized_request(cuda):
    from dgfs1d_gpu_amd import _lib
    L = _lib.lib()
    one = (C.c_double * 25)()
    n48 = 48 ** 3
    cap = (4 << 20) // L.dgfs_projection_work_bytes(1, n48)
    assert L.dgfs_projection_work_bytes(cap, n48) <= 4 << 20
    st = L.dgfs_project_conservative_noalloc(0, cap + 1, n48, 1, 1, 1, one, 1, 1, 0, None)
    assert st == 2
    assert b'no-allocation projection holds' in L.dgfs_last_error()
