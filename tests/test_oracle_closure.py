"""Pins of the closure oracle (CPU): the convolution against scipy.ndimage's
correlation, the analytic backward pass against central finite differences
(FP64, no bf16 rounding), the relative-L2 loss gradient, and the bf16
rounding helper against torch's bfloat16 conversion."""
import numpy as np
import pytest

from oracle import closure as OCL


def _net(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    C = OCL.C
    return OCL.Net(rng.standard_normal((27, C, 1)) * scale, rng.standard_normal(C) * 0.1,
                   rng.standard_normal((27, C, C)) * scale / 6, rng.standard_normal(C) * 0.1,
                   rng.standard_normal((27, C, C)) * scale / 6, rng.standard_normal(C) * 0.1,
                   rng.standard_normal((27, 1, C)) * scale / 6, rng.standard_normal(1) * 0.1)


def test_conv_matches_ndimage_correlate():
    from scipy import ndimage
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 5, 5, 3))
    W = rng.standard_normal((27, 4, 3))
    out = OCL.conv(x, W)
    for b in range(2):
        for o in range(4):
            ref = sum(ndimage.correlate(x[b, ..., i], W[:, o, i].reshape(3, 3, 3),
                                        mode='constant', cval=0.0) for i in range(3))
            assert np.abs(out[b, ..., o] - ref).max() < 1e-12


def test_conv_t_is_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 4, 4, 3))
    g = rng.standard_normal((1, 4, 4, 4, 5))
    W = rng.standard_normal((27, 5, 3))
    assert abs((OCL.conv(x, W) * g).sum() - (x * OCL.conv_t(g, W)).sum()) < 1e-10


def test_backward_finite_differences():
    net = _net()
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 1, (2, 4, 4, 4))
    dout = rng.standard_normal((2, 4, 4, 4))
    out, cache = net.forward(f)
    g = net.backward(dout, cache)
    L = lambda: (net.forward(f)[0] * dout).sum()
    eps = 1e-6
    for name, idx in [('W1', (13, 5, 0)), ('b1', (7,)), ('W2', (3, 4, 11)), ('b2', (30,)),
                      ('W3', (26, 0, 31)), ('b3', (2,)), ('W4', (9, 0, 17)), ('b4', (0,))]:
        P = getattr(net, name)
        old = P[idx]
        P[idx] = old + eps
        lp = L()
        P[idx] = old - eps
        lm = L()
        P[idx] = old
        fd = (lp - lm) / (2 * eps)
        an = g[name][idx]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (name, fd, an)


def test_loss_gradient():
    rng = np.random.default_rng(4)
    p, t = rng.standard_normal((3, 2, 2, 2)), rng.standard_normal((3, 2, 2, 2))
    loss, grad = OCL.rel_loss(p, t)
    eps = 1e-7
    p2 = p.copy()
    p2[1, 0, 1, 1] += eps
    fd = (OCL.rel_loss(p2, t)[0].mean() - loss.mean()) / eps
    assert abs(fd - grad[1, 0, 1, 1]) < 1e-5 * abs(grad[1, 0, 1, 1])


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(5).standard_normal(10000) * 10.0 ** np.random.default_rng(6)\
        .uniform(-8, 8, 10000)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OCL.bf16(x), ref)
