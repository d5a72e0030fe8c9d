"""Oracle: the learned-closure network, forward and backward (NumPy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

No reference counterpart (the closure's targets are the spectral operator
minus dgfs1D/std/scattering.py:1636-1833).  The network: conv3x3x3 1->32,
SiLU, 32->32, SiLU, 32->32, SiLU, 32->1 with zero padding on a G^3 grid,
tap = (dz+1)*9 + (dy+1)*3 + (dx+1), out[v] = sum_tap W[tap] . in[v + d(tap)].
Arrays are channels-last: act[b, z, y, x, c].

``bf16=True`` rounds to bfloat16 (round-to-nearest-even) exactly where the
HIP kernels round to bf16: the scaled input f and every forward weight (the
matrix cores take bf16 operands), every pre-activation Z and activation Y,
and the back-propagated dZ; arithmetic is FP64 in between.  The backward
uses the FP32 master W4 (the kernels' VALU path), as here.
``bf16=False`` is the exact FP64 network (pinned by finite differences).
"""
import numpy as np

C = 32
TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16(x):
    """Round to bfloat16 (nearest-even) and return as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_d(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv(x, W):
    """x (B, G, G, G, Ci), W (27, Co, Ci) -> (B, G, G, G, Co)."""
    G = x.shape[1]
    xp = _pad(x)
    out = 0.0
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = xp[:, 1 + dz:1 + dz + G, 1 + dy:1 + dy + G, 1 + dx:1 + dx + G, :]
        out = out + np.einsum('bzyxi,oi->bzyxo', sl, W[t])
    return out


def conv_t(g, W):
    """Adjoint of conv: g (B, G, G, G, Co) -> (B, G, G, G, Ci)."""
    G = g.shape[1]
    gp = _pad(g)
    out = 0.0
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = gp[:, 1 - dz:1 - dz + G, 1 - dy:1 - dy + G, 1 - dx:1 - dx + G, :]
        out = out + np.einsum('bzyxo,oi->bzyxi', sl, W[t])
    return out


def wgrad(g, x):
    """dW[t, o, i] = sum g[v, o] x[v + d(t), i]."""
    G = x.shape[1]
    xp = _pad(x)
    dW = np.zeros((27, g.shape[-1], x.shape[-1]))
    for t, (dz, dy, dx) in enumerate(TAPS):
        sl = xp[:, 1 + dz:1 + dz + G, 1 + dy:1 + dy + G, 1 + dx:1 + dx + G, :]
        dW[t] = np.einsum('bzyxo,bzyxi->oi', g, sl)
    return dW


class Net:
    """Weights as FP64 arrays: W1 (27, 32, 1), W2/W3 (27, 32, 32), W4 (27, 1, 32)."""

    def __init__(self, W1, b1, W2, b2, W3, b3, W4, b4):
        self.W1, self.b1 = np.asarray(W1, float).reshape(27, C, 1), np.asarray(b1, float)
        self.W2, self.b2 = np.asarray(W2, float).reshape(27, C, C), np.asarray(b2, float)
        self.W3, self.b3 = np.asarray(W3, float).reshape(27, C, C), np.asarray(b3, float)
        self.W4, self.b4 = np.asarray(W4, float).reshape(27, 1, C), np.asarray(b4, float)

    def forward(self, f, bf=False):
        """f (B, G, G, G) already scaled -> prediction (B, G, G, G), cache."""
        r = bf16 if bf else (lambda a: a)
        x0 = r(np.asarray(f, float))[..., None]
        W2 = r(self.W2)
        W3 = r(self.W3)
        z1 = conv(x0, r(self.W1)) + self.b1
        Z1, Y1 = r(z1), r(silu(z1))
        z2 = conv(Y1, W2) + self.b2
        Z2, Y2 = r(z2), r(silu(z2))
        z3 = conv(Y2, W3) + self.b3
        Z3, Y3 = r(z3), r(silu(z3))
        out = conv(Y3, r(self.W4))[..., 0] + self.b4[0]
        return out, dict(x0=x0, Z1=Z1, Y1=Y1, Z2=Z2, Y2=Y2, Z3=Z3, Y3=Y3, W2=W2, W3=W3)

    def backward(self, dout, cache, bf=False):
        """Gradients of sum(dout * out) w.r.t. all parameters."""
        r = bf16 if bf else (lambda a: a)
        c = cache
        g4 = np.asarray(dout, float)[..., None]
        dW4 = wgrad(g4, c['Y3'])
        db4 = np.array([g4.sum()])
        dZ3 = r(conv_t(g4, self.W4) * silu_d(c['Z3']))
        dW3 = wgrad(dZ3, c['Y2'])
        db3 = dZ3.sum(axis=(0, 1, 2, 3))
        dZ2 = r(conv_t(dZ3, c['W3']) * silu_d(c['Z2']))
        dW2 = wgrad(dZ2, c['Y1'])
        db2 = dZ2.sum(axis=(0, 1, 2, 3))
        dZ1 = r(conv_t(dZ2, c['W2']) * silu_d(c['Z1']))
        dW1 = wgrad(dZ1, c['x0'])
        db1 = dZ1.sum(axis=(0, 1, 2, 3))
        return dict(W1=dW1, b1=db1, W2=dW2, b2=db2, W3=dW3, b3=db3, W4=dW4, b4=db4)


def rel_loss(pred, target):
    """Mean over slices of |p - t|^2 / |t|^2 and its gradient."""
    B = pred.shape[0]
    d = (pred - target).reshape(B, -1)
    s = (target.reshape(B, -1) ** 2).sum(1)
    loss = (d * d).sum(1) / s
    grad = (2.0 * d / (s[:, None] * B)).reshape(pred.shape)
    return loss, grad
