"""Oracle: fast-spectral VHS collision operator Q(f, g) (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, step for step, the reference pipeline:
  * frequency vectors l (std/scattering.py:118-128)
  * a[p,q,k] = (pi/L) qz_p/2 (sz_q . l_k)          (vhs-gll.mako:43-67)
  * bb1, bb2 with sinc(s) = sin(s+1e-15)/(s+1e-15)  (vhs-gll.mako:9-12, :69-99)
  * type 'vhs'     fs(): std/scattering.py:346-429 with kernels
                   cosMul/sinMul/cplxMul/cplxMulAdd/computeQG/ax/output_append_
                   (vhs-gll.mako:101-433)
  * type 'vhs-gll' fs(): std/scattering.py:251-313 (f = g, complex squares
                   t3^2 + t4^2, vhs-gll.mako:140-167)
cuFFT conventions: forward e^{-i}, inverse e^{+i}, both unnormalised
(cufft.py:182-183); the explicit 1/n is applied in cosSinMul/cosMul.

``q_bruteforce`` evaluates the same truncated, aliased spectral sum directly,
with no FFT, and pins the pipeline (tests/test_oracle_collision.py).
"""
from math import gamma as _gamma

import numpy as np

EPS64 = 1e-15


def freq(N):
    """Signed FFT frequencies in index order (std/scattering.py:118)."""
    return np.concatenate((np.arange(0, N // 2), np.arange(-N // 2, 0))).astype(np.float64)


def lvec(N):
    """(3, N^3) integer frequency vectors, v = I*N^2 + J*N + K (:121-128)."""
    l0 = freq(N)
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing='ij')
    return np.stack([l0[I.ravel()], l0[J.ravel()], l0[K.ravel()]])


def sinc(s, eps=EPS64):
    return np.sin(s + eps) / (s + eps)


def vhs_prefactor(H0, n0, T0, omega, dRef, Tref):
    """prefac = invKn*alpha/(2^{2-omega+alpha} Gamma(2.5-omega) pi), alpha=1
    (std/scattering.py:70-83)."""
    alpha = 1.0
    invKn = H0 * np.sqrt(2.0) * np.pi * n0 * dRef * dRef * (Tref / T0) ** (omega - 0.5)
    return invKn * alpha / (2.0 ** (2.0 - omega + alpha) * _gamma(2.5 - omega) * np.pi), invKn


class Precomp:
    """a, bb1, bb2 of perform_precomputation (std/scattering.py:106-191)."""

    def __init__(self, N, L, qz, qw, sz, sw, gamma):
        self.N = N
        self.n = N ** 3
        self.L, self.qz, self.qw = L, np.asarray(qz), np.asarray(qw)
        self.sz, self.sw, self.gamma = np.asarray(sz), sw, gamma
        self.M, self.Nrho = self.sz.shape[0], len(self.qz)
        l = lvec(N)
        self.l = l
        fac = np.pi / L * self.qz / 2.0                              # mako:57
        sl = self.sz @ l                                             # (M, n)
        self.a = fac[:, None, None] * sl[None, :, :]                 # (Nrho, M, n)
        lnorm = np.sqrt((l * l).sum(0))                              # mako:84-86
        zg = self.qz ** (gamma + 2.0)
        self.bb1 = (zg * 4.0 * np.pi)[:, None] * sinc((np.pi / L * self.qz / 2.0)[:, None] * lnorm)
        self.bb2 = ((self.qw * zg * 16.0 * np.pi * np.pi)[:, None]
                    * sinc((np.pi / L * self.qz)[:, None] * lnorm)).sum(0)

    def weight(self):
        """computeQG weight intw_p = 2 qw_p sw / n * bb1[p] (mako:190-201)."""
        return (2.0 * self.qw * self.sw / self.n)[:, None] * self.bb1


# FFT backend: numpy (default) or scipy.fft with a worker count (used for the
# timed CPU baseline, where all host cores should be busy).
FFT_WORKERS = None


def set_fft_workers(workers):
    global FFT_WORKERS
    FFT_WORKERS = workers


def _fft(x, N):
    shp = x.shape[:-1] + (N, N, N)
    if FFT_WORKERS:
        import scipy.fft as sf
        return sf.fftn(x.reshape(shp), axes=(-3, -2, -1), workers=FFT_WORKERS).reshape(x.shape)
    return np.fft.fftn(x.reshape(shp), axes=(-3, -2, -1)).reshape(x.shape)


def _ifft_u(x, N):
    """Unnormalised inverse transform (cuFFT CUFFT_INVERSE)."""
    n = N ** 3
    shp = x.shape[:-1] + (N, N, N)
    if FFT_WORKERS:
        import scipy.fft as sf
        return n * sf.ifftn(x.reshape(shp), axes=(-3, -2, -1),
                            workers=FFT_WORKERS).reshape(x.shape)
    return n * np.fft.ifftn(x.reshape(shp), axes=(-3, -2, -1)).reshape(x.shape)


def q_vhs(pre, f, g, prefac):
    """One node, type 'vhs' (std/scattering.py:346-429). Returns the increment
    added to Q[modeout, elem]."""
    N, n = pre.N, pre.n
    FTf = _fft(f.astype(np.complex128), N) / n                       # cosMul scale
    FTg = _fft(g.astype(np.complex128), N) / n
    a = pre.a.reshape(-1, n)
    ca, sa = np.cos(a), np.sin(a)
    t3, t4 = _ifft_u(ca * FTf, N), _ifft_u(ca * FTg, N)
    t5 = t3 * t4                                                     # cplxMul
    t3, t4 = _ifft_u(sa * FTf, N), _ifft_u(sa * FTg, N)
    t5 = t5 + t3 * t4                                                # cplxMulAdd
    t1 = _fft(t5, N).reshape(pre.Nrho, pre.M, n)
    fC = (pre.weight()[:, None, :] * t1).sum(axis=(0, 1))            # computeQG
    QG = _ifft_u(fC, N)
    fC2 = _ifft_u(pre.bb2 * FTg, N)                                  # ax + inverse
    return prefac * (QG.real - f * fC2.real)                         # output_append_


def q_vhs_gll(pre, f, prefac):
    """One node, type 'vhs-gll' (std/scattering.py:251-313): f = g and the
    product is the complex square t3^2 + t4^2 (mako:140-167)."""
    N, n = pre.N, pre.n
    FTf = _fft(f.astype(np.complex128), N) / n
    a = pre.a.reshape(-1, n)
    t3 = _ifft_u(np.cos(a) * FTf, N)
    t4 = _ifft_u(np.sin(a) * FTf, N)
    t2 = t3 * t3 + t4 * t4
    t1 = _fft(t2, N).reshape(pre.Nrho, pre.M, n)
    fC = (pre.weight()[:, None, :] * t1).sum(axis=(0, 1))
    QG = _ifft_u(fC, N)
    fC2 = _ifft_u(pre.bb2 * FTf, N)
    return prefac * (QG.real - f * fC2.real)


def loss_field(pre, g):
    """L[g] = Re IDFT_u(bb2 * DFT(g) / n): the field that multiplies f in the
    loss term of output_append_ (vhs-gll.mako ax kernel and the inverse
    transform, std/scattering.py:418-425)."""
    N, n = pre.N, pre.n
    return _ifft_u(pre.bb2 * _fft(g.astype(np.complex128), N) / n, N).real


def q_reference(pre, f, f0, Q0, prefac):
    """Q(f, f) in the reference-state form used by the FP32 product path
    (no reference counterpart):

        Q0 + prefac (G(h, s) - h L[f] - f0 L[h]),   h = f - f0,  s = f + f0,

    equal to q_vhs(f, f) when Q0 = q_vhs(f0, f0), because the gain G of
    q_vhs is a symmetric bilinear form (its spectral kernel cos(a(l) - a(m))
    is even) and L is linear.  G(h, s) is taken from q_vhs(h, s) =
    prefac (G(h, s) - h L[s]) and L[s] - L[f0] = L[f]."""
    h, s = f - f0, f + f0
    return Q0 + q_vhs(pre, h, s, prefac) + prefac * (h * loss_field(pre, f0)
                                                    - f0 * loss_field(pre, h))


def q_bruteforce(pre, f, g, prefac):
    """The same operator as a direct spectral sum (no FFT of any product):

        G(k) = sum_p w_p bb1[p,k] sum_q (1/n) sum_{l+m = k mod N}
                   cos(a_pq(l) - a_pq(m)) F(l) G(m)
        Q    = prefac * (Re IDFT_u(G) - f * Re IDFT(bb2 * FFT(g)))

    with F = DFT(f), Gg = DFT(g) evaluated by explicit sums.  O(n^2 B): use
    only at N <= 8.  Derivation: IDFT_u(A).IDFT_u(B) has DFT n (A*B)(k), where
    A = cos(a) F/n etc., so t1t2 + t3t4 -> (1/n) sum cos(a(l) - a(m)) F(l) G(m).
    """
    N, n = pre.N, pre.n
    l = pre.l
    idx = np.arange(N)
    I, J, K = np.meshgrid(idx, idx, idx, indexing='ij')
    X = np.stack([I.ravel(), J.ravel(), K.ravel()]).astype(np.float64)   # (3, n)
    # explicit DFT matrices (forward e^{-i}, inverse e^{+i})
    phase = 2.0 * np.pi / N * (l.T @ X)                                  # (n_l, n_x)
    Wf = np.exp(-1j * phase)                                             # F(l) = sum_x f(x) Wf[l,x]
    Fl = Wf @ f
    Gl = Wf @ g
    # index of l+m mod N for all pairs
    li = np.stack([I.ravel(), J.ravel(), K.ravel()])                     # integer index per l
    kk = ((li[:, :, None] + li[:, None, :]) % N)
    kidx = (kk[0] * N + kk[1]) * N + kk[2]                               # (n, n)
    w = pre.weight()                                                     # (Nrho, n)
    Ghat = np.zeros(n, dtype=np.complex128)
    FG = Fl[:, None] * Gl[None, :]
    for p in range(pre.Nrho):
        S = np.zeros((n, n))
        for q in range(pre.M):
            ap = pre.a[p, q]
            S += np.cos(ap[:, None] - ap[None, :])
        contrib = np.bincount(kidx.ravel(), weights=(S * FG).real.ravel(), minlength=n) \
            + 1j * np.bincount(kidx.ravel(), weights=(S * FG).imag.ravel(), minlength=n)
        Ghat += w[p] * contrib / n
    Winv = np.conj(Wf).T                                                 # IDFT_u: x <- l
    QG = Winv @ Ghat
    loss = (Winv @ (pre.bb2 * Gl)) / n
    return prefac * (QG.real - f * loss.real)
