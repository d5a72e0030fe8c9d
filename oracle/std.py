"""Oracle: the dgfsStd1D RHS, boundaries, integrators, moments (FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, with explicit face arrays as the reference kernels use them:
  face maps mapL/mapR          dgfs1D/std/std.py:62-66
  rhs()                        dgfs1D/std/std.py:281-366
  extract/transfer/insert      dgfs1D/std/kernels/std.mako:4-90
  flux / mul_by_adv /
  totalFlux / mul_by_invjac    dgfs1D/std/kernels/std.mako:92-190
  boundaries                   dgfs1D/std/bc.py:25-324 + std/kernels/bcs/*.mako
  collision loop               dgfs1D/std/std.py:352-356
  integrators                  dgfs1D/std/integrator.py:36-157
  moments                      dgfs1D/std/moments.py:11-70
  residual                     dgfs1D/std/residual.py:21-29
"""
import numpy as np

from . import collision as OC
from . import relax as OR


def maxwellian(cv, rho, ux, uy, uz, T):
    """std/initcond.py:56-58"""
    u = np.array([ux, uy, uz]).reshape(3, 1)
    return rho / (np.pi * T) ** 1.5 * np.exp(-((cv - u) ** 2).sum(0) / T)


class Wall:
    """kind: 'copy' | 'diffuse' | 'diffuse-expr' | 'inlet' | 'vanish' (bc.py).
    'diffuse-expr' (std/bc.py:136-198, dgfs-wall-expr-diffuse.mako): wall
    velocity and temperature are functions of time (``state(t)`` returns
    non-dimensional (ux, uy, uz, T)); the wall distribution is the
    unnormalised exp(-|c - u|^2 / T), the density factor coming from the
    zero-net-mass-flux condition as for the fixed wall."""

    def __init__(self, kind, nl, cv, cw, f0=None, ushift=0.0, state=None):
        self.kind, self.nl, self.cv, self.cw = kind, nl, cv, cw
        self.f0, self.ushift, self.state = f0, ushift, state
        self.nw = None

    def apply(self, ul, t=0.0):
        c = self.cv[0]
        if self.kind == 'diffuse-expr':                             # mako:14-26, :40-52
            ux, uy, uz, T = self.state(t)
            u = np.array([ux, uy, uz]).reshape(3, 1)
            mx = np.exp(-((self.cv - u) ** 2).sum(0) / T)
            fac = (c - ux) * self.nl
            num = np.sum(np.where(fac >= 0, fac * ul * self.cw, 0.0))
            den = np.sum(np.where(fac < 0, fac * mx * self.cw, 0.0))
            self.nw = -num / den
            return np.where(fac < 0, mx * self.nw, ul)
        fac = (c - self.ushift) * self.nl
        if self.kind == 'copy':
            return ul.copy()
        if self.kind == 'diffuse':                                  # wall-diffuse.mako:24-37
            num = np.sum(np.where(fac >= 0, fac * ul * self.cw, 0.0))
            den = np.sum(np.where(fac < 0, fac * self.f0 * self.cw, 0.0))
            self.nw = -num / den                                    # bc.py:265-267
            return np.where(fac < 0, self.f0 * self.nw, ul)         # mako:10-21
        if self.kind == 'inlet':
            return np.where(fac < 0, self.f0, ul)
        if self.kind == 'vanish':
            return np.where(fac >= 0, 0.0, ul)
        raise ValueError(self.kind)


class StdOracle:
    """FP64 dgfsStd1D operator.  u has shape (K, Ne, n)."""

    def __init__(self, basis, invjac, vm, wall_l, wall_r, pre, prefac, model='vhs',
                 shakov=None, project=False):
        """shakov = (omega, Pr) for model 'shakov-direct-gll'; project applies
        the conservation projection to each output slice's collision sum."""
        self.b, self.invjac, self.vm = basis, np.asarray(invjac).ravel(), vm
        self.wl, self.wr = wall_l, wall_r
        self.pre, self.prefac, self.model = pre, prefac, model
        self.shakov, self.project = shakov, project
        self.K = basis.K
        self.Ne = len(self.invjac)
        self.Nqf = basis.face.shape[0]

    def transport(self, t, u):
        K, Ne, Nqf = self.K, self.Ne, self.Nqf
        n = u.shape[-1]
        c = self.vm.cv[0]
        usolF = (self.b.face @ u.reshape(K, -1)).reshape(Nqf * Ne, n)   # bwdTransFace
        mapL = np.arange(Ne + 1) + (Nqf - 1) * Ne - 1
        mapR = np.arange(Ne + 1)
        mapL[0], mapR[-1] = 0, Ne * Nqf - 1
        uL, uR = usolF[mapL].copy(), usolF[mapR].copy()                 # extract_left/right
        bcL, bcR = uL[0].copy(), uR[Ne].copy()                          # transfer_bc_*
        uL[0] = self.wl.apply(bcL, t)                                      # update/apply/insert
        uR[Ne] = self.wr.apply(bcR, t)
        jL = 0.5 * np.abs(c) * (uL - uR) + 0.5 * c * (uR - uL)          # flux
        jR = 0.5 * np.abs(c) * (uL - uR) + 0.5 * c * (uL - uR)
        ux = (self.b.deriv @ u.reshape(K, -1)).reshape(K, Ne, n)        # deriv
        ux *= -c                                                        # mul_by_adv
        for k in range(K):                                              # totalFlux
            for e in range(Ne):
                ux[k, e] += -jL[e + 1] * self.b.gRD[k] + jR[e] * self.b.gLD[k]
        ux *= self.invjac[None, :, None]                                # mul_by_invjac
        return ux

    def transport_upwind(self, t, u, Dp, Dm):
        """Split upwind summation-by-parts transport (astd/abstdu.py:295-380,
        astd/kernels/upwind.mako): f = c+ u, g = c- u, ux = -(Dm f + Dp g)
        with Dp, Dm in the 'ux = A u' orientation, jumps jL = c- (uR - uL),
        jR = -c+ (uR - uL), lift and 1/J as the standard transport."""
        K, Ne, Nqf = self.K, self.Ne, self.Nqf
        n = u.shape[-1]
        c = self.vm.cv[0]
        usolF = (self.b.face @ u.reshape(K, -1)).reshape(Nqf * Ne, n)
        mapL = np.arange(Ne + 1) + (Nqf - 1) * Ne - 1
        mapR = np.arange(Ne + 1)
        mapL[0], mapR[-1] = 0, Ne * Nqf - 1
        uL, uR = usolF[mapL].copy(), usolF[mapR].copy()
        bcL, bcR = uL[0].copy(), uR[Ne].copy()
        uL[0] = self.wl.apply(bcL, t)
        uR[Ne] = self.wr.apply(bcR, t)
        cp, cm = np.maximum(c, 0.0), np.minimum(c, 0.0)
        jL = cm * (uR - uL)                                              # uflux
        jR = -cp * (uR - uL)
        f, g = cp * u, cm * u                                            # splitFlux
        ux = -((Dm @ f.reshape(K, -1)) + (Dp @ g.reshape(K, -1))).reshape(K, Ne, n)
        for k in range(K):                                               # totalFlux
            for e in range(Ne):
                ux[k, e] += -jL[e + 1] * self.b.gRD[k] + jR[e] * self.b.gLD[k]
        ux *= self.invjac[None, :, None]
        return ux

    @staticmethod
    def log_mean(a, b):
        """aveL of the entropy driver (astd/kernels/limiters.mako): a stable
        evaluation of the logarithmic mean (a - b)/(ln a - ln b)."""
        xi = np.abs(a / (b + 1e-20))
        f = (xi - 1) / (xi + 1)
        u = f * f
        with np.errstate(divide='ignore', invalid='ignore'):
            den = np.where(u < 1e-2, 1 + u / 3 + u * u / 5 + u * u * u / 7,
                           np.log(xi + 1e-20) / 2 / (f + 1e-20))
        return 0.5 * (a + b) / den

    def transport_entropy(self, t, u, Dr, mean=None):
        """Entropy-stable transport (astd/abstde2.py:214-217, eDeriv): volume
        term -2 sum_l Dr[k][l] c L(u_k, u_l) with the logarithmic mean (or
        ``mean``), the standard upwind flux (std.mako:115-152), lift and 1/J."""
        mean = mean or self.log_mean
        K, Ne, Nqf = self.K, self.Ne, self.Nqf
        n = u.shape[-1]
        c = self.vm.cv[0]
        usolF = (self.b.face @ u.reshape(K, -1)).reshape(Nqf * Ne, n)
        mapL = np.arange(Ne + 1) + (Nqf - 1) * Ne - 1
        mapR = np.arange(Ne + 1)
        mapL[0], mapR[-1] = 0, Ne * Nqf - 1
        uL, uR = usolF[mapL].copy(), usolF[mapR].copy()
        bcL, bcR = uL[0].copy(), uR[Ne].copy()
        uL[0] = self.wl.apply(bcL, t)
        uR[Ne] = self.wr.apply(bcR, t)
        jL = 0.5 * np.abs(c) * (uL - uR) + 0.5 * c * (uR - uL)
        jR = 0.5 * np.abs(c) * (uL - uR) + 0.5 * c * (uL - uR)
        ux = np.zeros_like(u)
        for k in range(K):                                               # eDeriv
            for l in range(K):
                ux[k] += -2.0 * Dr[k, l] * c * mean(u[k], u[l])
        for k in range(K):                                               # totalFlux
            for e in range(Ne):
                ux[k, e] += -jL[e + 1] * self.b.gRD[k] + jR[e] * self.b.gLD[k]
        return ux * self.invjac[None, :, None]

    def collisions(self, u, ux):
        K, Ne = self.K, self.Ne
        vm = self.vm
        for m in range(K):
            f = np.einsum('ra,aen->ren', self.b.U[m].T, u)              # trans_U
            g = np.einsum('rb,ben->ren', self.b.V[m], u)                # trans_V
            for e in range(Ne):
                acc = np.zeros(u.shape[-1])
                for r in self.b.sig[m]:
                    if self.model == 'vhs':
                        acc += OC.q_vhs(self.pre, f[r, e], g[r, e], self.prefac)
                    elif self.model == 'vhs-gll':
                        acc += OC.q_vhs_gll(self.pre, f[r, e], self.prefac)
                    elif self.model == 'shakov-direct-gll':
                        omega, Pr = self.shakov
                        acc += OR.q_shakov(f[r, e], vm.cv, vm.cw, self.prefac, omega, Pr)
                    elif self.model == 'bgk-direct-gll':
                        acc += OR.q_bgk(f[r, e], vm.cv, vm.cw, self.prefac, self.shakov[0])
                    elif self.model == 'bgk-gll':
                        acc += OR.q_bgk_conservative(f[r, e], vm.cv, vm.cw, self.prefac,
                                                     self.shakov[0])
                    elif self.model == 'esbgk-gll':
                        omega, Pr = self.shakov
                        acc += OR.q_esbgk_conservative(f[r, e], vm.cv, vm.cw, self.prefac,
                                                       omega, Pr)
                    elif self.model == 'esbgk-direct-gll':
                        omega, Pr = self.shakov
                        acc += OR.q_esbgk(f[r, e], vm.cv, vm.cw, self.prefac, omega, Pr)
                    else:
                        raise ValueError(self.model)
                ux[m, e] += OR.project_conservative(acc, vm.cv) if self.project else acc
        return ux

    def rhs(self, t, u):
        ux = self.collisions(u, self.transport(t, u))
        K = self.K
        return (self.b.invmass @ ux.reshape(K, -1)).reshape(ux.shape)   # invMass


def integrate(op, scheme, t, dt, u):
    """One step of the explicit schemes (std/integrator.py:36-157)."""
    if scheme == 'euler':
        return u + dt * op.rhs(t, u)
    if scheme == 'ssp-rk2':
        r0 = u + dt * op.rhs(t, u)
        return 0.5 * u + 0.5 * r0 + 0.5 * dt * op.rhs(t + dt, r0)
    if scheme == 'ssp-rk3':
        r0 = u + dt * op.rhs(t, u)
        r0 = 0.75 * u + 0.25 * r0 + 0.25 * dt * op.rhs(t + dt, r0)
        return u / 3.0 + 2.0 / 3.0 * r0 + 2.0 / 3.0 * dt * op.rhs(t + 0.5 * dt, r0)
    if scheme in ('lse-rk45', 'lesrk-45'):
        a = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
             -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
        b = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
             1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
             2277821191437.0 / 14882151754819.0]
        c = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
             2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]
        r1 = np.zeros_like(u)
        for i in range(5):
            r1 = a[i] * r1 + dt * op.rhs(t + c[i] * dt, u)
            u = u + b[i] * r1
        return u
    raise ValueError(scheme)


def moments(u, interp, vm):
    """Non-dimensional (Nqr, Ne, 9): rho, ux, uy, T, qx, qy, Pxx, Pyy, Pxy
    (std/moments.py:26-61; u_z is taken as 0 in the peculiar velocity)."""
    cv, cw = vm.cv, vm.cw
    s = np.einsum('mr,mej->rej', interp, u)
    out = np.zeros(s.shape[:2] + (9,))
    rho = s.sum(-1) * cw
    ux = (s @ cv[0]) * cw / rho
    uy = (s @ cv[1]) * cw / rho
    cx = cv[0][None, None, :] - ux[..., None]
    cy = cv[1][None, None, :] - uy[..., None]
    cz = cv[2][None, None, :]
    c2 = cx * cx + cy * cy + cz * cz
    out[..., 0], out[..., 1], out[..., 2] = rho, ux, uy
    out[..., 3] = (s * c2).sum(-1) * (2.0 / 3.0 * cw) / rho
    out[..., 4] = (s * c2 * cx).sum(-1) * cw
    out[..., 5] = (s * c2 * cy).sum(-1) * cw
    out[..., 6] = 2 * (s * cx * cx).sum(-1) * cw
    out[..., 7] = 2 * (s * cy * cy).sum(-1) * cw
    out[..., 8] = 2 * (s * cx * cy).sum(-1) * cw
    return out


def residual(cur, prev):
    d = cur - prev
    return np.sqrt(np.sum(d * d) / np.sum(prev * prev))
