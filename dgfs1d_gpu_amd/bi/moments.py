"""Moment writer of the binary mixture (reference dgfs1D/bi/moments.py:10-193).

Per species p at the reconstruction points (FP64 on the device):
  rho_p = m_p cw sum f_p,  u_p = m_p cw sum c f_p / rho_p  (x, y),
  T_p   = 2/3 m_p^2 cw sum f_p |c - u_p|^2 / rho_p,
  q_p   = m_p^2 cw sum f_p |c - u|^2 (c_x - u_x)  with u the mixture velocity,
then dimensionalised; p_p = m_p R0/molarMass0 rho_p T_p, which is m_p^2 times
n_p k T, kept as the reference writes it (bi/moments.py:106-109), and
nden_p = NA/(m_p molarMass0) rho_p.  As in the reference the peculiar
velocities take u_z = 0 (moments.py:59-61, :83-85).  Columns: x0 then rho,
U:x, U:y, T, Q:x, p, nden with suffix :p for each species."""
import os

import numpy as np

PRIVARS = ['rho', 'U:x', 'U:y', 'T', 'Q:x', 'p', 'nden']


def bi_moments(coeffs, im, vm):
    """(Nqr, Ne, 7 * nspcs) dimensional moments of the coefficient arrays
    coeffs[p] (K*Ne*Nv, any float dtype), im = (K, Nqr) interpolation."""
    import torch
    dev = coeffs[0].device
    cv = torch.tensor(vm.cv(), dtype=torch.float64, device=dev)
    cw, N = vm.cw(), vm.vsize()
    K, Nqr = im.shape
    imd = torch.tensor(im, dtype=torch.float64, device=dev)
    ms = vm.masses()
    ns = len(coeffs)
    Ne = coeffs[0].numel() // (K * N)
    f = [torch.einsum('mr,mej->rej', imd, c.double().view(K, Ne, N)) for c in coeffs]
    out = np.zeros((Nqr, Ne, 7 * ns))
    rho_t = torch.zeros(Nqr, Ne, dtype=torch.float64, device=dev)
    mom_t = torch.zeros(2, Nqr, Ne, dtype=torch.float64, device=dev)
    per = []
    for p in range(ns):
        mr, fs = ms[p], f[p]
        mcw = mr * cw
        rho = fs.sum(-1) * mcw
        ux = (fs @ cv[0]) * mcw / rho
        uy = (fs @ cv[1]) * mcw / rho
        c2 = ((cv[0] - ux[..., None]) ** 2 + (cv[1] - uy[..., None]) ** 2 + cv[2] ** 2)
        T = (fs * c2).sum(-1) * (2.0 / 3.0 * mcw * mr) / rho
        per.append([rho, ux, uy, T])
        rho_t += rho
        mom_t[0] += rho * ux
        mom_t[1] += rho * uy
    Ux, Uy = mom_t[0] / rho_t, mom_t[1] / rho_t
    cx = cv[0] - Ux[..., None]
    c2 = cx ** 2 + (cv[1] - Uy[..., None]) ** 2 + cv[2] ** 2
    rho0, u0, T0, M0 = vm.rho0(), vm.u0(), vm.T0(), vm.molarMass0()
    for p in range(ns):
        mr = ms[p]
        rho, ux, uy, T = per[p]
        qx = mr * (f[p] * c2 * cx).sum(-1) * mr * cw
        cols = [rho * rho0, ux * u0, uy * u0, T * T0, qx * 0.5 * rho0 * u0 ** 3]
        o = 7 * p
        for i, c in enumerate(cols):
            out[..., o + i] = c.cpu().numpy()
        out[..., o + 5] = (mr * vm.R0 / M0) * out[..., o] * out[..., o + 3]
        out[..., o + 6] = (vm.NA / mr / M0) * out[..., o]
    return out


class DGFSMomWriterBi:
    def __init__(self, tcurr, im, xcoeff, coeffs, vm, cfg, cfgsect, **kwargs):
        self.vm = vm
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith('.txt'):
            self.basename += '.txt'
        names = ['%s:%d' % (v, p) for p in range(vm.nspcs()) for v in PRIVARS]
        self.fields = ', '.join(['x0'] + names)
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr
        self.im = im
        Kb, Nqr = im.shape
        Ne = xcoeff.shape[1]
        self.xsol = (np.einsum('mr,me->re', im, xcoeff) * vm.H0()).T.reshape(-1, 1)
        self.last = None
        self(1e-10, tcurr, coeffs)

    def __call__(self, dt, tcurr, coeffs):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        sol = bi_moments(coeffs, self.im, self.vm)
        self.last = (tcurr, sol)
        Nqr, Ne, nf = sol.shape
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        np.savetxt(fname, np.hstack((self.xsol, sol.swapaxes(0, 1).reshape(Nqr * Ne, nf))),
                   header='t={0} \n {1}'.format(tcurr, self.fields), comments='#')
        self.tout_next = tcurr + self.dt_out
