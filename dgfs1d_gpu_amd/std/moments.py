"""Moment writer (reference dgfs1D/std/moments.py:140-220): FP64 moments at the
reconstruction points computed on the GPU (dgfs_moments), dimensionalised on
the host and written as the same text table (header ``t=...`` + fields)."""
import os

import numpy as np

from ..transport import Moments

PRIVARS = ['rho', 'U:x', 'U:y', 'T', 'Q:x', 'Q:y', 'P:xx', 'P:yy', 'P:xy', 'p']


def dimensional(nd, vm):
    """(..., 9) non-dimensional -> (..., 10) dimensional (moments.py:63-70)."""
    rho0, u0, T0 = vm.rho0(), vm.u0(), vm.T0()
    sc = np.array([rho0, u0, u0, T0, 0.5 * rho0 * u0 ** 3, 0.5 * rho0 * u0 ** 3,
                   0.5 * rho0 * u0 ** 2, 0.5 * rho0 * u0 ** 2, 0.5 * rho0 * u0 ** 2])
    out = np.empty(nd.shape[:-1] + (10,))
    out[..., :9] = nd * sc
    out[..., 9] = (vm.R0 / vm.molarMass0()) * out[..., 0] * out[..., 3]
    return out


class DGFSMomWriterStd:
    def __init__(self, tcurr, im, xcoeff, coeff, vm, cfg, cfgsect, device=None, K=None,
                 Ne=None):
        self.vm, self.cfg = vm, cfg
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith('.txt'):
            self.basename += '.txt'
        self.fields = ', '.join(['x0'] + PRIVARS)
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr
        Kb, Nqr = im.shape
        Ne = xcoeff.shape[1]
        self.im = im
        self.xsol = (np.matmul(im.T, xcoeff.reshape(Kb, -1)).reshape(Nqr, Ne, -1) * vm.H0())
        self.xsol = self.xsol.swapaxes(0, 1).reshape(Nqr * Ne, -1)
        self.mom = Moments(im, Kb, Ne, vm, device)
        self.last = None
        self(1e-10, tcurr, coeff)

    def compute(self, coeff):
        nd = self.mom(coeff)                         # (Nqr, Ne, 9)
        return dimensional(nd, self.vm)              # (Nqr, Ne, 10)

    def __call__(self, dt, tcurr, coeff):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        sol = self.compute(coeff)
        self.last = (tcurr, sol)
        Nqr, Ne, _ = sol.shape
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        np.savetxt(fname, np.hstack((self.xsol, sol.swapaxes(0, 1).reshape(Nqr * Ne, -1))),
                   header='t={0} \n {1}'.format(tcurr, self.fields), comments='#')
        self.tout_next = tcurr + self.dt_out
