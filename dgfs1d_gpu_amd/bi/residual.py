"""Residual monitor of the binary mixture (reference dgfs1D/bi/residual.py:
9-30): per species sqrt(sum (f^{n+1} - f^n)^2 / sum (f^n)^2) every ``nsteps``
(and at step 1), FP64 reductions on the GPU."""
import logging

import numpy as np

from ..transport import Residual

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSResidualBi:
    def __init__(self, cfg, cfgsect, device=None):
        self.nsteps = cfg.lookupint(cfgsect, 'nsteps')
        self.res = Residual(device)
        self.history = []

    def __call__(self, tcurr, nsteps, solprevs, solcurrs):
        if nsteps % self.nsteps == 0 or nsteps == 1:
            r = np.array([self.res(c, p) for p, c in zip(solprevs, solcurrs)])
            self.history.append((tcurr, nsteps, r))
            log.info('residual at t = %s %s', tcurr, r)
            return r
        return None
