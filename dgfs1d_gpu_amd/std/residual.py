"""Residual monitor (reference dgfs1D/std/residual.py:9-31):
sqrt(sum (f^{n+1} - f^n)^2 / sum (f^n)^2) every ``nsteps`` (and at step 1),
FP64 reductions on the GPU."""
import logging

from ..transport import Residual

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSResidualStd:
    def __init__(self, cfg, cfgsect, device=None):
        self.nsteps = cfg.lookupordefault(cfgsect, 'nsteps', -1)
        self.res = Residual(device)
        self.history = []

    def __call__(self, tcurr, nsteps, solprev, solcurr):
        if self.nsteps > 0 and (nsteps % self.nsteps == 0 or nsteps == 1):
            r = self.res(solcurr, solprev)
            self.history.append((tcurr, nsteps, r))
            log.info('residual at t = %s %s', tcurr, r)
            return r
        return None
