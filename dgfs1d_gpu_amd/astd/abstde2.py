"""dgfsAstde21D: the entropy-stable IMEX driver.

Same INI schema as the reference dgfs1D/astd/abstde2.py:29-295 (the IMEX
driver's deck with a nodal basis).  The explicit part replaces the volume
term -c D f of the transport by entropy-conservative flux differencing with
the logarithmic mean L (abstde2.py:214-217, eDeriv):

    explicit(t, f, out) = invMass( invjac ( -2 sum_l Dr[k][l] c L(f_k, f_l) + lift(jL, jR) ) )

with the standard upwind face flux, in the same fused transport kernel.
[time-integrator] limiter = 1 applies the linear limiter and then the
positivity limiter after every step (abstde2.py:272-274).  The time loop is
the reference's ``while time < tend``.

Usage:  python -m dgfs1d_gpu_amd.astd.abstde2 run dgfs1D.ini [-v sect::key value]...
"""
import logging
import sys

from ..initialize import initialize
from ..transport import PositivityLimiter, Transport
from .abstd import DGFSAstd1D

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSAstde21D(DGFSAstd1D):
    def __init__(self, cfg, device='cuda', restart=None, writers=True):
        super().__init__(cfg, device=device, restart=restart, writers=writers)
        if not hasattr(self.basis, 'derivMat') or self.basis.Nq != self.basis.K:
            raise ValueError('the entropy-stable driver needs a nodal basis')
        self.transport = Transport(self.basis, self.mesh.invjac[:, 0], self.vm, self.bcl,
                                   self.bcr, self.tdtype, self.device, entropy=True)
        self.poslimiter = None
        if self.limiter is not None:
            self.poslimiter = PositivityLimiter(self.basis, self.Ne, self.Nv, self.tdtype,
                                                self.device)

    def _advance(self):
        super()._advance()                              # IMEX step, then limitLin
        if self.poslimiter is not None:                 # abstde2.py:274
            self.poslimiter(self.u, self.u)


def main(argv=None):
    logging.basicConfig(level=logging.INFO, format='%(asctime)s %(message)s')
    cfg, args = initialize(argv)
    restart = args.dist[0][0].name if hasattr(args, 'process_restart') else None
    DGFSAstde21D(cfg, restart=restart).run()


if __name__ == '__main__':
    sys.exit(main())
