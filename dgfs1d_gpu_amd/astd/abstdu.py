"""dgfsAstdu1D: the upwind summation-by-parts IMEX driver.

Same INI schema as the reference dgfs1D/astd/abstdu.py:37-522 (the IMEX
driver's deck; the basis must be nodal-sem-gll, whose [basis] sigma, default
-1, sets the upwind dissipation of basis.py:303-312).  The explicit part is
the split upwind transport (abstdu.py:295-380):

    explicit(t, f, out) = invMass( invjac ( -(D- c+ f + D+ c- f) + lift(jL, jR) ) )

with face jumps jL = c- (uR - uL), jR = -c+ (uR - uL) (upwind.mako), in the
same fused transport kernel as the other drivers.  [time-integrator]
limiter = 1 applies the linear limiter and then the positivity limiter after
every step (abstdu.py:498-500).

Usage:  python -m dgfs1d_gpu_amd.astd.abstdu run dgfs1D.ini [-v sect::key value]...
"""
import logging
import sys

from ..initialize import initialize
from ..transport import PositivityLimiter, Transport
from .abstd import DGFSAstd1D

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSAstdu1D(DGFSAstd1D):
    def __init__(self, cfg, device='cuda', restart=None, writers=True):
        super().__init__(cfg, device=device, restart=restart, writers=writers)
        sigma = float(cfg.lookupordefault('basis', 'sigma', -1.0))
        self.upwind = self.basis.upwind_ops(sigma)
        self.transport = Transport(self.basis, self.mesh.invjac[:, 0], self.vm, self.bcl,
                                   self.bcr, self.tdtype, self.device, upwind=self.upwind)
        self.poslimiter = None
        if self.limiter is not None:
            self.poslimiter = PositivityLimiter(self.basis, self.Ne, self.Nv, self.tdtype,
                                                self.device)

    def _advance(self):
        super()._advance()                              # IMEX step, then limitLin
        if self.poslimiter is not None:                 # abstdu.py:500
            self.poslimiter(self.u, self.u)


def main(argv=None):
    logging.basicConfig(level=logging.INFO, format='%(asctime)s %(message)s')
    cfg, args = initialize(argv)
    restart = args.dist[0][0].name if hasattr(args, 'process_restart') else None
    DGFSAstdu1D(cfg, restart=restart).run()


if __name__ == '__main__':
    sys.exit(main())
