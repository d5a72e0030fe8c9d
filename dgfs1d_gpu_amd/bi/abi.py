"""dgfsBiAstd1D: the binary mixture on the IMEX (asymptotic-preserving) path.

The mixture deck of dgfsBi1D (dgfs1D/bi/bi.py:31-421: mesh, basis, mixture
velocity mesh, [scattering-model] type = vss with the pair parameters,
per-species walls and initial state) with the single-species IMEX driver's
[time-integrator] schemes (dgfs1D/astd/abstd.py:37-480: imex-lm-bdf1..4 and
the other linear-multistep kinds, ars-111/222/443, sspl-222/443,
bdf-111/222/333, boltz-ars-111) and a [penalized-scattering-model] section
(type = shakov-direct-gll, optional Pr): each species is penalised by its own
Shakov model (dgfs1d_gpu_amd/bi/penalty.py).  The reference has no IMEX
mixture solver; the integrators and their update kernels are the
single-species ones, run on the species-major mixture state (both species in
one buffer), with

    explicit(t, f, out)   = invMass(transport_p(f_p))            per species p
    explicitQ(t, f, out)  = invMass(sum_q Q_pq(f_p, f_q))        per species p

(the mixture conservation projection applied to the collision terms when the
deck asks for it).

The penalisation is a pointwise operation on distributions.  On a nodal
basis the coefficients are point values and the integrators advance them
directly.  On a modal basis the integrators advance the values at the K
Gauss-Legendre points of each element instead, w = V c (V the K x K matrix of
the modes at those points, invertible): explicit parts are V L(V^-1 w), the
penalty acts on w, and the coefficients c = V^-1 w are refreshed after every
step.  (The single-species driver, like the reference, penalises coefficient
slices, which are distributions only on nodal bases.)

Usage:  python -m dgfs1d_gpu_amd.bi.abi run deck.ini [-v sect::key value]...
"""
import logging
import sys

import numpy as np

from ..astd.integrator import DGFSIntegratorAstd
from ..basis import ModalModifiedKarniadakis
from ..initialize import initialize
from ..nputil import subclass_where
from ..quadratures import ortho_basis_at, zwgj
from ..transport import MatOp
from .bi import DGFSBi1D
from .penalty import species_penalties

log = logging.getLogger('dgfs1d_gpu_amd')


def point_values_map(basis):
    """K x K matrix taking a modal basis' coefficients to its values at the K
    Gauss-Legendre points of the element; None for a nodal basis."""
    kind, K = basis.basis_kind, basis.K
    if kind.startswith('nodal'):
        return None
    z, _ = zwgj(K, 0.0, 0.0)
    if kind == 'modal-orthonormal-legendre':
        B = ortho_basis_at(K, z)
    elif kind == 'modal-modified-karniadakis':
        B = ModalModifiedKarniadakis.modified_basis_at(K, z)
    else:
        raise ValueError('no point-value map for basis %r' % kind)
    return B.T                                  # values[q] = sum_m B[m, q] c[m]


class DGFSBiAstd1D(DGFSBi1D):
    def _make_integrator(self, cfg, shape):
        K, Ne, Nv = shape
        ns = self.nspcs
        torch = self.torch
        self.psm = species_penalties(cfg, self.vm, K * Ne, self.device)
        n = self._n = K * Ne * Nv
        z = lambda k: torch.zeros(k, dtype=self.tdtype, device=self.device)
        self._ux, self._lx = z(n), z(n)
        V = point_values_map(self.basis)
        if V is None:
            self._to_points = self._to_coeffs = None
            self.w = self.u
        else:
            self._to_points = MatOp(V, self.device)
            self._to_coeffs = MatOp(np.linalg.inv(V), self.device)
            self.w = z(ns * n)
            self._cs = list(z(ns * n).view(ns, n))          # coefficient scratch
            for p in range(ns):
                self._to_points(self.us[p], self._species(self.w)[p])
        kind = cfg.lookup('time-integrator', 'scheme')
        odescls = subclass_where(DGFSIntegratorAstd, intg_kind=kind)
        return odescls(self.explicit, self.psm, (ns * K, Ne, Nv), self.tdtype,
                       device=self.device, t=self.time, dt=self.dt, f0=self.w,
                       explicitQ=self.explicitQ)

    def _species(self, x):
        return list(x.view(self.nspcs, -1))

    def _coeffs(self, w):
        """Species coefficient arrays of an integrator state."""
        if self._to_coeffs is None:
            return self._species(w)
        for p, wp in enumerate(self._species(w)):
            self._to_coeffs(wp, self._cs[p])
        return self._cs

    def _store(self, lx, out_p):
        """out_p = invMass(lx) (in point values on a modal basis)."""
        if self._to_points is None:
            if self.invmass.is_identity:
                out_p.copy_(lx)
            else:
                self.invmass(lx, out_p)
        elif self.invmass.is_identity:
            self._to_points(lx, out_p)
        else:
            self.invmass(lx, self._ux)
            self._to_points(self._ux, out_p)

    def explicit(self, t, f, out):
        """out_p = invMass(transport_p(f_p)) for every species."""
        cs, outs = self._coeffs(f), self._species(out)
        for p in range(self.nspcs):
            direct = self._to_points is None and self.invmass.is_identity
            self.transport[p](t, cs[p], outs[p] if direct else self._lx)
            if not direct:
                self._store(self._lx, outs[p])

    def explicitQ(self, t, f, out):
        """out_p = invMass(sum_q Q_pq(f_p, f_q)) for every species."""
        cs, outs = self._coeffs(f), self._species(out)
        out.zero_()
        for p in range(self.nspcs):
            direct = self._to_points is None and self.invmass.is_identity
            if direct:
                self._collide(p, cs, outs[p])
            else:
                self._lx.zero_()
                self._collide(p, cs, self._lx)
                self._store(self._lx, outs[p])

    def _advance(self):
        self.odes.integrate(self.time, self.dt, self.nacptsteps, self.w)
        if self._to_coeffs is not None:
            for p, wp in enumerate(self._species(self.w)):
                self._to_coeffs(wp, self.us[p])


def main(argv=None):
    logging.basicConfig(level=logging.INFO, format='%(asctime)s %(message)s')
    cfg, args = initialize(argv)
    restart = args.dist[0][0].name if hasattr(args, 'process_restart') else None
    DGFSBiAstd1D(cfg, restart=restart).run()


if __name__ == '__main__':
    sys.exit(main())
