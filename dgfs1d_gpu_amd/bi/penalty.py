"""Shakov penalisation of the binary mixture's IMEX path.

The reference's penalised IMEX schemes exist for a single species only
(dgfs1D/astd/integrator.py:623-849 with astd/scattering.py:355-397); its
mixture solver (dgfs1D/bi/bi.py) is explicit.  Here each species p is
penalised by its own 14-moment Shakov model,

    P_p = nu_p (S_p - f_p),   nu_p = prefac_p n_p T_p^(1 - omega_pp),

with S_p the Shakov distribution of species p (its number density n_p,
velocity u_p, temperature T_p and heat flux q_p) for molecular mass m_p:
S_p = n_p (m_p/(pi T_p))^(3/2) exp(-m_p |c - u_p|^2 / T_p) (1 + ...).
prefac_p = t0 n0 k T0 / mu_p(T0) is species p's own Shakov frequency, with
mu_p the variable-soft-sphere viscosity of the pure species (Bird's formula
for alpha_pp, omega_pp, dRef_pp, Tref_pp, m_p).

The single-species kernels evaluate these unchanged on the species' scaled
grid c' = sqrt(m_p) c with cell weight m_p^(3/2) cw: there the species'
Maxwellian is the unit-mass one, the moments are n' = m_p^(3/2) n_p,
u' = sqrt(m_p) u_p, T' = T_p, and the Shakov target of the scaled moments,
evaluated at c', is S_p at c.  The frequency prefactor passed to the kernels
is prefac_p m_p^(-3/2), so nu = prefac_p n_p T_p^(1 - omega_pp).

``MixturePenalty`` exposes the penalisation interface the IMEX integrators
call (moment, constructMaxwellian, collide, updateMomentLM, updateDistLM,
updateMomentARS, updateDistARS, new_moments) on the mixture's species-major
state (species 0's K*Ne slices, then species 1's) and moment arrays, and
forwards each species' half to its model.
"""
import math

from ..astd.scattering import DGFSShakovDirectGLLScatteringModelAstd

KB = 1.38064852e-23


def vss_viscosity(m, alpha, omega, dRef, Tref, T):
    """Viscosity of a variable-soft-sphere gas of molecular mass m [kg] at T
    (Bird, Molecular Gas Dynamics, eq. 3.74 with (T/Tref)^omega)."""
    mu_ref = (5.0 * (alpha + 1.0) * (alpha + 2.0) * math.sqrt(m * KB * Tref / math.pi)
              / (4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega) * dRef * dRef))
    return mu_ref * (T / Tref) ** omega


class DGFSShakovSpeciesPenalty(DGFSShakovDirectGLLScatteringModelAstd):
    """Species p's Shakov penalisation on its scaled grid (see the module
    docstring); p is the species index."""
    scattering_model = None            # not a single-species registry entry

    def __init__(self, cfg, vm, p, **kwargs):
        self._p = int(p)
        super().__init__(cfg, vm, **kwargs)

    def load_parameters(self):
        cfg, vm, p = self.cfg, self.vm, self._p
        sect = 'penalized-scattering-model'
        pp = '%d%d' % (p, p)
        self._Pr = float(cfg.lookupordefault(sect, 'Pr', 2. / 3.))
        self._omega = float(cfg.lookupfloat('scattering-model', 'omega' + pp))
        self._m = float(vm.masses()[p])
        mol = self._m * vm.molarMass0() / vm.NA                     # kg
        mu0 = vss_viscosity(mol, float(cfg.lookupfloat('scattering-model', 'alpha' + pp)),
                            self._omega, float(cfg.lookupfloat('scattering-model', 'dRef' + pp)),
                            float(cfg.lookupfloat('scattering-model', 'Tref' + pp)), vm.T0())
        t0 = vm.H0() / vm.u0()
        self.species_prefac = t0 * vm.n0() * KB * vm.T0() / mu0
        self._prefactor = self.species_prefac * self._m ** -1.5

    def perform_precomputation(self):
        super().perform_precomputation()
        s = math.sqrt(self._m)
        self.cx, self.cy, self.cz = (s * c for c in (self.cx, self.cy, self.cz))
        self.cw = float(self.vm.cw()) * self._m ** 1.5


class MixturePenalty:
    """The species' penalisations over the species-major mixture state."""
    nalph = 14

    def __init__(self, models, nslice):
        self.models = list(models)
        self.ns = len(self.models)
        self.nslice = int(nslice)                  # slices per species (K*Ne)
        self.torch = self.models[0].torch
        self._device = self.models[0]._device

    @property
    def prefac(self):
        return self.models[0].prefac

    def new_moments(self, nslice):
        if nslice != self.ns * self.nslice:
            raise ValueError('moment array for %d slices requested; the mixture has %d'
                             % (nslice, self.ns * self.nslice))
        return self.torch.zeros(nslice * self.nalph, dtype=self.torch.float64,
                                device=self._device)

    def _half(self, x, p):
        if x is None:
            return None
        return x.view(self.ns, -1)[p]

    def _halves(self, xs, p):
        return [self._half(x, p) for x in xs]

    def moment(self, t, f, U, stream=None):
        for p, m in enumerate(self.models):
            m.moment(t, self._half(f, p), self._half(U, p), stream)

    def constructMaxwellian(self, t, U, M, stream=None):
        for p, m in enumerate(self.models):
            m.constructMaxwellian(t, self._half(U, p), self._half(M, p), stream)

    def collide(self, t, U, M, f, Q, stream=None):
        for p, m in enumerate(self.models):
            m.collide(t, self._half(U, p), self._half(M, p), self._half(f, p), self._half(Q, p),
                      stream)

    def updateMomentLM(self, dt, A, B, C, LU, LM, U, nstages, stream=None):
        for p, m in enumerate(self.models):
            m.updateMomentLM(dt, A, B, C, self._halves(LU, p), self._halves(LM, p),
                             self._halves(U, p), nstages, stream)

    def updateDistLM(self, dt, A, B, C, L, M, F, fnew, U, nstages, stream=None):
        for p, m in enumerate(self.models):
            m.updateDistLM(dt, A, B, C, self._halves(L, p), self._halves(M, p),
                           self._halves(F, p), self._half(fnew, p), self._halves(U, p), nstages,
                           stream)

    def updateMomentARS(self, dt, ah, ai, LU, U, stream=None):
        for p, m in enumerate(self.models):
            m.updateMomentARS(dt, ah, ai, self._halves(LU, p), self._halves(U, p), stream)

    def updateDistARS(self, dt, ah, ai, L, U, M, F, stream=None):
        for p, m in enumerate(self.models):
            m.updateDistARS(dt, ah, ai, self._halves(L, p), self._halves(U, p),
                            self._halves(M, p), self._halves(F, p), stream)


def species_penalties(cfg, vm, nslice, device):
    """MixturePenalty of [penalized-scattering-model] type = shakov-direct-gll."""
    kind = cfg.lookup('penalized-scattering-model', 'type')
    if kind != 'shakov-direct-gll':
        raise ValueError("the mixture's penalized-scattering-model::type must be "
                         "'shakov-direct-gll' (got %r)" % kind)
    models = [DGFSShakovSpeciesPenalty(cfg, vm, p, Ne=None, device=device)
              for p in range(vm.nspcs())]
    return MixturePenalty(models, nslice)
