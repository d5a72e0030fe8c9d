"""Scattering model of the binary mixture (reference dgfs1D/bi/scattering.py).

``vss`` (DGFSBiVSSScatteringModel, :47-321) keeps the reference's deck keys
(``ssrulepre``, ``Mpre``, ``alpha<pq>``, ``omega<pq>``, ``dRef<pq>``,
``Tref<pq>`` for p, q in {0, 1}) and prefactors; the kernels b_pq, c_pq are
built on the device by one plan per ordered pair.  The solver uses the batched
``collide(p, us, Q)``: Q += sum_q Q_pq(U u_p, V u_q) for every (m, r, e) of
one RHS (bi.py:329-337), one launch sequence per partner q.
"""
import logging
from math import gamma
import itertools as it

import numpy as np

from .. import _lib
from ..collision import BiCollisionPlan, task_table
from ..sphericaldesign import get_sphquadrule

log = logging.getLogger('dgfs1d_gpu_amd')


class DGFSScatteringModelBi:
    scattering_model = None

    def __init__(self, cfg, velocitymesh, **kwargs):
        self.cfg = cfg
        self.vm = velocitymesh
        self._nspcs = velocitymesh.nspcs()
        self._Ne = kwargs.get('Ne')
        self._device = kwargs.get('device')
        self._ntask_max = kwargs.get('ntask_max')
        self.load_parameters()
        self.perform_precomputation()


class DGFSBiVSSScatteringModel(DGFSScatteringModelBi):
    scattering_model = 'vss'

    def load_parameters(self):
        cfg, vm = self.cfg, self.vm
        ssrulepre = cfg.lookup('scattering-model', 'ssrulepre')
        self._Mpre = cfg.lookupint('scattering-model', 'Mpre')
        self._szpre = get_sphquadrule('symmetric', rule=ssrulepre, npts=self._Mpre)
        self._cases = list(it.product(range(self._nspcs), repeat=2))
        masses = vm.masses()
        self._alpha, self._omega, self._prefactor, self._invKn = {}, {}, {}, {}
        for p, q in self._cases:
            pq = '%d%d' % (p, q)
            alpha = float(cfg.lookupfloat('scattering-model', 'alpha' + pq))
            omega = float(cfg.lookupfloat('scattering-model', 'omega' + pq))
            dRef = float(cfg.lookupfloat('scattering-model', 'dRef' + pq))
            Tref = float(cfg.lookupfloat('scattering-model', 'Tref' + pq))
            mp, mq = masses[p], masses[q]
            invKn = (np.sqrt(1.0 + mp / mq) * np.pi * vm.n0() * dRef * dRef
                     * (Tref / vm.T0()) ** (omega - 0.5) * vm.H0())          # :78-79
            mu = mp * mq / (mp + mq)
            self._prefactor[pq] = invKn * alpha / (                        # :83-86
                np.sqrt(1.0 + mp / mq) * mu ** (omega - 0.5) * 2.0 ** (1 + alpha)
                * gamma(2.5 - omega) * np.pi)
            self._alpha[pq], self._omega[pq], self._invKn[pq] = alpha, omega, invKn
            log.info('Kn%s: %s prefactor%s: %s', pq, 1.0 / invKn, pq, self._prefactor[pq])

    def perform_precomputation(self):
        vm = self.vm
        dt = _lib.F64 if self.cfg.dtype == np.float64 else _lib.F32
        masses = vm.masses()
        self.plans = {}
        for p, q in self._cases:
            pq = '%d%d' % (p, q)
            self.plans[p, q] = BiCollisionPlan(
                vm.Nv(), vm.M(), vm.Nrho(), vm.L(), vm.qz(), vm.qw(), vm.sz(), vm.sw(),
                self._szpre, masses[p], masses[q], self._alpha[pq], self._omega[pq],
                self._prefactor[pq], dt, max(1, int(self._ntask_max or 1)), device=self._device)
        self.tasks = None

    def prefactor(self, pq):
        return self._prefactor[pq]

    def set_tasks(self, K, Ne, tasks):
        td = self.plans[0, 0].tdtype
        self.tasks = task_table(K, Ne, tasks, td, self._device)
        for pl in self.plans.values():
            if self.tasks['ntask'] > pl.ntask_max:
                raise ValueError('too many collision tasks')
        return self.tasks

    def collide(self, p, us, Q, stream=None):
        """Q += sum_q Q_pq(U u_p, V u_q) over every task (bi.py:329-337)."""
        for q in range(self._nspcs):
            self.plans[p, q](us[p], us[q], Q, stream=stream, tasks=self.tasks)
