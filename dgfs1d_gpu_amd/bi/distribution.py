"""Distribution writer / restart of the binary mixture (reference
dgfs1D/bi/distribution.py:9-44, bi/bi.py:190-200).  The reference writes HDF5
datasets coeff0, coeff1 with attributes time, K, Ne, Nv; h5py is not part of
this stack, so the same payload is written as an uncompressed ``.npz``."""
import os

import numpy as np


class DGFSDistributionBi:
    def __init__(self, tcurr, shape, cfg, cfgsect, extn='.npz'):
        self.K, self.Ne, self.Nv = shape
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith(extn):
            self.basename += extn
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr + self.dt_out

    def __call__(self, dt, tcurr, coeffs):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        data = {'coeff%d' % p: c.cpu().numpy() for p, c in enumerate(coeffs)}
        np.savez(fname, time=tcurr, K=self.K, Ne=self.Ne, Nv=self.Nv, **data)
        self.tout_next = tcurr + self.dt_out


def load_distribution(path, K, Ne, Nv, nspcs):
    with np.load(path, allow_pickle=False) as d:
        for k, v in (('K', K), ('Ne', Ne), ('Nv', Nv)):
            if int(d[k]) != v:
                raise ValueError('Inconsistent distribution %s' % ('N' if k == 'Nv' else k))
        return float(d['time']), [np.array(d['coeff%d' % p]) for p in range(nspcs)]
