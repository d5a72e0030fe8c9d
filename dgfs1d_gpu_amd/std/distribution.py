"""Distribution writer / restart (reference dgfs1D/std/distribution.py:9-44,
std/std.py:235-247).  The reference writes HDF5 (h5py); h5py is not part of
this stack, so the same payload -- dataset ``coeff`` (K*Ne*Nv) with attributes
time, K, Ne, Nv -- is written as an uncompressed ``.npz``."""
import os

import numpy as np


class DGFSDistributionStd:
    def __init__(self, tcurr, shape, cfg, cfgsect, extn='.npz'):
        self.K, self.Ne, self.Nv = shape
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith(extn):
            self.basename += extn
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr + self.dt_out

    def __call__(self, dt, tcurr, coeff):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        np.savez(fname, coeff=coeff.cpu().numpy(), time=tcurr, K=self.K, Ne=self.Ne, Nv=self.Nv)
        self.tout_next = tcurr + self.dt_out


def load_distribution(path, K, Ne, Nv):
    """Read a written distribution; raise ValueError on a K/Ne/Nv mismatch."""
    with np.load(path, allow_pickle=False) as d:
        for k, v in (('K', K), ('Ne', Ne), ('Nv', Nv)):
            if int(d[k]) != v:
                raise ValueError('Inconsistent distribution %s' % ('N' if k == 'Nv' else k))
        return float(d['time']), np.array(d['coeff'])
