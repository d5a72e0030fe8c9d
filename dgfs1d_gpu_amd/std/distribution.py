"""Distribution writer / restart (reference dgfs1D/std/distribution.py:9-44,
std/std.py:235-247).

The files are HDF5, of the reference's shape: a flat dataset ``coeff`` of
K*Ne*Nv values in the run's precision, with scalar attributes time, K, Ne and
Nv, written and read by ``dgfs1d_gpu_amd.h5`` (h5py is not needed).  File
names follow the reference: basename + ``.h5`` + ``_rank=<rank>``, rank 0 for
this single-process solver.  Restart also accepts the ``.npz`` files (same
keys) that earlier versions of this package wrote.
"""
import os

import numpy as np

from .. import h5


class DGFSDistributionStd:
    def __init__(self, tcurr, shape, cfg, cfgsect, extn='.h5'):
        self.K, self.Ne, self.Nv = shape
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith(extn):
            self.basename += extn
        self.basename += '_rank=0'
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr + self.dt_out

    def __call__(self, dt, tcurr, coeff):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        write_distribution(fname, [coeff.cpu().numpy()], tcurr, self.K, self.Ne, self.Nv,
                           single=True)
        self.tout_next = tcurr + self.dt_out


def write_distribution(path, coeffs, tcurr, K, Ne, Nv, single=False):
    """One dataset per species: ``coeff`` (single=True) or ``coeff0``,
    ``coeff1``, ..., each flat, with attributes time, K, Ne, Nv."""
    attrs = {'time': float(tcurr), 'K': int(K), 'Ne': int(Ne), 'Nv': int(Nv)}
    names = ['coeff'] if single else ['coeff%d' % p for p in range(len(coeffs))]
    h5.write(path, {n: (np.ascontiguousarray(c).ravel(), attrs) for n, c in zip(names, coeffs)})


def read_datasets(path, names, K, Ne, Nv):
    """(time, [flat arrays]) of the named datasets; ValueError on a K/Ne/Nv
    mismatch, with the reference's messages (std/std.py:243-245)."""
    if h5.is_hdf5(path):
        data = h5.read(path)
        missing = [n for n in names if n not in data]
        if missing:
            raise ValueError('no dataset %s in %s' % (', '.join(missing), path))
        out, ti = [], None
        for n in names:
            arr, attrs = data[n]
            ti = float(attrs['time'])
            _check(attrs, K, Ne, Nv)
            out.append(np.array(arr).ravel())
        return ti, out
    with np.load(path, allow_pickle=False) as d:
        _check(d, K, Ne, Nv)
        return float(d['time']), [np.array(d[n]).ravel() for n in names]


def _check(attrs, K, Ne, Nv):
    for k, v in (('K', K), ('Ne', Ne), ('Nv', Nv)):
        if int(attrs[k]) != v:
            raise ValueError('Inconsistent distribution %s' % ('N' if k == 'Nv' else k))


def load_distribution(path, K, Ne, Nv):
    """(time, flat coeff) of a written distribution."""
    ti, (coeff,) = read_datasets(path, ['coeff'], K, Ne, Nv)
    return ti, coeff
