"""Golden HDF5 distribution files, written by h5py itself.

The reference writes its distribution files with h5py
(dgfs1D/std/distribution.py:36-41, dgfs1D/bi/distribution.py:36-42): a flat
dataset ``coeff`` (or ``coeff0``, ``coeff1`` for the binary mixture) of
K*Ne*Nv values in the configured precision, with scalar attributes time (a
Python float), K, Ne, Nv (Python ints).  This script writes files of exactly
that shape with h5py so that the package's own HDF5 reader is checked against
the library the reference uses.  It needs an interpreter that has h5py:

    /opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py

With ``--dump FILE`` it prints what h5py reads from FILE as JSON (datasets:
dtype, shape, values; attributes), which the tests use, when such an
interpreter is present, to check files the package writes.
"""
import json
import os
import sys

import h5py
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'h5')


def payload(K, Ne, Nv, dtype, shift=0.0):
    """The coeff values of every fixture: a closed form of the flat index."""
    i = np.arange(K * Ne * Nv, dtype=np.float64)
    return (np.sin(0.37 * i + shift) + 1e-3 * i).astype(dtype)


def write(name, K, Ne, Nv, dtype, nspcs=0, tcurr=0.0125, libver=None, **kw):
    path = os.path.join(HERE, name)
    with h5py.File(path, 'w', libver=libver) as h5f:
        names = ['coeff'] if nspcs == 0 else ['coeff%d' % p for p in range(nspcs)]
        for p, dname in enumerate(names):
            dst = h5f.create_dataset(dname, data=payload(K, Ne, Nv, dtype, shift=0.5 * p), **kw)
            dst.attrs['time'] = tcurr
            dst.attrs['K'] = K
            dst.attrs['Ne'] = Ne
            dst.attrs['Nv'] = Nv


def dump(path):
    out = {}
    with h5py.File(path, 'r') as h5f:
        for name, dst in h5f.items():
            attrs = {}
            for k, v in dst.attrs.items():
                attrs[k] = [str(np.asarray(v).dtype), np.asarray(v).item()]
            out[name] = {'dtype': str(dst.dtype), 'shape': list(dst.shape),
                         'values': np.asarray(dst[()], dtype=np.float64).ravel().tolist(),
                         'attrs': attrs}
    print(json.dumps(out))


def main():
    if len(sys.argv) == 3 and sys.argv[1] == '--dump':
        dump(sys.argv[2])
        return
    os.makedirs(HERE, exist_ok=True)
    write('std_k3_ne2_nv64.h5', 3, 2, 64, np.float64)
    write('std_k3_ne2_nv64_f32.h5', 3, 2, 64, np.float32)
    write('bi_k2_ne3_nv27.h5', 2, 3, 27, np.float64, nspcs=2, tcurr=1.5)
    # the newest file format h5py can write (object header v2, link messages)
    write('std_k2_ne2_nv8_latest.h5', 2, 2, 8, np.float64, libver='latest')
    # the size of a modest run's file (K=3, Ne=4, Nv=6^3)
    write('std_k3_ne4_nv216.h5', 3, 4, 216, np.float64, tcurr=3.0)
    # compressed storage, which the package's reader refuses
    write('std_k2_ne2_nv8_gzip.h5', 2, 2, 8, np.float64, compression='gzip')


if __name__ == '__main__':
    main()
