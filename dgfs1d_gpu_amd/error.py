"""L1 difference between a nodal DG solution on Ne elements and one on 2 Ne
elements (reference dgfs1D/error.py:9-48, ``dgfsStdError1D``), for the
distribution files the solvers write (HDF5 with dataset coeff and attributes
K, Ne, Nv, as the reference's; ``.npz`` files of the same keys also work).

    python -m dgfs1d_gpu_amd.error coarse.h5_rank=0 fine.h5_rank=0

The coarse solution's nodal polynomial on each element (GLL nodes) is
evaluated at the GLL nodes of its two halves and compared with the fine
solution there: 0.5 sum_q w_q |f_coarse - f_fine| per fine element and
velocity point, averaged over both.  (The reference prints the result with
``print(*value)`` of a scalar, which raises; the value is printed here.)
"""
import argparse
import sys
from pathlib import Path

import numpy as np

from . import h5
from .quadratures import nodal_basis_at, zwglj


def decode(path):
    """(Ne, K, coeff (K, Ne, Nv)) of a written distribution (error.py:9-14):
    an HDF5 file of the reference's layout, or a ``.npz`` of the same keys."""
    if h5.is_hdf5(path):
        arr, d = h5.read(path)['coeff']
    else:
        with np.load(path, allow_pickle=False) as z:
            d = {k: z[k] for k in ('K', 'Ne', 'Nv')}
            arr = np.array(z['coeff'])
    K, Ne, Nv = int(d['K']), int(d['Ne']), int(d['Nv'])
    return Ne, K, np.asarray(arr, dtype=np.float64).reshape(K, Ne, Nv)


def compare(K, f, f2):
    """f (K, Ne, Nv) on Ne elements, f2 (K, 2 Ne, Nv) on their halves."""
    _, Ne2, nv = f2.shape
    r, w = zwglj(K, 0.0, 0.0)
    half = 0.5 * r + 0.5                                   # nodes of the right half
    pts = np.concatenate([-half[::-1], half])              # both halves, left to right
    im = nodal_basis_at(K, r, pts)                         # (2K, K)
    fr = np.einsum('rm,mek->rek', im, f)                   # (2K, Ne, Nv)
    # element e's halves are fine elements 2e and 2e + 1
    fr = fr.reshape(2, K, f.shape[1], nv).transpose(1, 2, 0, 3).reshape(K, Ne2, nv)
    l1 = 0.5 * np.einsum('q,qek->ek', w, np.abs(fr - f2))
    return float(l1.sum() / Ne2 / nv)


def compute_error(file1, file2):
    Ne1, K1, u1 = decode(file1)
    Ne2, K2, u2 = decode(file2)
    if Ne2 != 2 * Ne1:
        raise ValueError('Elements in file2 != 2 * (elements in file1)')
    if K1 != K2:
        raise ValueError("Number of quadrature points don't match")
    return compare(K1, u1, u2)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('file1', type=Path, help='solution on Ne elements')
    ap.add_argument('file2', type=Path, help='solution on 2 Ne elements')
    args = ap.parse_args(argv)
    if not (args.file1.exists() and args.file2.exists()):
        raise SystemExit('Input files not found')
    print(compute_error(args.file1, args.file2))
    return 0


if __name__ == '__main__':
    sys.exit(main())
