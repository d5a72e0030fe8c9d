"""Distribution writer / restart of the binary mixture (reference
dgfs1D/bi/distribution.py:9-44, bi/bi.py:222-236): HDF5 datasets ``coeff0``,
``coeff1`` (flat, K*Ne*Nv each) with attributes time, K, Ne, Nv, written and
read by ``dgfs1d_gpu_amd.h5``; file names basename + ``.h5_rank=0`` as the
reference's single-rank runs.  Restart also accepts ``.npz`` files of the
same keys."""
import os

from ..std.distribution import read_datasets, write_distribution


class DGFSDistributionBi:
    def __init__(self, tcurr, shape, cfg, cfgsect, extn='.h5'):
        self.K, self.Ne, self.Nv = shape
        self.basedir = cfg.lookuppath(cfgsect, 'basedir', '.', abs=True)
        self.basename = cfg.lookup(cfgsect, 'basename')
        if not self.basename.endswith(extn):
            self.basename += extn
        self.basename += '_rank=0'
        self.dt_out = float(cfg.lookupfloat(cfgsect, 'dt-out'))
        self.tout_next = tcurr + self.dt_out

    def __call__(self, dt, tcurr, coeffs):
        if abs(self.tout_next - tcurr) > 0.5 * dt:
            return
        fname = os.path.join(self.basedir, self.basename.format(t=tcurr))
        write_distribution(fname, [c.cpu().numpy() for c in coeffs], tcurr, self.K, self.Ne,
                           self.Nv)
        self.tout_next = tcurr + self.dt_out


def load_distribution(path, K, Ne, Nv, nspcs):
    """(time, [flat coeff per species]) of a written distribution."""
    return read_datasets(path, ['coeff%d' % p for p in range(nspcs)], K, Ne, Nv)
