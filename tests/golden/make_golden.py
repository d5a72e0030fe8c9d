"""Write the golden trajectory of the small Couette deck with the FP64 oracle.

    python tests/golden/make_golden.py

Output tests/golden/couette_small_oracle.npz:
  steps      steps at which moments are stored (1, 13, 26, 39, 52)
  moments    (len(steps), Nqr, Ne, 9) non-dimensional rho ux uy T qx qy Pxx Pyy Pxy
  residual   (52,) sqrt(sum (f^{n+1}-f^n)^2 / sum (f^n)^2) per step
  u_final    (K, Ne, Nv) coefficients after 52 forward-Euler steps
The deck is dgfs1d_gpu_amd/configs/couette_small.ini (read as plain text).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import std as OS  # noqa: E402
from oracle.deck import Deck  # noqa: E402


def main():
    text = open(os.path.join(ROOT, 'dgfs1d_gpu_amd', 'configs', 'couette_small.ini')).read()
    d = Deck(text)
    nsteps = int(round((d.tend - d.t0) / d.dt))
    keep = (1, 13, 26, 39, 52)
    u, t = d.u0.copy(), d.t0
    moms, res = [], []
    for s in range(1, nsteps + 1):
        un = d.step(t, u)
        res.append(OS.residual(un, u))
        u, t = un, t + d.dt
        if s in keep:
            moms.append(OS.moments(u, d.basis.interp, d.vm))
    np.savez_compressed(os.path.join(HERE, 'couette_small_oracle.npz'), steps=np.array(keep),
                        moments=np.array(moms), residual=np.array(res), u_final=u)
    print('nsteps', nsteps, 'final residual', res[-1])


if __name__ == '__main__':
    main()
