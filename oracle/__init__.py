"""FP64 CPU oracle for the DGFS 1D3V single-species operator path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call or
execute anything in this directory.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker or
as the timed CPU baseline -- never as a code path of the product.

The oracle is a NumPy restatement, in FP64, of the algorithms of the reference
solver (jaisw7/dgfs1D_gpu, the ``dgfsStd1D`` path).  Each function cites the
reference ``file:line`` it follows.  The reference itself cannot run here (it
is PyCUDA + cuFFT only), so the oracle is pinned independently by brute force,
closed forms and invariants (see ``tests/test_oracle_*.py``):

* ``collision``: pinned by the brute-force spectral convolution sum at N <= 8
  (independent of any FFT), by conservation/equilibrium invariants and by
  bilinearity.
* ``quadratures``: pinned by ``numpy.polynomial.legendre.leggauss`` and the
  closed-form Gauss-Lobatto rules.
* ``basis``: pinned by polynomial exactness (derivative, mass, interpolation).
* ``std`` (advection/BC/integrators/moments): pinned by polynomial exactness,
  the free-molecular wall mass balance, scalar-ODE amplification factors and
  closed-form Maxwellian moments.

Functions whose behaviour only the running reference could fix bit-for-bit
(summation order of GiMMiK/cuFFT/PyCUDA reductions) are compared at tolerance,
never bitwise.
"""
