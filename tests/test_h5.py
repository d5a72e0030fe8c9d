"""The package's HDF5 distribution files (dgfs1d_gpu_amd/h5.py) against h5py.

Golden files under tests/golden/h5/ were written by h5py 3.3.0 on HDF5
1.10.6 (tests/golden/make_h5_fixtures.py) in the reference's layout
(dgfs1D/std/distribution.py:36-41, dgfs1D/bi/distribution.py:36-42); their
values are a closed form of the flat index, so reading them is checked bit
for bit.  Files the package writes are read back by h5py itself and by
h5dump when an interpreter with h5py is present (skipped otherwise), and by
the package's own reader always.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from dgfs1d_gpu_amd import h5

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, 'tests', 'golden', 'h5')
H5PY_PY = '/opt/conda/bin/python3.9'
H5DUMP = '/opt/conda/bin/h5dump'


def payload(K, Ne, Nv, dtype, shift=0.0):
    # the closed form of tests/golden/make_h5_fixtures.py:payload
    i = np.arange(K * Ne * Nv, dtype=np.float64)
    return (np.sin(0.37 * i + shift) + 1e-3 * i).astype(dtype)


GOLDEN = [('std_k3_ne2_nv64.h5', 3, 2, 64, np.float64, 0, 0.0125),
          ('std_k3_ne2_nv64_f32.h5', 3, 2, 64, np.float32, 0, 0.0125),
          ('bi_k2_ne3_nv27.h5', 2, 3, 27, np.float64, 2, 1.5),
          ('std_k2_ne2_nv8_latest.h5', 2, 2, 8, np.float64, 0, 0.0125),
          ('std_k3_ne4_nv216.h5', 3, 4, 216, np.float64, 0, 3.0)]


@pytest.mark.parametrize('name,K,Ne,Nv,dt,nspcs,t', GOLDEN)
def test_reads_h5py_files(name, K, Ne, Nv, dt, nspcs, t):
    data = h5.read(os.path.join(GOLD, name))
    names = ['coeff'] if nspcs == 0 else ['coeff%d' % p for p in range(nspcs)]
    assert sorted(data) == names
    for p, n in enumerate(names):
        arr, attrs = data[n]
        assert arr.dtype == dt and arr.shape == (K * Ne * Nv,)
        assert np.array_equal(arr, payload(K, Ne, Nv, dt, shift=0.5 * p))
        assert attrs == {'time': t, 'K': K, 'Ne': Ne, 'Nv': Nv}
        assert np.asarray(attrs['K']).dtype == np.int64
        assert np.asarray(attrs['time']).dtype == np.float64


def test_refuses_what_it_cannot_read(tmp_path):
    with pytest.raises(ValueError, match='filtered'):
        h5.read(os.path.join(GOLD, 'std_k2_ne2_nv8_gzip.h5'))
    bad = tmp_path / 'x.h5'
    bad.write_bytes(b'not hdf5 at all' * 10)
    with pytest.raises(ValueError, match='not an HDF5 file'):
        h5.read(str(bad))
    with pytest.raises(ValueError, match='unsupported dtype'):
        h5.write(str(tmp_path / 'c.h5'), {'coeff': (np.zeros(4, np.complex128), {})})


def _sample():
    a = payload(3, 2, 64, np.float64)
    attrs = {'time': 0.0125, 'K': 3, 'Ne': 2, 'Nv': 64}
    return {'coeff0': (a.astype(np.float32), attrs), 'coeff1': (-a, dict(attrs, time=1.5))}


def test_writer_round_trip(tmp_path):
    path = str(tmp_path / 'w.h5')
    src = _sample()
    h5.write(path, src)
    assert h5.is_hdf5(path)
    got = h5.read(path)
    assert sorted(got) == sorted(src)
    for n, (arr, attrs) in src.items():
        assert got[n][0].dtype == arr.dtype
        assert np.array_equal(got[n][0], arr)
        assert got[n][1] == attrs
    # the same names and values as h5py's own file of this layout
    ref = h5.read(os.path.join(GOLD, 'std_k3_ne2_nv64.h5'))['coeff']
    h5.write(path, {'coeff': ref})
    again = h5.read(path)['coeff']
    assert np.array_equal(again[0], ref[0]) and again[1] == ref[1]


def _have_h5py():
    if not os.path.exists(H5PY_PY):
        return False
    r = subprocess.run([H5PY_PY, '-c', 'import h5py'], capture_output=True, timeout=120)
    return r.returncode == 0


@pytest.mark.skipif(not _have_h5py(), reason='no interpreter with h5py')
def test_h5py_reads_written_files(tmp_path):
    path = str(tmp_path / 'w.h5')
    src = _sample()
    h5.write(path, src)
    out = subprocess.run([H5PY_PY, os.path.join(ROOT, 'tests', 'golden', 'make_h5_fixtures.py'),
                          '--dump', path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    assert sorted(got) == sorted(src)
    for n, (arr, attrs) in src.items():
        g = got[n]
        assert g['dtype'] == str(arr.dtype) and g['shape'] == list(arr.shape)
        assert np.array_equal(np.array(g['values']), arr.astype(np.float64))
        assert {k: v[1] for k, v in g['attrs'].items()} == attrs
        assert {k: v[0] for k, v in g['attrs'].items()} == {
            'time': 'float64', 'K': 'int64', 'Ne': 'int64', 'Nv': 'int64'}


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason='no h5dump')
def test_h5dump_reads_written_files(tmp_path):
    path = str(tmp_path / 'w.h5')
    h5.write(path, _sample())
    out = subprocess.run([H5DUMP, '-p', '-A', path], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    for s in ('DATASET "coeff0"', 'H5T_IEEE_F32LE', 'DATASET "coeff1"', 'H5T_IEEE_F64LE',
              'ATTRIBUTE "Nv"', 'H5T_STD_I64LE', 'CONTIGUOUS', 'SIMPLE { ( 384 ) / ( 384 ) }'):
        assert s in out.stdout, s


def test_distribution_writer_and_restart(tmp_path):
    """DGFSDistributionStd writes basename.h5_rank=0 at the output times; the
    restart reader returns (time, coeff) and rejects a mismatched K/Ne/Nv with
    the reference's messages (std/std.py:243-245); an h5py file restarts."""
    import torch
    from dgfs1d_gpu_amd.dictionary import Dictionary
    from dgfs1d_gpu_amd.std.distribution import DGFSDistributionStd, load_distribution
    cfg = Dictionary('[w]\ndt-out = 0.5\nbasedir = %s\nbasename = dist-{t:.1f}\n' % tmp_path)
    w = DGFSDistributionStd(0.0, (3, 2, 64), cfg, 'w')
    u = torch.tensor(payload(3, 2, 64, np.float64))
    w(0.1, 0.2, u)                                       # not an output time
    assert not list(tmp_path.iterdir())
    w(0.1, 0.5, u)
    path = tmp_path / 'dist-0.5.h5_rank=0'
    assert path.exists()
    t, c = load_distribution(str(path), 3, 2, 64)
    assert t == 0.5 and np.array_equal(c, u.numpy())
    for shape, msg in (((2, 2, 64), 'K'), ((3, 4, 64), 'Ne'), ((3, 2, 8), 'N')):
        with pytest.raises(ValueError, match='Inconsistent distribution %s$' % msg):
            load_distribution(str(path), *shape)
    t, c = load_distribution(os.path.join(GOLD, 'std_k3_ne2_nv64.h5'), 3, 2, 64)
    assert t == 0.0125 and np.array_equal(c, payload(3, 2, 64, np.float64))


def test_bi_distribution_reads_h5py_file():
    from dgfs1d_gpu_amd.bi.distribution import load_distribution
    t, cs = load_distribution(os.path.join(GOLD, 'bi_k2_ne3_nv27.h5'), 2, 3, 27, 2)
    assert t == 1.5
    for p in range(2):
        assert np.array_equal(cs[p], payload(2, 3, 27, np.float64, shift=0.5 * p))
    with pytest.raises(ValueError, match='no dataset coeff2'):
        load_distribution(os.path.join(GOLD, 'bi_k2_ne3_nv27.h5'), 2, 3, 27, 3)


def test_error_tool_reads_h5(tmp_path):
    """The convergence tool gives the same number on HDF5 and .npz files."""
    from dgfs1d_gpu_amd.error import compute_error
    from dgfs1d_gpu_amd.quadratures import zwglj
    from dgfs1d_gpu_amd.std.distribution import write_distribution
    r, _ = zwglj(3, 0.0, 0.0)
    for Ne in (4, 8):
        x = (np.arange(Ne)[None, :] + 0.5 * (r[:, None] + 1.0)) / Ne
        coeff = np.stack([np.sin(3 * x + v) for v in range(3)], axis=-1)
        write_distribution(str(tmp_path / ('%d.h5' % Ne)), [coeff], 0.0, 3, Ne, 3, single=True)
        np.savez(tmp_path / ('%d.npz' % Ne), coeff=coeff.ravel(), time=0.0, K=3, Ne=Ne, Nv=3)
    e_h5 = compute_error(tmp_path / '4.h5', tmp_path / '8.h5')
    e_npz = compute_error(tmp_path / '4.npz', tmp_path / '8.npz')
    assert e_h5 == e_npz and 0 < e_h5 < 1e-2
