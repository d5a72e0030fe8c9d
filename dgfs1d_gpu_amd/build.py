"""Build libdgfs_hip.so (gfx950) from dgfs1d_gpu_amd/csrc with hipcc.

Each ``*.hip`` translation unit is compiled in parallel to an object under
``dgfs1d_gpu_amd/_build/`` and the objects are linked into
``dgfs1d_gpu_amd/libdgfs_hip.so`` (in-tree, so the built library travels with
the source tree).

    python -m dgfs1d_gpu_amd.build [--force] [-j JOBS]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, 'csrc')
BUILD = os.path.join(PKG, '_build')
LIB = os.path.join(PKG, 'libdgfs_hip.so')
ARCH = 'gfx950'

HIPCC_FLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
    # packed FP32 instructions are kept out of every kernel
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I', os.path.join(ROOT, 'include'), '-I', CSRC,
    '-Wno-unused-command-line-argument',
]


def _hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the HIP toolchain is required to build libdgfs_hip.so')


def _sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.h')]
    inc = os.path.join(ROOT, 'include')
    hs += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith('.h')]
    return hs


def _compile(src, force, hdr_mtime):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + '.o')
    if (not force and os.path.exists(obj)
            and os.path.getmtime(obj) >= max(os.path.getmtime(src), hdr_mtime)):
        return obj, None
    cmd = [_hipcc()] + HIPCC_FLAGS + ['-c', src, '-o', obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        return obj, '$ ' + ' '.join(cmd) + '\n' + p.stdout + p.stderr
    return obj, None


def build(force=True, jobs=None, verbose=False):
    """Compile every HIP source and link the shared library. Raises on error."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    hdr_mtime = max(os.path.getmtime(h) for h in _headers())
    jobs = jobs or min(16, os.cpu_count() or 4)
    errors, objs = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, force, hdr_mtime), srcs):
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError('HIP compilation failed:\n' + '\n'.join(errors))
    tmp = LIB + '.tmp'
    cmd = [_hipcc(), '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError('link failed:\n' + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--force', action='store_true')
    ap.add_argument('-j', '--jobs', type=int, default=None)
    a = ap.parse_args(argv)
    build(force=a.force, jobs=a.jobs, verbose=True)


if __name__ == '__main__':
    sys.exit(main())
