"""Spherical point sets for the fast-spectral collision operator.

* Stored rules: ``<name>-ss<deg>-m<npts>.txt`` tables in ``symmetric/``.  The
  shipped tables are the published Womersley antipodally symmetric spherical
  designs (the same data the reference ships, loaded the same way:
  dgfs1D/sphericaldesign/__init__.py:174-210).
* Tabulated rules passed inline as ``ssrule`` text (``(x, y, z)`` per line,
  :138-171, :216-230).
* ``octahedral``: a 24-point antipodal spherical 5-design built here, the
  O_h orbit of (a, a, b) with 2a^4 + b^4 = 3/5, 2a^2 + b^2 = 1.  It supplies
  the M = 12 half-sphere (first 12 points, z > 0; the rest are their
  negatives), for which no Womersley table is shipped.
"""
import os
import re

import numpy as np

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'symmetric')


def octahedral24():
    s = (4.0 - np.sqrt(16.0 - 9.6)) / 12.0          # a^2 root of 6 s^2 - 4 s + 0.4 = 0
    a, b = np.sqrt(s), np.sqrt(1.0 - 2.0 * s)
    half = []
    for base in ((a, a, b), (a, b, a), (b, a, a)):
        for sx in (1.0, -1.0):
            for sy in (1.0, -1.0):
                half.append((sx * base[0], sy * base[1], base[2]))
    half = np.array(half)
    return np.vstack([half, -half])


def _stored(name, npts):
    best = None
    for fn in os.listdir(_DIR):
        m = re.match(r'([a-zA-Z0-9\-~+]+)-ss(\d+)(?:-m(\d+))?\.txt$', fn)
        if not m:
            continue
        if (not name or name == m.group(1)) and int(m.group(3) or -1) == npts:
            if best is None or int(m.group(2)) > best[1]:
                best = (fn, int(m.group(2)))
    if best is None:
        raise ValueError('No suitable spherical quadrature rule found')
    return np.loadtxt(os.path.join(_DIR, best[0]), dtype=np.float64).reshape(-1, 3)


def _tabulated(rule):
    rule = re.sub(r'(?<=\))\s*,?\s*(?!$)', r'\n', rule)
    rule = re.sub(r'\(|\)|,', '', rule).strip()
    rule = rule[1:-1] if rule.startswith('[') else rule
    pts = []
    for line in rule.splitlines():
        if not line.strip():
            continue
        vals = [float(x) for x in line.split()]
        if len(vals) not in (3, 4):
            raise ValueError('Invalid points in quadrature rule')
        pts.append(vals[:3])
    return np.array(pts, dtype=np.float64)


def get_sphquadrule(sstype, rule=None, npts=None):
    """Return the (npts, 3) point set named by ``rule`` (symmetric rules only)."""
    if sstype != 'symmetric':
        raise ValueError('only symmetric spherical rules are supported')
    if rule == 'octahedral':
        pts = octahedral24()
    elif rule and not re.match(r'[a-zA-z0-9\-~+]+$', rule):
        pts = _tabulated(rule)
    else:
        return _stored(rule, npts)
    if npts and npts != len(pts):
        raise ValueError('Invalid number of points in provided rule')
    return pts
