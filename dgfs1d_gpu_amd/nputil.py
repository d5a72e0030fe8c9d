"""Host utilities: plugin registry, restricted expression evaluation.

Mirrors the semantics of the reference's registry and expression helpers
(dgfs1D/nputil.py:22-58): plugins are selected by a class attribute equal to
the INI string; expressions may use a fixed set of NumPy functions, no
attribute access and no '**'/'^'.
"""
import itertools as it
import re

import numpy as np

_NPEVAL_SYMS = {
    '__builtins__': None,
    'exp': np.exp, 'log': np.log, 'sin': np.sin, 'asin': np.arcsin,
    'cos': np.cos, 'acos': np.arccos, 'tan': np.tan, 'atan': np.arctan,
    'atan2': np.arctan2, 'abs': np.abs, 'pow': np.power, 'sqrt': np.sqrt,
    'tanh': np.tanh, 'pi': np.pi, 'linspace': np.linspace, 'logspace': np.logspace,
}


def npeval(expr, local_vars):
    """Evaluate a restricted arithmetic expression (nputil.py:22-36 semantics)."""
    if '^' in expr or '**' in expr:
        raise ValueError('powers are written pow(a, b); the operators ** and ^ are refused')
    if not re.match(r'[A-Za-z0-9 \t\n\r.,+\-*/%()<>=]+$', expr):
        raise ValueError('expression %r holds characters outside the arithmetic set' % expr)
    names = '|'.join(re.escape(k) for k in it.chain(_NPEVAL_SYMS, local_vars))
    if re.search(r'(%s|\))\s*\.' % names, expr):
        raise ValueError('expression %r reads an attribute, which is refused' % expr)
    return eval(expr, dict(_NPEVAL_SYMS), dict(local_vars))  # noqa: S307 (restricted)


def ndrange(*args):
    return it.product(*map(range, args))


def subclasses(cls):
    out = []
    for s in cls.__subclasses__():
        out.append(s)
        out.extend(subclasses(s))
    return out


def subclass_where(cls, **kwargs):
    """First subclass (depth-first) whose attribute equals the value."""
    k, v = next(iter(kwargs.items()))
    for s in subclasses(cls):
        if getattr(s, k, None) == v:
            return s
    raise KeyError('no plugin derived from %s has %s = %r' % (cls.__name__, k, v))
