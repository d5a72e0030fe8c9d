"""INI configuration (same schema and lookup semantics as the reference's
dgfs1D/dictionary.py:18-137).  Uses ConfigParser (SafeConfigParser is gone
from Python 3.12), case-preserving keys and ';' inline comments."""
import io
import json
import os
import re
from configparser import ConfigParser

import numpy as np

_NP_MAP = {'float': np.float32, 'single': np.float32, 'double': np.float64}


def _ensure_float(m):
    m = m.group(0)
    return m if any(c in m for c in '.eE') else m + '.'


class Dictionary:
    def __init__(self, inistr=None, defaults=None):
        self._cp = cp = ConfigParser(inline_comment_prefixes=[';'])
        cp.optionxform = str
        if inistr:
            cp.read_string(inistr)
        if defaults:
            cp.read_dict(defaults)
        self._derive()

    def _derive(self):
        self._dtypename = self.lookupordefault('config', 'precision', 'double')
        if self._dtypename not in _NP_MAP:
            raise ValueError('precision must be double or float')
        self._dtype = _NP_MAP[self._dtypename]
        self._dim = self.lookupordefault('config', 'dim', 1)

    @staticmethod
    def load(file, defaults=None):
        if isinstance(file, str):
            with open(file) as fh:
                return Dictionary(fh.read(), defaults=defaults)
        return Dictionary(file.read(), defaults=defaults)

    def has_section(self, section):
        return self._cp.has_section(section)

    def has_option(self, section, option):
        return self._cp.has_option(section, option)

    def set(self, section, option, value):
        if not self._cp.has_section(section):
            self._cp.add_section(section)
        self._cp.set(section, option, str(value))
        if section == 'config':
            self._derive()

    def lookup(self, section, option, vars=None):
        return os.path.expandvars(self._cp.get(section, option, vars=vars))

    def lookupordefault(self, section, option, default, vars=None):
        try:
            return type(default)(self._cp.get(section, option, vars=vars))
        except Exception:
            return default

    def lookuppath(self, section, option, default, vars=None, abs=False):
        path = os.path.expanduser(self.lookupordefault(section, option, default, vars))
        return os.path.abspath(path) if abs else path

    def lookupexpr(self, section, option, subs=None):
        expr = self.lookup(section, option)
        if not re.match(r'[A-Za-z0-9 \t\n\r.,+\-*/%()<>=\{\}\$]+$', expr):
            raise ValueError('expression %r holds characters outside the arithmetic set' % expr)
        if subs:
            expr = re.sub(r'\b({0})\b'.format('|'.join(subs)), lambda m: subs[m.group(1)], expr)
        expr = re.sub(r'\b((\d+\.?\d*)|(\.\d+))([eE][+-]?\d+)?(?!\s*])', _ensure_float, expr)
        return '({0})'.format(expr)

    def lookupfloat(self, section, option):
        return self._dtype(self.lookup(section, option))

    def lookupfloats(self, section, options):
        return [self.lookupfloat(section, o) for o in options]

    def lookupint(self, section, option):
        return int(self.lookup(section, option))

    def lookupints(self, section, options):
        return [self.lookupint(section, o) for o in options]

    def lookup_list(self, section, option, dtype):
        return np.array([dtype(x) for x in json.loads(self.lookup(section, option))])

    def lookupfloat_list(self, section, option):
        return self.lookup_list(section, option, self._dtype)

    def lookupint_list(self, section, option):
        return self.lookup_list(section, option, int)

    def section_values(self, section, type):
        out = {}
        for k, v in self._cp.items(section):
            try:
                if '[' in v:
                    out[k] = self.lookup_list(section, k, type)
                else:
                    out[k] = type(v)
            except ValueError:
                pass
        return out

    def __str__(self):
        buf = io.StringIO()
        self._cp.write(buf)
        return buf.getvalue()

    @property
    def dtypename(self):
        return 'float' if self._dtype == np.float32 else 'double'

    @property
    def dtype(self):
        return self._dtype

    @property
    def dim(self):
        return self._dim
