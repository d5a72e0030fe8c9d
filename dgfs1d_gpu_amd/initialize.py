"""Command line of the drivers: ``run inp [-v sect::key value]...`` and
``restart inp dist... [-v ...]`` (same grammar as dgfs1D/initialize.py:7-48)."""
from argparse import ArgumentParser, FileType

from .dictionary import Dictionary


def overrides_to_dict(pairs):
    out = {}
    for key, value in pairs:
        scope = key.split('::')
        node = out
        for k in scope[:-1]:
            node = node.setdefault(k, {})
        node[scope[-1]] = value
    return out


def initialize(argv=None):
    ap = ArgumentParser()
    sp = ap.add_subparsers(dest='cmd', help='sub-command help')
    ap_run = sp.add_parser('run', help='run --help')
    ap_run.add_argument('inp', type=FileType('r'), help='input file')
    ap_run.add_argument('-v', nargs=2, action='append', default=[],
                        help='substitute variables. Example: -v mesh::Ne 4')
    ap_run.set_defaults(process_run=True)
    ap_rs = sp.add_parser('restart', help='restart --help')
    ap_rs.add_argument('inp', type=FileType('r'), help='input file')
    ap_rs.add_argument('dist', nargs='+', action='append', type=FileType('rb'), default=[],
                       help='sorted list of distributions')
    ap_rs.add_argument('-v', nargs=2, action='append', default=[],
                       help='substitute variables. Example: -v mesh::Ne 4')
    ap_rs.set_defaults(process_restart=True)
    args = ap.parse_args(argv)
    if hasattr(args, 'process_run') or hasattr(args, 'process_restart'):
        return Dictionary.load(args.inp, defaults=overrides_to_dict(args.v)), args
    ap.print_help()
    raise SystemExit(2)
