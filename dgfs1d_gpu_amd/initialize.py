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
    sp = ap.add_subparsers(dest='cmd', help='what to do with the deck')
    ap_run = sp.add_parser('run', help='start a run from the deck\'s initial condition')
    ap_run.add_argument('inp', type=FileType('r'), help='the INI deck')
    ap_run.add_argument('-v', nargs=2, action='append', default=[],
                        help='override one deck entry, e.g. -v mesh::Ne 4 (repeatable)')
    ap_run.set_defaults(process_run=True)
    ap_rs = sp.add_parser('restart', help='continue a run from distribution files')
    ap_rs.add_argument('inp', type=FileType('r'), help='the INI deck')
    ap_rs.add_argument('dist', nargs='+', action='append', type=FileType('rb'), default=[],
                       help='distribution files, one per rank, in rank order')
    ap_rs.add_argument('-v', nargs=2, action='append', default=[],
                       help='override one deck entry, e.g. -v mesh::Ne 4 (repeatable)')
    ap_rs.set_defaults(process_restart=True)
    args = ap.parse_args(argv)
    if hasattr(args, 'process_run') or hasattr(args, 'process_restart'):
        return Dictionary.load(args.inp, defaults=overrides_to_dict(args.v)), args
    ap.print_help()
    raise SystemExit(2)
