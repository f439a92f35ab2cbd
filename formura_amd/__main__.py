"""Command line of the compiler, after the reference's
`formura FILES -o OUT --nc NC` (src/Formura/CommandLineOption.hs:25-45).

    python -m formura_amd prog.fmr --nc prog.yaml -o out/prog.hip [--precision fp32]

The output extension selects the backend as in MPICxx/Language.hs:17-22;
`.hip` writes `prog.h` and `prog.hip` (the single-GPU step library).
"""
from __future__ import annotations

import argparse
import os
import sys

from .compiler.driver import compile_files, write_hip


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="formura_amd")
    ap.add_argument("fmr")
    ap.add_argument("--nc", required=True, help="numerical config (.yaml)")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("--precision", choices=["faithful", "fp32"], default="faithful")
    ap.add_argument("--params", default="", help="comma-separated scalar names made per-launch parameters")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    base, ext = os.path.splitext(a.output)
    if ext != ".hip":
        print(f"formura_amd: unsupported output extension {ext!r} (use .hip)", file=sys.stderr)
        return 2
    params = [p for p in a.params.split(",") if p]
    c = compile_files(os.path.basename(base), a.fmr, a.nc, a.precision, params)
    if a.verbose:
        for k, n in c.om.step_graph.items():
            print(f"{'M' if n.manifest else ' '} {n.name or '':8s} ({k},{n.inst}) :: {n.typ}")
    write_hip(c, os.path.dirname(a.output) or ".")
    return 0


if __name__ == "__main__":
    sys.exit(main())
