"""Compiler pipeline: `.fmr` + `.yaml` -> header + HIP step library source.

Stage order mirrors exe-src/formura.hs:34-84: parse (with the numerical
config and the injected extents), desugar, OM graph, manifestation, backend.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

from .abi import Layout, header_text, make_layout
from .emit_hip import emit_hip
from .manifest import MMProgram, gen_mm_program
from .numerical_config import NumericalConfig
from .omgraph import OMProgram, gen_om_program
from .parser import parse_program
from .pointir import lower_stages, PointIR, lower_step


@dataclass
class Compiled:
    name: str
    om: OMProgram
    mm: MMProgram
    ir: PointIR
    layout: Layout
    precision: str
    source_path: Optional[str] = None    # the .fmr file


def extern_functions(c: "Compiled"):
    """Names of the extern functions the step calls."""
    out = set()
    for n in c.om.step_graph.values():
        inst = n.inst
        if inst[0] == "Uniop" and str(inst[1]).startswith("external-call/"):
            out.add(str(inst[1])[len("external-call/"):])
    return sorted(out)


def extern_header_of(c: "Compiled") -> Optional[str]:
    """The header defining the program's extern functions: `<stem>.extern.h`
    beside the `.fmr` (host and device definitions, `FMR_HD` before each),
    or None when the step calls none.  A program that calls an extern
    function without such a header is refused, as an undefined symbol would
    be at link time."""
    names = extern_functions(c)
    if not names:
        return None
    if not c.source_path:
        raise ValueError(f"extern functions {names} need the program's .fmr path to find <stem>.extern.h")
    stem = os.path.splitext(c.source_path)[0]
    path = stem + ".extern.h"
    if not os.path.exists(path):
        raise ValueError(f"extern functions {names}: no definitions in {path}")
    return path


def copy_extern_header(c: "Compiled", outdir: str) -> Optional[str]:
    """Copy the extern header next to the generated source; its file name."""
    path = extern_header_of(c)
    if path is None:
        return None
    name = os.path.basename(path)
    with open(path, encoding="utf-8") as f:
        _write_if_changed(os.path.join(outdir, name), f.read())
    return name


def compile_source(name: str, fmr_text: str, nc: NumericalConfig, precision: str = "faithful",
                   params: Sequence[str] = ()) -> Compiled:
    prog = parse_program(fmr_text, nc)
    om = gen_om_program(prog)
    mm = gen_mm_program(om)
    ir = lower_step(om, precision=precision, params=params)
    lay = make_layout(name, mm, ir)
    return Compiled(name, om, mm, ir, lay, precision)


def compile_files(name: str, fmr_path: str, yaml_path: str, precision: str = "faithful",
                  params: Sequence[str] = (), nc_override: Optional[dict] = None) -> Compiled:
    with open(fmr_path, encoding="utf-8") as f:
        text = f.read()
    nc = NumericalConfig.from_file(yaml_path)
    for k, v in (nc_override or {}).items():
        setattr(nc, k, v)
    c = compile_source(name, text, nc, precision, params)
    c.source_path = os.path.abspath(fmr_path)
    return c


def write_hip(c: Compiled, outdir: str, march_variants=None) -> Tuple[str, str]:
    os.makedirs(outdir, exist_ok=True)
    h = os.path.join(outdir, f"{c.name}.h")
    src = os.path.join(outdir, f"{c.name}.hip")
    _write_if_changed(h, header_text(c.layout))
    stages = lower_stages(c.mm, c.precision, c.ir.params)
    ext = copy_extern_header(c, outdir)
    _write_if_changed(src, emit_hip(c.layout, c.ir, c.precision, march_variants, stages, extern_header=ext))
    return h, src


def _write_if_changed(path: str, text: str) -> None:
    if os.path.exists(path):
        with open(path, encoding="utf-8") as f:
            if f.read() == text:
                return
    with open(path, "w", encoding="utf-8") as f:
        f.write(text)
