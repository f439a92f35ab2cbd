"""The generated library ABI: `<name>.h`, shared by the HIP library and the
CPU oracle so that one unchanged `main.cpp` links against either.

Layout of the header follows the reference emitter, MPICxx/Translate.hs:
1000-1139: `#pragma once`, `extern "C"` guard, `#include <mpi.h>`, one
`#define N<AXIS>` per axis, the `rsc_surface` typedef, one `extern` state
array per state variable (interior at `[lower, upper)` on every axis), the
`Formura_Navigator` struct (time_step; lower/upper/offset per axis; mpi_comm,
mpi_my_rank and the neighbour ranks), the global `navi`, `Formura_Init` and
`Formura_Forward`.

State-array margins: the reference sizes them from its region plan
(MPICxx/Translate.hs:935-969 via Cut.hs); here every axis gets a lower and an
upper margin equal to the deepest stencil reach of one step. The margins hold
no interior data; a driver only touches `[lower, upper)`.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass
from typing import List

from .manifest import MMProgram
from .pointir import PointIR


@dataclass
class Layout:
    name: str
    dim: int
    axes: List[str]
    intra: List[int]            # interior extent per axis (one rank)
    margin: List[int]           # lower margin per axis (= upper margin)
    field_names: List[str]
    field_ctype: List[str]      # element type of the host arrays
    steps_per_forward: int      # monitorInterval2

    def shape(self) -> List[int]:
        return [n + 2 * m for n, m in zip(self.intra, self.margin)]

    def lower(self) -> List[int]:
        return list(self.margin)

    def upper(self) -> List[int]:
        return [m + n for m, n in zip(self.margin, self.intra)]


def reach_of(mm: MMProgram) -> List[int]:
    """Deepest cursor offset per axis over every micro node of one step."""
    r = [0] * mm.dimension
    for n in mm.step.values():
        for mn in n.micro.values():
            if mn.inst[0] in ("LoadCursorStatic", "LoadCursor"):
                for a, c in enumerate(mn.inst[1]):
                    r[a] = max(r[a], abs(c))
    return [max(1, v) for v in r]


def make_layout(name: str, mm: MMProgram, ir: PointIR) -> Layout:
    nc = mm.numerical_config
    if nc is None or len(nc.intra_node_shape) != mm.dimension:
        raise ValueError("numerical config must give intra_node_shape for every axis")
    if any(g != 1 for g in nc.mpi_grid_shape):
        raise ValueError("one GPU per program: mpi_grid_shape must be all ones")
    ftypes = []
    for f in mm.state_order:
        e = mm.state_signature[f][2]
        ftypes.append("float" if e == "float" else "double")
    return Layout(name, mm.dimension, list(mm.axes), list(nc.intra_node_shape), reach_of(mm),
                  list(mm.state_order), ftypes, nc.monitor_interval2())


def neighbour_deltas(dim: int):
    return [d for d in itertools.product((-1, 0, 1), repeat=dim) if any(d)]


def delta_name(d) -> str:
    return "mpi_rank_" + "_".join(("m%d" % -v) if v < 0 else str(v) for v in d)


def header_text(lay: Layout) -> str:
    shape = lay.shape()
    dims = "".join(f"[{s}]" for s in shape)
    out = [
        f"/* {lay.name}.h -- Formura_Navigator ABI of the generated step library. */",
        "#pragma once",
        "#ifdef __cplusplus",
        'extern "C"',
        "{",
        "#endif",
        "#include <mpi.h>",
        "",
    ]
    for a, n in zip(lay.axes, lay.intra):
        out.append(f"#define N{a.upper()}  {n}")
    out.append("typedef double rsc_surface" + "".join(f"[{s}]" for s in shape[1:]) + ";")
    for f, t in zip(lay.field_names, lay.field_ctype):
        out.append(f"extern {t} {f}{dims};")
    out.append("struct Formura_Navigator {")
    out.append("int time_step;")
    for a in lay.axes:
        out += [f"int lower_{a};", f"int upper_{a};", f"int offset_{a};"]
    out += ["MPI_Comm mpi_comm;", "int mpi_my_rank;"]
    for d in neighbour_deltas(lay.dim):
        out.append(f"int {delta_name(d)};")
    out.append("};")
    out.append("extern struct Formura_Navigator navi;")
    out.append("int Formura_Init (struct Formura_Navigator *navi, MPI_Comm comm);")
    out.append("int Formura_Forward (struct Formura_Navigator *navi);")
    out += ["", "#ifdef __cplusplus", "}", "#endif", ""]
    return "\n".join(out)


def init_body(lay: Layout) -> str:
    """`Formura_Init` for one rank of a periodic single-rank grid
    (MPICxx/Translate.hs:1070-1092 with mpi_grid_shape all ones: every
    neighbour rank is rank 0, offset = i*intra - o = -o)."""
    lines = ["navi->mpi_comm = comm;", "navi->mpi_my_rank = 0;"]
    for d in neighbour_deltas(lay.dim):
        lines.append(f"navi->{delta_name(d)} = 0;")
    lines.append("navi->time_step = 0;")
    for a, o, n in zip(lay.axes, lay.margin, lay.intra):
        lines += [f"navi->offset_{a} = {-o};", f"navi->lower_{a} = {o};", f"navi->upper_{a} = {o + n};"]
    return "\n  ".join(lines)
