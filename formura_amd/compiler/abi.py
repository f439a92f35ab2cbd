"""The generated library ABI: `<name>.h`, shared by the HIP library and the
CPU oracle so that one unchanged `main.cpp` links against either.

Layout of the header follows the reference emitter, MPICxx/Translate.hs:
1000-1139: `#pragma once`, `extern "C"` guard, `#include <mpi.h>`, one
`#define N<AXIS>` per axis, the `rsc_surface` typedef, one `extern` state
array per state variable (interior at `[lower, upper)` on every axis), the
`Formura_Navigator` struct (time_step; lower/upper/offset per axis; mpi_comm,
mpi_my_rank and the neighbour ranks), the global `navi`, `Formura_Init` and
`Formura_Forward`.

State-array extents and neighbour ranks, as the reference derives them:
the region plans (normal and wall-inverted, cut_plan.py restating Cut.hs) of
the T-fold step graph are made, the state arrays get the hull of both plans'
static allocations with the nanoblock margin `nbu - 1 + 2` added to its upper
vertex (`collaboratePlans`, MPICxx/Translate.hs:935-969), and the navigator
declares one `mpi_rank_<delta>` field per rank delta of the plans' ridges and
its negation, in sorted order (:1032-1035, names by `toCName`, :317-345).
When no plan can be made -- a config without `initial_walls`, or walls that
leave regions narrower than the stencil reach, which the reference rejects at
generation time (Cut.hs:182-214) -- the arrays get a margin of the deepest
stencil reach of one step on both sides and the navigator every neighbour
delta.  The margins hold no interior data; a driver only touches
`[lower, upper)`.  Nanoblock unrolling (`nbu<axis>N`) changes no value; the
GPU kernels ignore it, and it enters the header only through the margins.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass
from typing import List, Optional, Tuple

from .manifest import MMProgram
from .pointir import PointIR


@dataclass
class Layout:
    name: str
    dim: int
    axes: List[str]
    intra: List[int]            # interior extent per axis (one rank)
    margin: List[int]           # lower margin per axis
    field_names: List[str]
    field_ctype: List[str]      # element type of the host arrays
    steps_per_forward: int      # monitorInterval2
    margin_hi: Optional[List[int]] = None      # upper margin per axis (default: the lower one)
    deltas: Optional[List[Tuple[int, ...]]] = None   # navigator rank deltas (default: every neighbour)

    def __post_init__(self):
        if self.margin_hi is None:
            self.margin_hi = list(self.margin)
        if self.deltas is None:
            self.deltas = neighbour_deltas(self.dim)

    def shape(self) -> List[int]:
        return [n + a + b for n, a, b in zip(self.intra, self.margin, self.margin_hi)]

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


def plan_geometry(mm: MMProgram):
    """(lower vertex, upper vertex, rank deltas) of the state arrays from the
    region plans (collaboratePlans, MPICxx/Translate.hs:935-969; deltaMPIs,
    :1032-1035), or None when no plan can be made for this config."""
    from .cut_plan import make_plan
    from .manifest import temporal_blocking
    nc = mm.numerical_config
    T = max(1, nc.temporal_blocking_interval)
    step = temporal_blocking(mm, T) if T > 1 else dict(mm.step)
    try:
        plans = [make_plan(step, mm.dimension, list(mm.axes), list(nc.intra_node_shape), nc.initial_walls or {},
                           mm.state_order, inv, nc.option_strings or []) for inv in (False, True)]
    except ValueError:
        return None
    hull = None
    for p in plans:
        for k, b in p.array_alloc.items():
            if k.static:
                hull = b if hull is None else hull.hull(b)
    if hull is None:
        return None
    nbu = [nc.nbu_size(a) for a in mm.axes]
    lo = list(hull.lo)
    hi = [u + n - 1 + 2 for u, n in zip(hull.hi, nbu)]
    deltas = set()
    for p in plans:
        for r in p.ridge_alloc:
            d = tuple(r.dmpi)
            deltas.add(d)
            deltas.add(tuple(-x for x in d))
    return lo, hi, sorted(deltas)


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
    intra = list(nc.intra_node_shape)
    reach = reach_of(mm)
    lay = Layout(name, mm.dimension, list(mm.axes), intra, reach, list(mm.state_order), ftypes,
                 nc.monitor_interval2())
    g = plan_geometry(mm)
    if g is not None:
        lo, hi, deltas = g
        m_lo = [-x for x in lo]
        m_hi = [h - n for h, n in zip(hi, intra)]
        if all(a >= r and b >= r for a, b, r in zip(m_lo, m_hi, reach)):
            lay.margin, lay.margin_hi, lay.deltas = m_lo, m_hi, [tuple(d) for d in deltas]
    return lay


def neighbour_deltas(dim: int):
    return [d for d in itertools.product((-1, 0, 1), repeat=dim) if any(d)]


def delta_name(d) -> str:
    """`nameDeltaMPIRank` (MPICxx/Translate.hs:406-407): `mpi_rank_` and the
    `toCName` of the rank delta -- its shown form `MPIRank (-1,0)` with `-`
    written `m`, runs of other characters as one `_`, `MPIRank` dropped and
    trailing `_` removed, e.g. `mpi_rank__m1_0`."""
    return "mpi_rank__" + "_".join(("m%d" % -v) if v < 0 else str(v) for v in d)


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
    for d in lay.deltas:
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
    for d in lay.deltas:
        lines.append(f"navi->{delta_name(d)} = 0;")
    lines.append("navi->time_step = 0;")
    for a, o, n in zip(lay.axes, lay.margin, lay.intra):
        lines += [f"navi->offset_{a} = {-o};", f"navi->lower_{a} = {o};", f"navi->upper_{a} = {o + n};"]
    return "\n  ".join(lines)
