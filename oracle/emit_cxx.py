"""ORACLE (test infrastructure only; never part of the product path).

C/OpenMP restatement of the reference's MPI C++ emitter for one rank with
periodic boundaries. Given the manifest-machine program of a `.fmr`, it emits a
C99 library with the same ABI as the reference (`<name>.h`: state arrays,
`struct Formura_Navigator`, `Formura_Init`, `Formura_Forward`) whose arithmetic
is the reference's statement for statement:

  * one loop nest per manifest node, innermost axis last, OpenMP on the
    outer axes (MPICxx/Translate.hs:671-737, 647-650);
  * inside it, one statement `T a_k = (x op y);` per micro node, with T the
    C type of the node (MPICxx/Translate.hs:218-230, 481-629); literals are
    the nearest double printed shortest round-trip (:585), `**` is pow
    (:595), `>?`/`<?`/`<%` are fmax/fmin chains (:606-609), `ite` is `?:`
    (:597-601), index variables are `navi.offset_x + ix` (:611-615);
  * manifest nodes are stored into their own arrays (`Om_<id>`) and read back
    at cursors; state stores go to the state arrays;
  * `Formura_Forward` advances by monitorInterval2 = monitor_interval rounded
    up to a multiple of 2T and adds it to `time_step` (:1106-1131).

What differs from the reference, and why it changes no value: the region
plan, ridge staging and MPI self-messages of Cut.hs / MPICxx/Translate.hs:742-836
are replaced by a periodic halo fill of each array before it is read at a
cursor, and the skewed in-place store by a double buffer. Both only move
values; no arithmetic statement changes. The temporal-blocking fold
(TemporalBlocking.hs:49-134) is likewise arithmetic-free and is not applied.

Build with: g++/gcc -O3 -fopenmp -ffp-contract=off (no -ffast-math).
"""
from __future__ import annotations

import os
import sys
from typing import Dict, List

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))

from formura_amd.compiler.abi import Layout, header_text, init_body  # noqa: E402
from formura_amd.compiler.manifest import MMProgram  # noqa: E402
from formura_amd.compiler.pointir import C_TYPES, c_expr_of, imm_literal, Op  # noqa: E402


def _ctype(elem: str) -> str:
    return C_TYPES.get(elem, "double")


def _param_nodes(mm: MMProgram, params) -> Dict[int, int]:
    """OM node id -> parameter index for the named scalars made parameters
    (same rule as the point IR: a non-grid node bound to that name)."""
    out = {}
    for nid, n in mm.om.step_graph.items():
        if n.name in params and n.typ[0] == "elem":
            out[nid] = list(params).index(n.name)
    return out


def emit_cxx(lay: Layout, mm: MMProgram, params=(), extern_header=None) -> str:
    dim = lay.dim
    axes = lay.axes
    shape = lay.shape()
    lo, hi = lay.lower(), lay.upper()
    ivars = ["i" + a for a in axes]
    dims_decl = "".join(f"[{s}]" for s in shape)
    out: List[str] = []
    w = out.append
    w("/* ORACLE: reference-faithful CPU step library (generated). */")
    w("#include <math.h>")
    w("#include <stdbool.h>")
    w("#include <string.h>")
    w(f'#include "{lay.name}.h"')
    if extern_header:
        w(f'#include "{extern_header}"')
    w("")
    for f, t in zip(lay.field_names, lay.field_ctype):
        w(f"{t} {f}{dims_decl};")
        w(f"static {t} Next_{f}{dims_decl};")
    w("struct Formura_Navigator navi;")
    pmap = _param_nodes(mm, params)
    if params:
        # per-run values of the named scalars (batched-launch parity tests)
        w(f"static double fmr_param[{len(params)}];")
        w("int Formura_Set_Params(const double *p, int np)")
        w("{")
        w(f"  if (np != {len(params)}) return -1;")
        w(f"  for (int i = 0; i < {len(params)}; ++i) fmr_param[i] = p[i];")
        w("  return 0;")
        w("}")
    stores = {}
    om_nodes = []
    for k, n in mm.step.items():
        root = n.micro[n.root()]
        if root.inst[0] == "Store":
            stores[k] = root.inst[1]
        else:
            om_nodes.append(k)
            w(f"static {_ctype(n.typ[2] if n.typ[0] == 'grid' else n.typ[1])} Om_{k}{dims_decl};")
    w("")
    w("int Formura_Init (struct Formura_Navigator *navi, MPI_Comm comm)")
    w("{")
    w("  " + init_body(lay))
    w("  return 0;")
    w("}")
    w("")

    # periodic halo fill: copy the opposite interior slab into each margin
    w("#define FMR_WRAP(i, lo, n) ((((i) - (lo)) % (n) + (n)) % (n) + (lo))")
    for a in range(dim):
        pass
    w("static void halo_fill(void *base, int elem)")
    w("{")
    w("  char *p = (char *)base;")
    strides = []
    s = 1
    for a in reversed(range(dim)):
        strides.insert(0, s)
        s *= shape[a]
    for a in range(dim):
        # fill margin slabs along axis a over the full extent of the other axes
        loops = []
        for b in range(dim):
            if b == a:
                continue
            loops.append(b)
        w("  #pragma omp parallel for")
        w(f"  for (int m = 0; m < {shape[a]}; ++m) {{")
        w(f"    if (m >= {lo[a]} && m < {hi[a]}) continue;")
        w(f"    int src = FMR_WRAP(m, {lo[a]}, {lay.intra[a]});")
        if loops:
            for b in loops:
                w(f"    for (int j{b} = 0; j{b} < {shape[b]}; ++j{b})")
            idx_dst = " + ".join([f"(long)m * {strides[a]}"] + [f"(long)j{b} * {strides[b]}" for b in loops])
            idx_src = " + ".join([f"(long)src * {strides[a]}"] + [f"(long)j{b} * {strides[b]}" for b in loops])
            w(f"      memcpy(p + ({idx_dst}) * elem, p + ({idx_src}) * elem, elem);")
        else:
            w(f"    memcpy(p + (long)m * elem, p + (long)src * elem, elem);")
        w("  }")
    w("}")
    w("")

    def acc(cursor) -> str:
        return "".join(f"[{iv}+({c})]" for iv, c in zip(ivars, cursor))

    def body_of(k: int) -> List[str]:
        n = mm.step[k]
        lines = []
        names: Dict[int, str] = {}
        for j in sorted(n.micro):
            mn = n.micro[j]
            inst = mn.inst
            ct = _ctype(mn.elem)
            tag = inst[0]
            if mn.om_id in pmap and tag not in ("LoadCursorStatic", "LoadCursor", "Store"):
                names[j] = f"fmr_param[{pmap[mn.om_id]}]"
                continue
            if tag == "LoadCursorStatic":
                names[j] = f"{inst[2]}{acc(inst[1])}"
                continue
            if tag == "LoadCursor":
                names[j] = f"Om_{inst[2]}{acc(inst[1])}"
                continue
            if tag == "Imm":
                names[j] = imm_literal(inst[1], "double")
                continue
            if tag == "LoadIndex":
                ax = inst[1]
                code = f"(navi.offset_{axes[ax]}+{ivars[ax]})"
            elif tag == "Store":
                names[j] = names[inst[2]]
                continue
            else:
                kind = {"Uniop": "uniop", "Binop": "binop", "Triop": "triop", "Naryop": "naryop"}[tag]
                code = c_expr_of(Op(kind, inst[1:], ct), lambda i: names[i])
            v = f"a_{j}"
            lines.append(f"{ct} {v} = {code};")
            names[j] = v
        return lines, names[n.root()]

    w("int Formura_Forward (struct Formura_Navigator *navi_p)")
    w("{")
    w(f"  for (int timestep = 0; timestep < {lay.steps_per_forward}; ++timestep) {{")
    for f, t in zip(lay.field_names, lay.field_ctype):
        w(f"    halo_fill({f}, sizeof({t}));")
    for k in list(om_nodes) + list(stores):
        lines, root = body_of(k)
        coll = f" collapse({dim - 1})" if dim > 1 else ""
        w(f"    #pragma omp parallel for{coll}")
        for a in range(dim):
            w(f"    for (int {ivars[a]} = {lo[a]}; {ivars[a]} < {hi[a]}; ++{ivars[a]})")
        w("    {")
        for ln in lines:
            w("      " + ln)
        idx = "".join(f"[{iv}]" for iv in ivars)
        if k in stores:
            w(f"      Next_{stores[k]}{idx} = {root};")
        else:
            w(f"      Om_{k}{idx} = {root};")
        w("    }")
        if k not in stores:
            n = mm.step[k]
            w(f"    halo_fill(Om_{k}, sizeof(Om_{k}{''.join('[0]' for _ in range(dim))}));")
    for f in lay.field_names:
        w(f"    memcpy({f}, Next_{f}, sizeof({f}));")
    w("  }")
    w(f"  navi_p->time_step += {lay.steps_per_forward};")
    w("  return 0;")
    w("}")
    return "\n".join(out) + "\n"


def write_oracle(outdir: str, lay: Layout, mm: MMProgram, params=(), extern_header=None) -> str:
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f"{lay.name}.h"), "w") as f:
        f.write(header_text(lay))
    path = os.path.join(outdir, f"{lay.name}_cpu.c")
    with open(path, "w") as f:
        f.write(emit_cxx(lay, mm, params, extern_header))
    return path
