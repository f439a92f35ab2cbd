"""ORACLE (test infrastructure only; never part of the product path).

C/OpenMP step library on the reference's own loop and region structure:
the region plans of formura_amd/compiler/cut_plan.py (normal and wall-inverted, Cut.hs)
turned into C as the reference's MPI C++ backend does
(src/Formura/MPICxx/Translate.hs):

  * plans collaborate (:935-969): the state arrays get the hull of both
    plans' static allocations plus the nanoblock margin, which fixes their
    shape and the navigator's lower/upper/offset (:1070-1092);
  * each manifest-node array of a region lives in a shared buffer (the
    plan's buffer sharing) sized for the largest region;
  * a computation is a loop nest over the region box in the coordinates of
    the destination array, one statement per micro node (:671-737,
    481-629); state stores are skewed by the plan's system offset (:714-719);
  * ridges are copied in and out of halo slabs (:742-796); ridges that cross
    ranks travel in facet structs by MPI_Irecv / MPI_Isend / MPI_Wait
    (:798-836) -- with one rank these are self-messages, delivered by the
    serial MPI header;
  * consecutive non-communication instructions form one `Formura_internal_N`
    function with `#pragma omp parallel` and `#pragma omp for
    collapse(dim-1)` on each loop nest (:840-931, 647-650);
  * `Formura_Forward` runs monitorInterval2 / (2T) times the normal plan then
    the inverted one, over the T-fold step graph (:1106-1131).

  * nanoblock unroll `nbu<axis>N` (Manifestation.hs:146-158, :698-719): the
    loops step by the nanoblock, one iteration computes every cursor of it
    (an operation shared by several cursors once) and stores each; regions
    that are not whole nanoblocks, where the reference's loop would run past
    the region, are refused.

Names follow the reference's patterns (`Rsc<k>`, `Ridge_...`, `Facet_...`,
`req_...`) but are not guaranteed to be its exact text; the loop-template
deduplication into `Formura_subroutine_N` (a code-size measure) is left out.

Build with: gcc -O3 -fopenmp -ffp-contract=off (no -ffast-math).
"""
from __future__ import annotations

import os
import re
import sys
from typing import Dict, List

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))

from formura_amd.compiler.abi import Layout, delta_name, header_text  # noqa: E402
from formura_amd.compiler.manifest import MMProgram, temporal_blocking  # noqa: E402
from formura_amd.compiler.pointir import C_TYPES, Op, c_expr_of, imm_literal  # noqa: E402
from formura_amd.compiler.cut_plan import Box, FacetID, IRank, Plan, RidgeID, Rsc, make_plan  # noqa: E402


def PlannedLayout(base: Layout, lower, shape, deltas) -> Layout:
    """Layout of a library with the plan's (asymmetric) array margins and
    rank deltas."""
    hi = [sh - n - lo for sh, n, lo in zip(shape, base.intra, lower)]
    return Layout(base.name, base.dim, base.axes, base.intra, list(lower), base.field_names, base.field_ctype,
                  base.steps_per_forward, hi, [tuple(d) for d in deltas])


def _ctype(elem: str) -> str:
    return C_TYPES.get(elem, "double")


def _cname(s: str) -> str:
    s = s.replace("-", "m")
    s = re.sub(r"[^0-9A-Za-z]+", "_", s)
    return s.strip("_")


def _vec(v) -> str:
    return "_".join(str(x) for x in v)


def ridge_name(r: RidgeID) -> str:
    if r.delta.static:
        d = f"St_{r.delta.name}"
    else:
        s, t = r.delta.extra
        d = f"Om_{r.delta.name}_r_{_vec(s.v)}_r_{_vec(t.v)}"
    return _cname(f"Ridge_{_vec(r.dmpi)}_{d}")


def facet_name(f: FacetID) -> str:
    return _cname(f"Facet_{_vec(f.dmpi)}_src_{_vec(f.src.v)}_dest_{_vec(f.dest.v)}")


def delta_field(d) -> str:
    return delta_name(d)


def _nbu_size(options, axis: str) -> int:
    """Nanoblock unroll `nbu<axis>N` (NumericalConfig.hs:48-53): the first
    option with that prefix, else 1."""
    kwd = "nbu" + axis
    for opt in options:
        if opt.startswith(kwd):
            return int(opt[len(kwd):])
    return 1


def _nanoblock_cursors(nbu) -> List[tuple]:
    """Cursors of one nanoblock in the order the reference manifests them
    (Manifestation.hs:146-158: x outer, z inner, then reversed)."""
    cs = [()]
    for n in nbu:
        cs = [c + (k,) for c in cs for k in range(n)]
    return list(reversed(cs))


class _Emitter:
    def __init__(self, lay: Layout, mm: MMProgram, T: int, walls, options):
        self.lay = lay
        self.mm = mm
        self.dim = lay.dim
        self.axes = lay.axes
        self.ivars = ["i" + a for a in self.axes]
        self.step = temporal_blocking(mm, T) if T > 1 else dict(mm.step)
        self.T = T
        self.options = list(options)
        self.plans: List[Plan] = [make_plan(self.step, self.dim, self.axes, lay.intra, walls, mm.state_order,
                                            inv, self.options) for inv in (False, True)]
        # collaboratePlans (:935-969): one static box for both plans, plus the
        # nanoblock margin nbu-1+2 per axis
        hull = None
        for p in self.plans:
            for k, b in p.array_alloc.items():
                if k.static:
                    hull = b if hull is None else hull.hull(b)
        self.nbu = [_nbu_size(self.options, a) for a in self.axes]
        self.static_box = Box(hull.lo, tuple(u + n - 1 + 2 for u, n in zip(hull.hi, self.nbu)))
        for p in self.plans:
            for k in list(p.array_alloc):
                if k.static:
                    p.array_alloc[k] = self.static_box
        # shared buffers: one extent for both plans (the larger of the two)
        ext = [0] * self.dim
        for p in self.plans:
            ext = [max(a, b) for a, b in zip(ext, p.shared_extent.shape())]
        self.shared_shape = ext
        self.funcs: List[str] = []
        self.nfunc = 0
        # ridge and facet names per plan; a name the normal plan already
        # uses gets the suffix _1 in the inverted plan (genFreeName, :150-171)
        self.names: Dict[tuple, str] = {}
        used = set()
        for pi, p in enumerate(self.plans):
            for kind, keys, fn in (("r", sorted(p.ridge_alloc), ridge_name), ("f", sorted(p.facets), facet_name)):
                for k in keys:
                    base = fn(k)
                    name = base if base not in used else f"{base}_{pi}"
                    used.add(name)
                    self.names[(pi, kind, k)] = name

    # -- names -----------------------------------------------------------------
    def rname(self, p: Plan, r: RidgeID) -> str:
        return self.names[(self.plans.index(p), "r", r)]

    def fname(self, p: Plan, f: FacetID) -> str:
        return self.names[(self.plans.index(p), "f", f)]

    def rsc_name(self, p: Plan, k: Rsc) -> str:
        if k.static:
            return k.name
        return f"Rsc{p.sharing[k]}" if k in p.sharing else ""

    def acc(self, offs) -> str:
        return "".join(f"[{iv}+({o})]" for iv, o in zip(self.ivars, offs))

    # -- one computation (:671-737) ---------------------------------------------
    def computation(self, p: Plan, ir: IRank, nid: int, dest: Rsc) -> List[str]:
        region = p.region[(ir, nid)]
        margin = p.array_alloc[dest]
        frm = [a - b for a, b in zip(region.lo, margin.lo)]
        to = [a - b for a, b in zip(region.hi, margin.lo)]
        ind = margin.lo
        n = self.step[nid]
        # nanoblock unroll (Manifestation.hs:146-158, :698-719): one loop
        # iteration computes the sites i + c for every cursor c of the
        # nanoblock, a micro node shared by several cursors (the same
        # operation on the same shifted operands) once, and stores each
        # cursor's root value
        nbu = self.nbu
        cursors = _nanoblock_cursors(nbu) if any(k > 1 for k in nbu) else [(0,) * self.dim]
        if len(cursors) > 1:
            if any((h - l) % k for l, h, k in zip(frm, to, nbu)):
                raise ValueError(f"region {region} is not a whole number of nanoblocks {nbu}: the reference's "
                                 "loop would run past the region")
            if any(mn.inst[0] == "LoadIndex" for mn in n.micro.values()):
                raise ValueError("index variables under nanoblock unroll are not restated")
        names: Dict[tuple, str] = {}
        lines: List[str] = []
        roots: List[tuple] = []
        seen: Dict[tuple, str] = {}
        for c in cursors:
            key_of: Dict[int, tuple] = {}
            for j in sorted(n.micro):
                mn = n.micro[j]
                inst = mn.inst
                tag = inst[0]
                ct = _ctype(mn.elem)
                if tag in ("LoadCursorStatic", "LoadCursor"):
                    vc = tuple(v + d for v, d in zip(inst[1], c))
                    key = (tag, vc, inst[2])
                elif tag == "Imm":
                    key = (tag, inst[1], mn.elem)
                elif tag == "Store":
                    key_of[j] = key_of[inst[2]]
                    continue
                elif tag == "LoadIndex":
                    key = (tag, inst[1], c)
                else:
                    args = tuple(key_of[a] if isinstance(a, int) else tuple(key_of[x] for x in a)
                                 for a in inst[2:])
                    key = (tag, inst[1], args, mn.elem)
                key_of[j] = key
                if key in seen:
                    continue
                if tag == "LoadCursorStatic":
                    ab = p.array_alloc[Rsc(True, inst[2])]
                    seen[key] = inst[2] + self.acc([i + v - l for i, v, l in zip(ind, key[1], ab.lo)])
                    continue
                if tag == "LoadCursor":
                    rk = Rsc(False, inst[2], ir)
                    src = self.step[inst[2]]
                    if src.typ[0] != "grid":
                        seen[key] = self.rsc_name(p, rk)
                    else:
                        ab = p.array_alloc[rk]
                        seen[key] = self.rsc_name(p, rk) + self.acc([i + v - l for i, v, l in zip(ind, key[1], ab.lo)])
                    continue
                if tag == "Imm":
                    seen[key] = imm_literal(inst[1], "double")
                    continue
                if tag == "LoadIndex":
                    ax = inst[1]
                    code = f"(navi.offset_{self.axes[ax]}+{self.ivars[ax]}+({ind[ax]}))"
                else:
                    kind = {"Uniop": "uniop", "Binop": "binop", "Triop": "triop", "Naryop": "naryop"}[tag]
                    code = c_expr_of(Op(kind, inst[1:], ct), lambda i: seen[key_of[i]])
                v = f"a_{len(lines)}"
                lines.append(f"{ct} {v} = {code};")
                seen[key] = v
            roots.append((c, seen[key_of[n.root()]]))
        out = []
        if "omp-collapse" in self.options and self.dim > 1:
            out.append(f"#pragma omp for collapse({self.dim - 1})")
        for a in range(self.dim):
            out.append(f"for (int {self.ivars[a]} = {frm[a]}; {self.ivars[a]} < {to[a]}; "
                       f"{self.ivars[a]} += {nbu[a]}) {{")
        out += lines
        for c, root in roots:
            if dest.static:
                lhs = dest.name + self.acc([d - s for d, s in zip(c, p.system_offset)])
            else:
                lhs = self.rsc_name(p, dest) + self.acc(list(c))
            out.append(f"{lhs} = {root};")
        out += ["}"] * self.dim
        return out

    # -- ridge staging (:742-796) ------------------------------------------------
    def staging(self, p: Plan, r: RidgeID, is_stage: bool) -> List[str]:
        box0 = p.ridge_alloc[r]
        if r.delta.static:
            src = Rsc(True, r.delta.name)
        else:
            s, t = r.delta.extra
            src = Rsc(False, r.delta.name, s if is_stage else t)
        box1 = p.array_alloc[src]
        shift = [d * n for d, n in zip(r.dmpi, self.lay.intra)] if is_stage else [0] * self.dim
        other = [a - b - c for a, b, c in zip(box0.lo, box1.lo, shift)]
        arr = self.rsc_name(p, src) + self.acc(other)
        if r.needs_mpi():
            rdg = f"{self.fname(p, p.facet_of[r])}_{'Send' if is_stage else 'Recv'}.{self.rname(p, r)}"
        else:
            rdg = self.rname(p, r)
        rdg += "".join(f"[{iv}]" for iv in self.ivars)
        out = []
        if "omp-collapse" in self.options and self.dim > 1:
            out.append(f"#pragma omp for collapse({self.dim - 1})")
        for a, e in enumerate(box0.shape()):
            out.append(f"for (int {self.ivars[a]} = 0; {self.ivars[a]} < {e}; ++{self.ivars[a]}) {{")
        out.append(f"{rdg} = {arr};" if is_stage else f"{arr} = {rdg};")
        out += ["}"] * self.dim
        return out

    # -- facets (:798-836) -------------------------------------------------------
    def sendrecv(self, p: Plan, f: FacetID) -> List[str]:
        fn = self.fname(p, f)
        tag = p.facet_tag[f]
        neg = tuple(-d for d in f.dmpi)
        return [f"MPI_Irecv((void *)&{fn}_Recv, sizeof(struct {fn}), MPI_BYTE, navi->{delta_field(f.dmpi)}, "
                f"{tag}, navi->mpi_comm, &req_{fn});",
                f"MPI_Isend((void *)&{fn}_Send, sizeof(struct {fn}), MPI_BYTE, navi->{delta_field(neg)}, "
                f"{tag}, navi->mpi_comm, &req_{fn});"]

    def program(self, p: Plan) -> List[str]:
        """The plan's instructions: runs of non-communication instructions
        become one Formura_internal_N function (:860-931)."""
        calls: List[str] = []
        run: List[List[str]] = []

        def flush():
            if not run:
                return
            name = f"Formura_internal_{self.nfunc}"
            self.nfunc += 1
            body = [f"static void {name}(void)", "{"]
            if "omp" in self.options:
                body += ["#pragma omp parallel", "{"]
            for blk in run:
                body += blk
            if "omp" in self.options:
                body.append("}")
            body.append("}")
            self.funcs.append("\n".join(body))
            calls.append(f"{name}();")
            run.clear()

        for inst in p.program:
            k = inst[0]
            if k == "free":
                continue
            if k in ("wait", "sendrecv"):
                flush()
                if k == "wait":
                    calls.append(f"MPI_Wait(&req_{self.fname(p, inst[1])}, MPI_STATUS_IGNORE);")
                else:
                    calls += self.sendrecv(p, inst[1])
                continue
            if k == "compute":
                ir, nid = inst[1]
                run.append(self.computation(p, ir, nid, inst[2]))
            elif k == "stage":
                run.append(self.staging(p, inst[1], True))
            elif k == "unstage":
                run.append(self.staging(p, inst[1], False))
        flush()
        return calls

    # -- the library -------------------------------------------------------------
    def deltas(self):
        ds = set()
        for p in self.plans:
            for r in p.ridge_alloc:
                ds.add(r.dmpi)
                ds.add(tuple(-d for d in r.dmpi))
        return sorted(ds)

    def layout(self) -> Layout:
        return PlannedLayout(self.lay, [-x for x in self.static_box.lo], self.static_box.shape(), self.deltas())

    def header(self) -> str:
        return header_text(self.layout())

    def source(self) -> str:
        lay = self.layout()
        dim = self.dim
        decl: List[str] = ["/* ORACLE: step library on the reference's region plan (generated). */",
                           "#include <math.h>", "#include <stdbool.h>", "#include <mpi.h>",
                           f'#include "{lay.name}.h"', ""]
        dims = "".join(f"[{s}]" for s in lay.shape())
        for f, t in zip(lay.field_names, lay.field_ctype):
            decl.append(f"{t} {f}{dims};")
        decl.append("struct Formura_Navigator navi;")
        ctype_of_share: Dict[int, str] = {}
        for p in self.plans:
            for k, sid in p.sharing.items():
                n = self.step[k.name]
                t = n.typ[2] if n.typ[0] == "grid" else n.typ[1]
                ctype_of_share[sid] = _ctype(t)
        sdims = "".join(f"[{s}]" for s in self.shared_shape)
        for sid in sorted(ctype_of_share):
            decl.append(f"static {ctype_of_share[sid]} Rsc{sid}{sdims};")
        for p in self.plans:
            for f, rs in sorted(p.facets.items()):
                fn = self.fname(p, f)
                decl.append(f"struct {fn} {{")
                for r in rs:
                    t = self._rsc_ctype(r)
                    decl.append(f"  {t} {self.rname(p, r)}" + "".join(f"[{s}]" for s in p.ridge_alloc[r].shape()) + ";")
                decl.append("};")
                decl.append(f"static struct {fn} {fn}_Send, {fn}_Recv;")
                decl.append(f"static MPI_Request req_{fn};")
            for r, b in sorted(p.ridge_alloc.items()):
                if not r.needs_mpi():
                    decl.append(f"static {self._rsc_ctype(r)} {self.rname(p, r)}" +
                                "".join(f"[{s}]" for s in b.shape()) + ";")
        decl.append("")
        # Formura_Init (:1070-1092) for one rank: every neighbour is rank 0
        o = [-x for x in self.static_box.lo]
        init = ["int Formura_Init (struct Formura_Navigator *navi, MPI_Comm comm)", "{",
                "  navi->mpi_comm = comm;", "  { int r; MPI_Comm_rank(comm, &r); navi->mpi_my_rank = r; }"]
        for d in self.deltas():
            init.append(f"  navi->{delta_field(d)} = 0;")
        init.append("  navi->time_step = 0;")
        for a, oa, n in zip(self.axes, o, lay.intra):
            init += [f"  navi->offset_{a} = {-oa};", f"  navi->lower_{a} = {oa};", f"  navi->upper_{a} = {oa + n};"]
        init += ["  return 0;", "}", ""]
        calls0 = self.program(self.plans[0])
        calls1 = self.program(self.plans[1])
        mi2 = self.lay.steps_per_forward
        fwd = ["int Formura_Forward (struct Formura_Navigator *navi)", "{",
               f"  for (int timestep = 0; timestep < {mi2 // (2 * self.T)}; ++timestep) {{"]
        fwd += ["    " + c for c in calls0] + ["    /* HALFWAYS */"] + ["    " + c for c in calls1]
        fwd += ["  }", f"  navi->time_step += {mi2};", "  return 0;", "}", ""]
        return "\n".join(decl + self.funcs + [""] + init + fwd)

    def _rsc_ctype(self, r: RidgeID) -> str:
        if r.delta.static:
            i = self.lay.field_names.index(r.delta.name)
            return self.lay.field_ctype[i]
        n = self.step[r.delta.name]
        return _ctype(n.typ[2] if n.typ[0] == "grid" else n.typ[1])


def write_planned(outdir: str, lay: Layout, mm: MMProgram, T: int = 1, walls=None, options=None):
    """Header and C source of the region-plan library; returns (path of the
    C source, its layout).  `walls` defaults to the program's initial_walls,
    or one region per axis when there are none."""
    nc = mm.numerical_config
    if walls is None:
        walls = getattr(nc, "initial_walls", None) or {}
    walls = {a: list(walls.get(a) or []) for a in lay.axes}
    for a, n in zip(lay.axes, lay.intra):
        if not walls[a]:
            walls[a] = [n // 2]
    if options is None:
        options = list(getattr(nc, "option_strings", None) or ["omp", "omp-collapse"])
    e = _Emitter(lay, mm, T, walls, options)
    src = e.source()
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, f"{lay.name}.h"), "w") as f:
        f.write(e.header())
    path = os.path.join(outdir, f"{lay.name}_plan.c")
    with open(path, "w") as f:
        f.write(src)
    return path, e.layout()
