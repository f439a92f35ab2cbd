"""Point IR: the per-site arithmetic of one explicit step, fully inlined.

The reference evaluates each manifest node into its own array and reads it
back (MPICxx/Translate.hs:671-737); every micro node is one C statement
`T a_k = (x op y);` whose declared type T comes from the node type
(MPICxx/Translate.hs:218-230, 481-629). Storing a value of type T into an
array of type T and reading it back changes no bit, so evaluating the whole
step as one expression DAG per site -- inlining manifest nodes at their
cursors -- computes exactly the same values. This module builds that DAG from
the OM step graph, with one op per OM node and cursor, each carrying the C
type the reference would declare for it.

Ops (in topological order):
  ('load', field, offset)      state field at cursor + offset
  ('imm', Fraction)            literal, printed as the nearest double
  ('param', index)             a per-instance scalar (batched launches only)
  ('index', axis)              global coordinate (`LoadIndex`, cursor-independent
                               as in Manifestation.hs:130)
  ('uniop', op, a) ('binop', op, a, b) ('triop', 'ite', c, a, b)
  ('naryop', op, [a..])
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Sequence, Tuple

from .omgraph import FormuraCompileError, OMProgram, elem_of

# element type name -> C type, faithful mode (Rational nodes are `double`,
# MPICxx/Translate.hs:221)
C_TYPES = {"Rational": "double", "double": "double", "float": "float", "real": "double",
           "Real": "double", "bool": "bool", "int": "int"}
# precision override used by the FP32 bulk path: every floating node is float
C_TYPES_FP32 = dict(C_TYPES, Rational="float", double="float", real="float", Real="float")


@dataclass
class Op:
    kind: str
    args: tuple
    ctype: str


@dataclass
class PointIR:
    dim: int
    axes: List[str]
    extents: List[int]                 # global grid extents (NX, NY, ...)
    fields: List[str]                  # state variables in `step` order
    field_ctypes: List[str]            # element C type of each state array
    ops: List[Op]
    outputs: List[int]                 # op index of the new value of each field
    params: List[str] = field(default_factory=list)

    def radius(self) -> List[int]:
        r = [0] * self.dim
        for op in self.ops:
            if op.kind == "load":
                for a, o in enumerate(op.args[1]):
                    r[a] = max(r[a], abs(o))
        return r

    def loads(self):
        return [(i, op) for i, op in enumerate(self.ops) if op.kind == "load"]

    def flops(self) -> int:
        """Floating-point operations per site (unary minus / plus excluded)."""
        n = 0
        for op in self.ops:
            if op.kind == "binop" and op.args[0] in "+-*/":
                n += 1
            elif op.kind == "naryop":
                n += len(op.args[1]) - 1
        return n


def lower_step(om: OMProgram, precision: str = "faithful",
               params: Sequence[str] = ()) -> PointIR:
    """Fully inline the step graph at cursor 0 for every `Store`.

    precision: 'faithful' keeps the reference's C types; 'fp32' maps every
    floating type to float (the FP32 bulk path).
    params: names of scalar (non-grid) variables to turn into per-instance
    parameters.
    """
    ctypes = C_TYPES if precision == "faithful" else C_TYPES_FP32
    g = om.step_graph
    fields = list(om.state_order)
    field_ctypes = [ctypes[om.state_signature[f][2]] for f in fields]
    param_list = list(params)
    param_nodes = {}
    for nid, n in g.items():
        if n.name in param_list and n.typ[0] == "elem":
            param_nodes[nid] = param_list.index(n.name)
    missing = set(param_list) - {g[n].name for n in param_nodes}
    if missing:
        raise FormuraCompileError(f"parameters not found as scalar variables: {sorted(missing)}")

    ops: List[Op] = []
    memo: Dict[tuple, int] = {}
    dim = om.dimension

    def emit(op: Op, key) -> int:
        memo[key] = len(ops)
        ops.append(op)
        return memo[key]

    def at(cursor: tuple, nid: int) -> int:
        key = (cursor, nid)
        if key in memo:
            return memo[key]
        node = g[nid]
        inst = node.inst
        tag = inst[0]
        ct = ctypes.get(elem_of(node.typ), "double") if node.typ[1] != "void" else "void"
        if nid in param_nodes:
            return emit(Op("param", (param_nodes[nid],), ct), ((0,) * dim, nid))
        if tag == "Imm":
            k2 = ((0,) * dim, nid)           # literals do not depend on the cursor
            if k2 in memo:
                return memo[k2]
            return emit(Op("imm", (inst[1],), ct), k2)
        if tag == "Shift":
            c2 = tuple(c + v for c, v in zip(cursor, inst[1]))
            r = at(c2, inst[2])
            memo[key] = r
            return r
        if tag == "Load":
            f = fields.index(inst[1])
            return emit(Op("load", (f, cursor), ct), key)
        if tag == "LoadIndex":
            return emit(Op("index", (inst[1],), ct), ((0,) * dim, nid))
        if tag == "Uniop":
            a = at(cursor, inst[2])
            return emit(Op("uniop", (inst[1], a), ct), key)
        if tag == "Binop":
            a = at(cursor, inst[2])
            b = at(cursor, inst[3])
            return emit(Op("binop", (inst[1], a, b), ct), key)
        if tag == "Triop":
            a = at(cursor, inst[2])
            b = at(cursor, inst[3])
            c = at(cursor, inst[4])
            return emit(Op("triop", (inst[1], a, b, c), ct), key)
        if tag == "Naryop":
            xs = [at(cursor, x) for x in inst[2]]
            return emit(Op("naryop", (inst[1], xs), ct), key)
        raise FormuraCompileError(f"cannot lower {tag}")

    outputs = [None] * len(fields)
    for nid, n in g.items():
        if n.inst[0] == "Store":
            f = fields.index(n.inst[1])
            outputs[f] = at((0,) * dim, n.inst[2])
    if any(o is None for o in outputs):
        raise FormuraCompileError("every state variable needs a Store")
    nc = om.numerical_config
    extents = nc.global_extents() if nc is not None else []
    return PointIR(dim, list(om.axes), extents, fields, field_ctypes, ops, outputs, param_list)


@dataclass
class Stage:
    """One manifest node of the step, evaluated over the whole grid by one
    kernel: `ir` has one output; its loads index the arrays [state fields...,
    stage arrays...]; the result goes to stage array `index` (target
    'temp') or to the new value of state field `index` (target 'state')."""
    ir: PointIR
    target: str
    index: int
    om_id: int


def stage_order(mm) -> Tuple[List[int], List[int]]:
    """Manifest nodes in evaluation order -- the non-Store ones in MM order,
    then the Stores (the order the CPU oracle uses)."""
    temps, stores = [], []
    for k, n in mm.step.items():
        (stores if n.micro[n.root()].inst[0] == "Store" else temps).append(k)
    return temps, stores


def lower_stages(mm, precision: str = "faithful", params: Sequence[str] = ()) -> List[Stage]:
    """The step as a sequence of manifest-node kernels, as the reference
    evaluates it (one loop nest per manifest node, MPICxx/Translate.hs:671-737):
    each node's micro DAG (Manifestation.hs:93-158) becomes a point function
    whose `LoadCursorStatic` reads a state field and whose `LoadCursor` reads
    the array of an earlier manifest node.  Same operations as the inlined
    step, in stages: storing a value of type T and reading it back changes no
    bit, so the results are the same."""
    ctypes = C_TYPES if precision == "faithful" else C_TYPES_FP32
    fields = list(mm.state_order)
    nf = len(fields)
    temps, stores = stage_order(mm)
    temp_index = {k: nf + i for i, k in enumerate(temps)}
    g = mm.om.step_graph
    param_list = list(params)
    param_nodes = {nid: param_list.index(n.name) for nid, n in g.items()
                   if n.name in param_list and n.typ[0] == "elem"}
    field_ctypes = [ctypes[mm.state_signature[f][2]] for f in fields]
    for k in temps:
        t = mm.step[k].typ
        field_ctypes.append(ctypes.get(t[2] if t[0] == "grid" else t[1], "double"))
    names = fields + [f"Om_{k}" for k in temps]
    nc = mm.numerical_config
    extents = nc.global_extents() if nc is not None else []
    out: List[Stage] = []
    for k in temps + stores:
        n = mm.step[k]
        ops: List[Op] = []
        ref: Dict[int, int] = {}
        for j in sorted(n.micro):
            mn = n.micro[j]
            inst = mn.inst
            tag = inst[0]
            ct = ctypes.get(mn.elem, "double")
            if mn.om_id in param_nodes and tag not in ("LoadCursorStatic", "LoadCursor", "Store"):
                op = Op("param", (param_nodes[mn.om_id],), ct)
            elif tag == "LoadCursorStatic":
                op = Op("load", (fields.index(inst[2]), tuple(inst[1])), ct)
            elif tag == "LoadCursor":
                op = Op("load", (temp_index[inst[2]], tuple(inst[1])), ct)
            elif tag == "Imm":
                op = Op("imm", (inst[1],), ct)
            elif tag == "LoadIndex":
                op = Op("index", (inst[1],), ct)
            elif tag == "Store":
                ref[j] = ref[inst[2]]
                continue
            elif tag == "Uniop":
                op = Op("uniop", (inst[1], ref[inst[2]]), ct)
            elif tag == "Binop":
                op = Op("binop", (inst[1], ref[inst[2]], ref[inst[3]]), ct)
            elif tag == "Triop":
                op = Op("triop", (inst[1], ref[inst[2]], ref[inst[3]], ref[inst[4]]), ct)
            elif tag == "Naryop":
                op = Op("naryop", (inst[1], [ref[x] for x in inst[2]]), ct)
            else:
                raise FormuraCompileError(f"cannot lower micro node {tag}")
            ref[j] = len(ops)
            ops.append(op)
        ir = PointIR(mm.dimension, list(mm.axes), extents, names, field_ctypes, ops, [ref[n.root()]], param_list)
        if k in temp_index:
            out.append(Stage(ir, "temp", temp_index[k], k))
        else:
            out.append(Stage(ir, "state", fields.index(n.micro[n.root()].inst[1]), k))
    return out


@dataclass
class StageGroup:
    """Manifest nodes evaluated by one kernel: none of them reads an array
    another member writes, so they can run over the grid together.  `ir` has
    one output per member, in member order; equal loads and equal operations
    of different members are computed once (the same IEEE operations on the
    same operands: the same bits as evaluating each member alone)."""
    ir: PointIR
    members: List[Stage]


def _op_key(op: Op, remap: Dict[int, int]) -> tuple:
    k, a = op.kind, op.args
    if k == "uniop":
        args = (a[0], remap[a[1]])
    elif k == "binop":
        args = (a[0], remap[a[1]], remap[a[2]])
    elif k == "triop":
        args = (a[0],) + tuple(remap[x] for x in a[1:4])
    elif k == "naryop":
        args = (a[0], tuple(remap[x] for x in a[1]))
    elif k == "load":
        args = (a[0], tuple(a[1]))
    else:
        args = tuple(a)
    return (k, args, op.ctype)


def _merge(members: Sequence[Stage]) -> PointIR:
    ops: List[Op] = []
    seen: Dict[tuple, int] = {}
    outputs: List[int] = []
    for st in members:
        remap: Dict[int, int] = {}
        for i, op in enumerate(st.ir.ops):
            key = _op_key(op, remap)
            j = seen.get(key)
            if j is None:
                k, a = key[0], key[1]
                if k == "naryop":
                    args = (a[0], list(a[1]))
                elif k == "load":
                    args = (a[0], a[1])
                else:
                    args = a
                j = len(ops)
                ops.append(Op(k, args, op.ctype))
                seen[key] = j
            remap[i] = j
        outputs.append(remap[st.ir.outputs[0]])
    ir0 = members[0].ir
    return PointIR(ir0.dim, ir0.axes, ir0.extents, ir0.fields, ir0.field_ctypes, ops, outputs, ir0.params)


def group_stages(stages: Sequence[Stage], max_loads: int = 64, max_ops: int = 1200,
                 max_members: int = 16) -> List[StageGroup]:
    """Independent manifest nodes side by side: a node's level is one more
    than the highest level of the stage arrays it reads (0 when it reads only
    state fields); nodes of one level never read each other's arrays and
    become one kernel, split where a group would exceed `max_loads` distinct
    loads, `max_ops` operations or `max_members` outputs.  New state values
    go to the other buffer of the double-buffered state, so a node writing
    one never conflicts with a node reading the current state."""
    nf = len(stages[0].ir.fields) - sum(1 for s in stages if s.target == "temp") if stages else 0
    level_of_array: Dict[int, int] = {}
    levels: Dict[int, List[Stage]] = {}
    for st in stages:
        reads = {op.args[0] for op in st.ir.ops if op.kind == "load"}
        lv = 1 + max((level_of_array[a] for a in reads if a >= nf), default=-1)
        if st.target == "temp":
            level_of_array[st.index] = lv
        levels.setdefault(lv, []).append(st)
    out: List[StageGroup] = []
    for lv in sorted(levels):
        cur: List[Stage] = []
        for st in levels[lv]:
            trial = cur + [st]
            ir = _merge(trial)
            nloads = sum(1 for op in ir.ops if op.kind == "load")
            if cur and (nloads > max_loads or len(ir.ops) > max_ops or len(trial) > max_members):
                out.append(StageGroup(_merge(cur), cur))
                cur = [st]
            else:
                cur = trial
        if cur:
            out.append(StageGroup(_merge(cur), cur))
    return out


def imm_literal(r: Fraction, ctype: str) -> str:
    """C literal of a rational: the nearest double (MPICxx/Translate.hs:585,
    `realToFrac r :: Double` is correctly rounded), printed shortest
    round-trip; `float` nodes (FP32 override) round that double to float."""
    d = float(r)
    s = repr(d)
    if s in ("inf", "-inf", "nan"):
        raise FormuraCompileError(f"literal {r} is not representable")
    if "e" not in s and "." not in s:
        s += ".0"
    if ctype == "float":
        return f"(float){s}"
    return s


def c_expr_of(op: Op, name) -> str:
    """C text of one op given `name(i)` for its operands
    (MPICxx/Translate.hs:585-609: `**` -> pow, `>?`/`<?` -> fmax/fmin chains,
    `<%` -> minmod, `ite` -> ?:, extern -> call)."""
    k = op.kind
    if k == "uniop":
        o, a = op.args
        if o.startswith("external-call/"):
            return f"({o[len('external-call/'):]}({name(a)}))"
        return f"({o}{name(a)})"
    if k == "binop":
        o, a, b = op.args
        if o == "**":
            return f"pow({name(a)},{name(b)})"
        return f"({name(a)}{o}{name(b)})"
    if k == "triop":
        _, c, a, b = op.args
        return f"({name(c)}?{name(a)}:{name(b)})"
    if k == "naryop":
        o, xs = op.args
        parts = [name(x) for x in xs]

        def chain(fn, cs):
            acc = cs[-1]
            for c in reversed(cs[:-1]):
                acc = f"{fn}({c},{acc})"
            return acc
        if o == ">?":
            return chain("fmax", parts)
        if o == "<?":
            return chain("fmin", parts)
        if o == "<%":
            return chain("fmin", ["0.0", chain("fmax", parts)]) + "+" + chain("fmax", ["0.0", chain("fmin", parts)])
        raise FormuraCompileError(f"unsupported N-ary operator: {o}")
    raise FormuraCompileError(f"not an arithmetic op: {k}")
