"""Manifestation: OM graph -> manifest-machine (MM) program, and the
temporal-blocking fold.

Restates src/Formura/OrthotopeMachine/Manifestation.hs:93-245: every node marked
manifest (`manifest ::` declarations and the `Store` nodes) becomes one
micro-DAG evaluated at a cursor; non-manifest ("delayed") nodes are inlined at
the cursor, `Shift v x` moves the cursor by v, `Load name` becomes
`LoadCursorStatic cursor name`, and a reference to another manifest node
becomes `LoadCursor cursor nid`. Micro nodes are shared per (cursor, OM node)
(Manifestation.hs:72-82).

`temporal_blocking` restates TemporalBlocking.hs:49-134: T copies of the step
graph with node ids strided, state loads of copy i>0 rewired to the stores of
copy i-1, and the inner stores replaced by unary `+` (a value no-op).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

from .omgraph import FormuraCompileError, Node, OMProgram, elem_of


@dataclass
class MicroNode:
    inst: tuple
    elem: str                      # element type name of the value
    om_id: int                     # MMLocation: originating OM node
    cursor: Tuple[int, ...]        # MMLocation: cursor


@dataclass
class MMNode:
    micro: Dict[int, MicroNode]
    typ: tuple                     # OM node type (grid / elem / void)
    name: str
    om_id: int

    def root(self) -> int:
        """The micro node that holds the value of the manifest node: the last
        inserted one (the reference takes `M.findMax`)."""
        return max(self.micro)


@dataclass
class MMProgram:
    dimension: int
    axes: List[str]
    step: Dict[int, MMNode]                 # manifest OM id -> micro DAG
    state_signature: Dict[str, tuple]
    state_order: List[str]
    numerical_config: object = None
    om: OMProgram = None


class _Manifester:
    def __init__(self, graph: Dict[int, Node], dim: int):
        self.g = graph
        self.dim = dim
        self.is_manifest = {k for k, n in graph.items() if n.manifest}

    def build(self, nid: int) -> MMNode:
        self.micro: Dict[int, MicroNode] = {}
        self.idmap: Dict[tuple, int] = {}
        self.delayed(tuple([0] * self.dim), nid)
        n = self.g[nid]
        return MMNode(self.micro, n.typ, n.name or f"node{nid}", nid)

    def ins(self, cursor, nid, inst) -> int:
        key = (cursor, nid)
        if key not in self.idmap:
            self.idmap[key] = len(self.micro)
        j = self.idmap[key]
        t = self.g[nid].typ
        self.micro[j] = MicroNode(inst, elem_of(t), nid, cursor)
        return j

    def at(self, cursor, nid) -> int:
        if nid in self.is_manifest:
            return self.ins(cursor, nid, ("LoadCursor", cursor, nid))
        return self.delayed(cursor, nid)

    def delayed(self, cursor, nid) -> int:
        inst = self.g[nid].inst
        tag = inst[0]
        if tag == "Imm":
            return self.ins(cursor, nid, inst)
        if tag == "Uniop":
            a = self.at(cursor, inst[2])
            return self.ins(cursor, nid, ("Uniop", inst[1], a))
        if tag == "Binop":
            a = self.at(cursor, inst[2])
            b = self.at(cursor, inst[3])
            return self.ins(cursor, nid, ("Binop", inst[1], a, b))
        if tag == "Triop":
            a = self.at(cursor, inst[2])
            b = self.at(cursor, inst[3])
            c = self.at(cursor, inst[4])
            return self.ins(cursor, nid, ("Triop", inst[1], a, b, c))
        if tag == "Naryop":
            xs = [self.at(cursor, x) for x in inst[2]]
            return self.ins(cursor, nid, ("Naryop", inst[1], xs))
        if tag == "Shift":
            c2 = tuple(c + v for c, v in zip(cursor, inst[1]))
            return self.at(c2, inst[2])
        if tag == "Load":
            return self.ins(cursor, nid, ("LoadCursorStatic", cursor, inst[1]))
        if tag == "LoadIndex":
            return self.ins(cursor, nid, inst)
        if tag == "Store":
            a = self.at(cursor, inst[2])
            return self.ins(cursor, nid, ("Store", inst[1], a))
        raise FormuraCompileError(f"manifestation path unimplemented for keyword: {tag}")


def gen_mm_program(om: OMProgram) -> MMProgram:
    nc = om.numerical_config
    # Nanoblock unroll (`nbu<axis>N`, Manifestation.hs:146-158) replicates
    # the micro DAG at nbu cursors per loop iteration: it changes no value.
    # It is honoured where it shows -- the state arrays' upper margin grows
    # by nbu - 1 (abi.plan_geometry, as MPICxx/Translate.hs:942) -- and the
    # GPU kernels, which choose their own unrolling, need no more.
    m = _Manifester(om.step_graph, om.dimension)
    step = {k: m.build(k) for k in sorted(m.is_manifest)}
    return MMProgram(om.dimension, om.axes, step, om.state_signature, om.state_order, nc, om)


def temporal_blocking(mm: MMProgram, T: int) -> Dict[int, MMNode]:
    """The T-fold step graph of TemporalBlocking.hs:49-134 (returned, not
    stored: the GPU kernels fuse time steps without it)."""
    g0 = mm.step
    lo, hi = min(g0), max(g0)
    stride = hi - lo + 1
    store_of = {}
    for k, n in g0.items():
        for mn in n.micro.values():
            if mn.inst[0] == "Store":
                store_of[mn.inst[1]] = k
    out: Dict[int, MMNode] = {}
    for i in range(T):
        for k, n in g0.items():
            micro = {}
            is_store = any(mn.inst[0] == "Store" for mn in n.micro.values())
            for j, mn in n.micro.items():
                inst = mn.inst
                if inst[0] == "LoadCursor":
                    inst = ("LoadCursor", inst[1], inst[2] + i * stride)
                elif inst[0] == "LoadCursorStatic" and i > 0:
                    inst = ("LoadCursor", inst[1], store_of[inst[2]] + (i - 1) * stride)
                if inst[0] == "Store" and i < T - 1:
                    inst = ("Uniop", "+", inst[2])
                micro[j] = MicroNode(inst, mn.elem, mn.om_id + i * stride, mn.cursor)
            typ = n.typ
            if is_store and i < T - 1:
                st_name = [mn.inst[1] for mn in n.micro.values() if mn.inst[0] == "Store"][0]
                typ = mm.state_signature[st_name]
            out[k + i * stride] = MMNode(micro, typ, n.name, k + i * stride)
    return out
