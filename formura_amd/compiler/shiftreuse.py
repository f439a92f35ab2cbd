"""Shift reuse: values of the point function that a neighbouring site has
already computed.

Every op of the point function is evaluated at every site.  If op B, moved by
one site along an axis, is op A -- B's expression with every load offset
shifted by +1 along that axis is A's expression, and neither depends on the
site's coordinates -- then B at site s equals A at site s - e: the same IEEE
operations on the same operands, so the same bits.  A kernel that has A's
value at s - e at hand (the march kernel holds the previous z-site in the
previous lane) can take it instead of evaluating B.  For the Laplacian of the
reference programs, B = c - c[k-1] is the previous site's A = c[k+1] - c.

Expressions are hash-consed: each op gets an integer id for its structure
(kind, operator, C type, ids of its operands; loads by field and offset), and
shifted ids are interned on demand, so the analysis is linear in the op count.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

from .pointir import PointIR


def op_children(op) -> Sequence[int]:
    k = op.kind
    if k == "uniop":
        return (op.args[1],)
    if k == "binop":
        return (op.args[1], op.args[2])
    if k == "triop":
        return tuple(op.args[1:4])
    if k == "naryop":
        return tuple(op.args[1])
    return ()


class _Interner:
    def __init__(self) -> None:
        self.ids: Dict[tuple, int] = {}
        self.keys: List[tuple] = []

    def get(self, key: tuple) -> int:
        i = self.ids.get(key)
        if i is None:
            i = len(self.keys)
            self.ids[key] = i
            self.keys.append(key)
        return i


class ShiftAnalysis:
    """Structure ids of the ops of `ir` and their shifts along the axes."""

    def __init__(self, ir: PointIR) -> None:
        self.ir = ir
        self.dim = ir.dim
        self.I = _Interner()
        self.op_id: List[Optional[int]] = []      # None: depends on the site's coordinates
        for op in ir.ops:
            k = op.kind
            if k == "load":
                f, off = op.args
                key = ("load", f, tuple(off) + (0,) * (3 - len(off)))
            elif k == "imm":
                key = ("imm", op.args[0], op.ctype)
            elif k == "param":
                key = ("param", op.args[0], op.ctype)
            elif k == "index":
                self.op_id.append(None)
                continue
            else:
                ch = [self.op_id[c] for c in op_children(op)]
                if any(c is None for c in ch):
                    self.op_id.append(None)
                    continue
                key = (k, op.args[0], op.ctype, tuple(ch))
            self.op_id.append(self.I.get(key))
        self.first: Dict[int, int] = {}
        for i, sid in enumerate(self.op_id):
            if sid is not None and sid not in self.first:
                self.first[sid] = i
        self._memo: Dict[Tuple[int, int, int], int] = {}

    def shift(self, sid: int, axis: int, d: int) -> int:
        """Id of the expression `sid` with every load moved by d along axis."""
        m = self._memo.get((sid, axis, d))
        if m is not None:
            return m
        key = self.I.keys[sid]
        if key[0] == "load":
            off = list(key[2])
            off[axis] += d
            out = self.I.get(("load", key[1], tuple(off)))
        elif key[0] in ("imm", "param"):
            out = sid
        else:
            out = self.I.get((key[0], key[1], key[2], tuple(self.shift(c, axis, d) for c in key[3])))
        self._memo[(sid, axis, d)] = out
        return out

    def pairs(self, axis: int) -> Dict[int, int]:
        """{B: A}: op B at site s equals op A at site s - e_axis.  Only ops
        with operands (not loads, literals or parameters) are listed as B."""
        out: Dict[int, int] = {}
        for b, op in enumerate(self.ir.ops):
            sid = self.op_id[b]
            if sid is None or op.kind in ("load", "imm", "param", "index"):
                continue
            moved = self.shift(sid, axis, +1)
            if moved == sid:          # no load under it: the same value at every site
                continue
            a = self.first.get(moved)
            if a is not None and a != b and self.ir.ops[a].ctype == op.ctype:
                out[b] = a
        return out


def depends_on(ir: PointIR, a: int, b: int) -> bool:
    """Whether op a uses op b (transitively)."""
    seen = set()
    stack = [a]
    while stack:
        i = stack.pop()
        if i == b:
            return True
        if i in seen:
            continue
        seen.add(i)
        stack.extend(op_children(ir.ops[i]))
    return False
