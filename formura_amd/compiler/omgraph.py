"""Orthotope-machine (OM) graph generation: `.fmr` AST -> dataflow DAG of
whole-grid operations.

Restates src/Formura/OrthotopeMachine/Translate.hs:
  * node insertion without constant folding or CSE (Translate.hs:103-116):
    every literal, every operator application is a fresh node;
  * unary / binary / n-ary / if-then-else operators and the element-type meet
    (Translate.hs:140-229, Type.hs:16-45, Graph.hs:112-127);
  * grid indexing turns half-integer offsets into `Shift` nodes
    (Translate.hs:245-270);
  * tuple indexing, function application, lexical closures (Translate.hs:280-330);
  * `let` / statement blocks bound one by one with typed casts and the
    `manifest` / `extern` modifiers (Translate.hs:390-483);
  * global `init` / `step` functions with `Load` / `Store` nodes
    (Translate.hs:505-584).

Node instructions are tuples:
  ('Imm', Fraction) ('Uniop', op, a) ('Binop', op, a, b) ('Triop', op, a, b, c)
  ('Naryop', op, [a..]) ('Shift', (dx,..), a) ('Load', name) ('LoadIndex', axis)
  ('Store', name, a)
Types are ('elem', name) or ('grid', (offsets...), name).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from fractions import Fraction
from typing import Callable, Dict, List, Optional, Tuple

from .parser import COMPARISON_OPS, ELEMENT_TYPENAMES, Program


class FormuraCompileError(Exception):
    pass


_TYPE_RANK = {n: i for i, n in enumerate(ELEMENT_TYPENAMES)}


def meet_elem(a: str, b: str) -> Optional[str]:
    """Element-type meet: the larger of the two in the table
    void < bool < int < Rational < float < double < real < Real (Type.hs:16-45)."""
    ra, rb = _TYPE_RANK.get(a), _TYPE_RANK.get(b)
    if ra is None or rb is None:
        return None
    return ELEMENT_TYPENAMES[max(ra, rb)]


def meet_type(a, b):
    """Meet of node types (Graph.hs:112-127). None stands for TopType."""
    if a is None or b is None:
        return None
    if a[0] == "elem" and b[0] == "elem":
        e = meet_elem(a[1], b[1])
        return None if e is None else ("elem", e)
    if a[0] == "elem" and b[0] == "grid":
        e = meet_elem(a[1], b[2])
        return None if e is None else ("grid", b[1], e)
    if a[0] == "grid" and b[0] == "elem":
        return meet_type(b, a)
    if a[1] != b[1]:
        return None
    e = meet_elem(a[2], b[2])
    return None if e is None else ("grid", a[1], e)


def elem_of(t) -> str:
    return t[1] if t[0] == "elem" else t[2]


@dataclass
class Node:
    inst: tuple
    typ: tuple
    manifest: bool = False
    name: Optional[str] = None      # source variable name (annotation)


# ---------------------------------------------------------------- values
@dataclass
class NodeV:
    nid: int
    typ: tuple


@dataclass
class TupleV:
    xs: list


class FunV:
    """A function value. `fn(arg)` generates the application."""

    def __init__(self, fn: Callable, desc: str = "fun"):
        self.fn = fn
        self.desc = desc

    def __repr__(self):
        return f"FunV({self.desc})"


class OMGenerator:
    """Builds one OM graph (the init graph or the step graph)."""

    def __init__(self, prog: Program):
        self.prog = prog
        self.dim = prog.dimension
        self.graph: Dict[int, Node] = {}

    # -- nodes ---------------------------------------------------------
    def insert(self, inst: tuple, typ) -> NodeV:
        nid = len(self.graph)
        self.graph[nid] = Node(inst, typ)
        return NodeV(nid, typ)

    def imm_value(self, v) -> Optional[Fraction]:
        if isinstance(v, NodeV):
            n = self.graph[v.nid]
            if n.inst[0] == "Imm":
                return n.inst[1]
        return None

    # -- casts -----------------------------------------------------------
    def cast(self, t, v):
        """castVal (Translate.hs:126-134): only retypes, never inserts nodes."""
        if t is None:
            return v
        if isinstance(v, NodeV):
            t0 = v.typ
            if t == t0:
                return v
            if t[0] == "elem" and t0[0] == "elem":
                return v
            if t[0] == "grid" and t0[0] == "elem":
                return NodeV(v.nid, t)
            if t[0] == "grid" and t0[0] == "grid" and t[1] == t0[1]:
                return v
        if isinstance(v, TupleV) and isinstance(t, tuple) and t[0] == "tuple":
            if len(t[1]) != len(v.xs):
                raise FormuraCompileError("tuple length mismatch in cast")
            return TupleV([self.cast(ti, vi) for ti, vi in zip(t[1], v.xs)])
        if isinstance(v, FunV) and t == ("functype",):
            return v
        raise FormuraCompileError(f"cannot convert type {v} to {t}")

    def type_from_ast(self, t):
        """Convert a parsed type expression to a node type (or tuple/function)."""
        tag = t[0]
        if tag == "elemtype":
            return ("elem", t[1])
        if tag == "gridtype":
            inner = self.type_from_ast(t[2])
            if inner[0] != "elem":
                raise FormuraCompileError("nested grid types are not supported")
            return ("grid", tuple(t[1]), inner[1])
        if tag == "functype":
            return ("functype",)
        if tag == "tuple":
            return ("tuple", [self.type_from_ast(x) for x in t[1]])
        if tag == "toptype":
            return None
        raise FormuraCompileError(f"unsupported type {t}")

    # -- operators -------------------------------------------------------
    def go_uniop(self, op, a):
        if isinstance(a, NodeV):
            return self.insert(("Uniop", op, a.nid), a.typ)
        if isinstance(a, FunV):
            return FunV(lambda x: self.go_uniop(op, self.apply(a, x)), f"{op}{a.desc}")
        if isinstance(a, TupleV):
            return TupleV([self.go_uniop(op, x) for x in a.xs])
        raise FormuraCompileError("unexpected path in unary operator")

    def go_binop(self, op, a, b):
        if op == ".":
            return FunV(lambda x: self.apply(a, self.apply(b, x)), "compose")
        if isinstance(a, NodeV) and isinstance(b, NodeV):
            ct = meet_type(a.typ, b.typ)
            if ct is None:
                raise FormuraCompileError(
                    f"there is no common type that can accomodate both hand side: {a.typ} {op} {b.typ}")
            rt = ct
            if op in COMPARISON_OPS:
                rt = ("elem", "bool") if ct[0] == "elem" else ("grid", ct[1], "bool")
            return self.insert(("Binop", op, a.nid, b.nid), rt)
        if isinstance(a, FunV) and isinstance(b, FunV):
            return FunV(lambda x: self.go_binop(op, self.apply(a, x), self.apply(b, x)), "binop-ff")
        if isinstance(a, FunV) and isinstance(b, NodeV):
            return FunV(lambda x: self.go_binop(op, self.apply(a, x), b), "binop-fx")
        if isinstance(a, NodeV) and isinstance(b, FunV):
            return FunV(lambda x: self.go_binop(op, a, self.apply(b, x)), "binop-xf")
        if isinstance(a, TupleV) and isinstance(b, TupleV):
            if len(a.xs) != len(b.xs):
                raise FormuraCompileError("tuple length mismatch.")
            return TupleV([self.go_binop(op, x, y) for x, y in zip(a.xs, b.xs)])
        if isinstance(a, NodeV) and isinstance(b, TupleV):
            return TupleV([self.go_binop(op, a, y) for y in b.xs])
        if isinstance(a, TupleV) and isinstance(b, NodeV):
            return TupleV([self.go_binop(op, x, b) for x in a.xs])
        raise FormuraCompileError(f"unexpected path in binary operator {op}")

    def go_triop(self, op, a, b, c):
        if isinstance(a, TupleV) or isinstance(b, TupleV) or isinstance(c, TupleV):
            n = max(len(v.xs) for v in (a, b, c) if isinstance(v, TupleV))
            pick = lambda v, i: v.xs[i] if isinstance(v, TupleV) else v
            return TupleV([self.go_triop(op, pick(a, i), pick(b, i), pick(c, i)) for i in range(n)])
        if op != "ite" or not all(isinstance(v, NodeV) for v in (a, b, c)):
            raise FormuraCompileError(f"unexpected path in trinary operator {op}")
        if elem_of(a.typ) != "bool":
            raise FormuraCompileError("the first argument of if-expr must be of type bool")
        bct = meet_type(b.typ, c.typ)
        if meet_type(a.typ, bct) is None:
            raise FormuraCompileError("Type mismatch in if-then-else expr")
        return self.insert(("Triop", op, a.nid, b.nid, c.nid), bct)

    def go_naryop(self, op, xs):
        # Translate.hs:158-181: the node type is the meet of all operands
        if any(isinstance(x, TupleV) for x in xs):
            n = max(len(x.xs) for x in xs if isinstance(x, TupleV))
            return TupleV([self.go_naryop(op, [x.xs[i] if isinstance(x, TupleV) else x for x in xs])
                           for i in range(n)])
        t = None
        for i, x in enumerate(reversed(xs)):
            if not isinstance(x, NodeV):
                raise FormuraCompileError("unexpected path in N-ary operator")
            t = x.typ if i == 0 else meet_type(x.typ, t)
            if t is None:
                raise FormuraCompileError(f"there is no common type for N-ary operator {op}")
        return self.insert(("Naryop", op, [x.nid for x in xs]), t)

    # -- grids -------------------------------------------------------------
    def go_grid(self, npks, v):
        if isinstance(v, TupleV):
            return TupleV([self.go_grid(npks, x) for x in v.xs])
        if not isinstance(v, NodeV):
            raise FormuraCompileError(f"unexpected pattern in gen of grid: {v}")
        if v.typ[0] == "elem":
            return v
        offs0 = v.typ[1]
        ks = [k for _, k in npks]
        if len(ks) != len(offs0):
            raise FormuraCompileError("grid index dimension mismatch")
        new_pos = [o - k for o, k in zip(offs0, ks)]
        int_off = [math.floor(p) for p in new_pos]
        new_off = tuple(p - i for p, i in zip(new_pos, int_off))
        t1 = ("grid", new_off, v.typ[2])
        if all(i == 0 for i in int_off):
            return NodeV(v.nid, t1)
        return self.insert(("Shift", tuple(-i for i in int_off), v.nid), t1)

    # -- application -----------------------------------------------------
    def apply(self, f, x):
        if isinstance(f, TupleV):
            r = self.imm_value(x)
            if r is None:
                raise FormuraCompileError("tuple applied to non-constant integer")
            if r.denominator != 1:
                raise FormuraCompileError("non-integer indexing in tuple access")
            n = int(r)
            if n < 0 or n >= len(f.xs):
                raise FormuraCompileError("tuple access out of bounds")
            return f.xs[n]
        if isinstance(f, FunV):
            return f.fn(x)
        raise FormuraCompileError("unexpected combination of application")

    # -- l-expressions -------------------------------------------------------
    @staticmethod
    def names_of_lhs(l):
        tag = l[0]
        if tag == "ident":
            return ("ident", l[1])
        if tag in ("grid",):
            return OMGenerator.names_of_lhs(l[2])
        if tag == "vector":
            return OMGenerator.names_of_lhs(l[2])
        if tag == "tuple":
            return ("tuple", [OMGenerator.names_of_lhs(x) for x in l[1]])
        raise FormuraCompileError(f"bad l-expression {l}")

    def match_to_idents(self, names, v) -> List[Tuple[str, object]]:
        if names[0] == "tuple":
            if isinstance(v, TupleV):
                if len(names[1]) != len(v.xs):
                    raise FormuraCompileError("tuple length mismatch.")
                out = []
                for n, x in zip(names[1], v.xs):
                    out += self.match_to_idents(n, x)
                return out
            if isinstance(v, tuple) and v and v[0] == "tuple":
                if len(names[1]) != len(v[1]):
                    raise FormuraCompileError("tuple length mismatch.")
                out = []
                for n, x in zip(names[1], v[1]):
                    out += self.match_to_idents(n, x)
                return out
            raise FormuraCompileError("the LHS expects a tuple, but RHS is not a tuple.")
        return [(names[1], v)]

    def match_to_lhs(self, l, v):
        return self.match_to_idents(self.names_of_lhs(l), v)

    # -- expressions -------------------------------------------------------
    def gen(self, e, env: dict):
        tag = e[0]
        if tag == "imm":
            return self.insert(("Imm", e[1]), ("elem", "Rational"))
        if tag == "ident":
            if e[1] not in env:
                raise FormuraCompileError(f"undefined variable: {e[1]}")
            return env[e[1]]
        if tag == "tuple":
            return TupleV([self.gen(x, env) for x in e[1]])
        if tag == "uniop":
            return self.go_uniop(e[1], self.gen(e[2], env))
        if tag == "binop":
            a = self.gen(e[2], env)
            b = self.gen(e[3], env)
            return self.go_binop(e[1], a, b)
        if tag == "triop":
            a = self.gen(e[2], env)
            b = self.gen(e[3], env)
            c = self.gen(e[4], env)
            return self.go_triop(e[1], a, b, c)
        if tag == "naryop":
            return self.go_naryop(e[1], [self.gen(x, env) for x in e[2]])
        if tag == "grid":
            return self.go_grid(e[1], self.gen(e[2], env))
        if tag == "apply":
            f = self.gen(e[1], env)
            x = self.gen(e[2], env)
            return self.apply(f, x)
        if tag == "lambda":
            lhs, body = e[1], e[2]
            closure = env

            def fn(arg, lhs=lhs, body=body, closure=closure):
                binds = dict(self.match_to_lhs(lhs, arg))
                env2 = dict(closure)
                env2.update(binds)
                return self.gen(body, env2)
            return FunV(fn, "lambda")
        if tag == "let":
            env2 = self.with_bindings(e[1], env)
            return self.gen(e[2], env2)
        raise FormuraCompileError(f"cannot generate code for {tag}")

    def with_bindings(self, stmts: List[tuple], env: dict) -> dict:
        """Bind statements one by one (Translate.hs:413-483)."""
        type_dict: Dict[str, object] = {}
        mod_dict: Dict[str, List[str]] = {}
        for s in stmts:
            if s[0] != "typedecl":
                continue
            mods, t = s[1]
            names = self.names_of_lhs(s[2])
            for n in _flatten_names(names):
                for m in mods:
                    mod_dict.setdefault(n[1], []).append(m)
            if t[0] != "toptype":
                for name, ti in self.match_to_idents(names, t):
                    type_dict[name] = self.type_from_ast(ti)
        env = dict(env)
        for name, mods in mod_dict.items():
            if "extern" in mods and type_dict.get(name) == ("functype",):
                env[name] = FunV(lambda q, name=name: self.go_uniop("external-call/" + name, q),
                                 "extern " + name)
        for s in stmts:
            if s[0] != "subst":
                continue
            lhs, rhs = s[1], s[2]
            idx_env = dict(env)
            if lhs[0] == "grid":
                for ax, (iname, _) in enumerate(lhs[1]):
                    if iname:
                        idx_env[iname] = self.insert(("LoadIndex", ax), ("elem", "Rational"))
            v0 = self.gen(rhs, idx_env)
            if lhs[0] == "grid":
                neg = [(n, -k) for n, k in lhs[1]]
                lvs = [(n, self.go_grid(neg, v)) for n, v in self.match_to_lhs(lhs[2], v0)]
            else:
                lvs = self.match_to_lhs(lhs, v0)
            for name, v in lvs:
                if name in type_dict:
                    v = self.cast(type_dict[name], v)
                self._annotate(v, name, "manifest" in mod_dict.get(name, []))
                env[name] = v
        return env

    def _annotate(self, v, name, is_manifest):
        if isinstance(v, NodeV):
            node = self.graph[v.nid]
            if node.name is None:
                node.name = name
            if is_manifest:
                node.manifest = True
                node.name = name
        elif isinstance(v, TupleV):
            for x in v.xs:
                self._annotate(x, name, is_manifest)


def _flatten_names(n):
    if n[0] == "ident":
        return [n]
    out = []
    for x in n[1]:
        out += _flatten_names(x)
    return out


@dataclass
class OMProgram:
    """Output of `gen_om_program` (Translate.hs:556-584)."""
    dimension: int
    axes: List[str]
    init_graph: Dict[int, Node]
    step_graph: Dict[int, Node]
    state_signature: Dict[str, tuple]     # state name -> grid type
    state_order: List[str]                # in `step` argument order
    numerical_config: object = None


def _lookup_toplevel(prog: Program, name: str):
    found = [s[2] for s in prog.statements if s[0] == "subst" and s[1] == ("ident", name)]
    if not found:
        raise FormuraCompileError(f"Identifier `{name}` not found.")
    if len(found) > 1:
        raise FormuraCompileError(f"Multiple declaration of identifier `{name}` found.")
    return found[0]


def _gen_global_function(prog: Program, input_types, output_lhs, fundef):
    """genGlobalFunction (Translate.hs:505-535)."""
    if fundef[0] != "lambda":
        raise FormuraCompileError("Identifier specified for function generation is not of function type.")
    g = OMGenerator(prog)
    env = g.with_bindings(prog.statements, {})
    lhs, body = fundef[1], fundef[2]
    binds = {}
    for name, t in g.match_to_lhs(lhs, input_types):
        binds[name] = g.insert(("Load", name), t)
    env2 = dict(env)
    env2.update(binds)
    ret = g.gen(body, env2)
    out_types = []
    for name, v in g.match_to_lhs(output_lhs, ret):
        if not isinstance(v, NodeV):
            raise FormuraCompileError("The return type of a global function must be a tuple of grids.")
        st = g.insert(("Store", name, v.nid), ("elem", "void"))
        g.graph[st.nid].manifest = True
        g.graph[st.nid].name = name + "_next"
        out_types.append(v.typ)
    ret_type = ("tuple", out_types) if len(out_types) != 1 else out_types[0]
    return g.graph, ret, ret_type


def gen_om_program(prog: Program) -> OMProgram:
    step_def = _lookup_toplevel(prog, "step")
    if step_def[0] != "lambda":
        raise FormuraCompileError("`step` must be a function")
    lhs_of_step = step_def[1]
    init_def = _lookup_toplevel(prog, "init")
    init_graph, init_ret, init_type = _gen_global_function(prog, ("tuple", []), lhs_of_step, init_def)
    names = [n[1] for n in _flatten_names(OMGenerator.names_of_lhs(lhs_of_step))]
    in_types = ("tuple", [v.typ for v in (init_ret.xs if isinstance(init_ret, TupleV) else [init_ret])])
    if len(names) == 1:
        in_types = in_types[1][0]
    step_graph, _, step_type = _gen_global_function(prog, in_types, lhs_of_step, step_def)
    if step_type != init_type:
        raise FormuraCompileError(
            f"the return type of step : {step_type} must match the return type of init : {init_type}")
    sig_types = init_type[1] if names and len(names) > 1 else [init_type]
    sig = dict(zip(names, sig_types))
    for n, t in sig.items():
        if t[0] != "grid":
            raise FormuraCompileError(f"state variable {n} must be a grid")
    return OMProgram(prog.dimension, prog.axes, init_graph, step_graph, sig, names, prog.numerical_config)
