"""Evaluate the reference's committed generated C, examples-generated/pearson.c,
without compiling it.

That file cannot be built here: it includes a generated header `pearson.h`
that the reference does not contain and <mpi.h>, which this image lacks.  Its
`Formura_Forward` is, however, a flat sequence of doubly nested `for` loops
with constant bounds whose bodies are single IEEE-double statements
(`double a_k = (x op y);`, array copies).  This module parses that text and
evaluates every loop nest with numpy float64 array operations -- one numpy
operation per C statement, each an IEEE-754 double operation like the C one --
so the result is the reference's own generated arithmetic, region plan, ridge
staging and skewed in-place stores, evaluated on the CPU.

Used only by tests (reads the reference tree when it is present).
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

import numpy as np

_DECL = re.compile(r"^double\s+(\w+)((?:\[\d+\])+);", re.M)
_FOR = re.compile(r"for \(int (\w+) = (-?\d+); \1 < (-?\d+); \+\+\1\) \{")
_REF = re.compile(r"^(\w+)\[(\w+)(?: ([+-]) (-?\d+))?\]\[(\w+)(?: ([+-]) (-?\d+))?\]$")
_NUM = re.compile(r"^-?\d+(\.\d+)?(e-?\d+)?$")


class PearsonC:
    def __init__(self, text: str):
        self.arrays: Dict[str, np.ndarray] = {}
        for m in _DECL.finditer(text):
            dims = [int(d) for d in re.findall(r"\[(\d+)\]", m.group(2))]
            self.arrays[m.group(1)] = np.zeros(dims)
        fwd = text[text.index("Formura_Forward (struct Formura_Navigator *navi)"):]
        body = fwd[fwd.index("{") + 1:]
        m = re.search(r"for \(int timestep = 0; timestep < (\d+); \+\+timestep\) \{", body)
        self.outer = int(m.group(1))
        inner = body[m.end():]
        self.nests = self._parse_nests(inner)
        m2 = re.search(r"navi->time_step \+= (\d+);", body)
        self.steps_per_forward = int(m2.group(1))
        init = text[text.index("Formura_Init"):]
        init = init[:init.index("}")]
        self.navi = {k: int(v) for k, v in re.findall(r"navi->(\w+) = (-?\d+);", init)}

    def _parse_nests(self, s: str) -> List[Tuple]:
        nests = []
        pos = 0
        while True:
            m1 = _FOR.search(s, pos)
            if not m1:
                break
            m2 = _FOR.match(s, s.index("for", m1.end()))
            if m2 is None or s[m1.end():m2.start()].strip():
                raise ValueError("unexpected loop structure")
            end = s.index("}", m2.end())
            stmts = [t.strip() for t in s[m2.end():end].split(";") if t.strip()]
            nests.append(((m1.group(1), int(m1.group(2)), int(m1.group(3))),
                          (m2.group(1), int(m2.group(2)), int(m2.group(3))), stmts))
            pos = s.index("}", end + 1) + 1
            if "/* HALFWAYS */" in s[end:pos] or pos >= len(s):
                pass
        return nests

    def _slice(self, ref: str, ax, ay):
        m = _REF.match(ref)
        if not m:
            raise ValueError(f"bad array reference {ref!r}")
        name = m.group(1)

        def rng(var, sign, k, loop):
            lv, lo, hi = loop
            if var != lv:
                raise ValueError(f"index {var} is not the loop variable {lv}")
            off = 0 if sign is None else (int(k) if sign == "+" else -int(k))
            return slice(lo + off, hi + off)
        return name, (rng(m.group(2), m.group(3), m.group(4), ax), rng(m.group(5), m.group(6), m.group(7), ay))

    def _eval(self, e: str, env, ax, ay):
        e = e.strip()
        if e in env:
            return env[e]
        if _NUM.match(e):
            return float(e)
        if e.startswith("(") and e.endswith(")"):
            inner = e[1:-1].strip()
            if inner.startswith("-") and " " not in inner:
                return -self._eval(inner[1:], env, ax, ay)
            parts = inner.split(" ")
            if len(parts) == 3 and parts[1] in "+-*/":
                a = self._eval(parts[0], env, ax, ay)
                b = self._eval(parts[2], env, ax, ay)
                return {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide}[parts[1]](a, b)
            raise ValueError(f"unsupported expression {e!r}")
        name, sl = self._slice(e, ax, ay)
        return self.arrays[name][sl]

    def run_nest(self, nest) -> None:
        ax, ay, stmts = nest
        env = {}
        writes = []
        reads = set()
        for st in stmts:
            lhs, rhs = [t.strip() for t in st.split("=", 1)]
            reads.update(re.findall(r"(\w+)\[", rhs))
            val = self._eval(rhs, env, ax, ay)
            if lhs.startswith("double "):
                env[lhs[len("double "):].strip()] = val
            else:
                name, sl = self._slice(lhs, ax, ay)
                writes.append((name, sl, val))
        for name, sl, val in writes:
            if name in reads:
                raise ValueError("a loop nest that reads and writes one array is not vectorisable")
            target = self.arrays[name][sl]
            target[...] = np.broadcast_to(val, target.shape)

    def forward(self) -> int:
        for _ in range(self.outer):
            for nest in self.nests:
                self.run_nest(nest)
        return self.steps_per_forward
