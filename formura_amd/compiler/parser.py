"""Parser for the Formura stencil language (`.fmr`).

Follows the reference grammar in src/Formura/Parser.hs:
  * tokens and identifiers: Parser.hs:59-176 (alphabetic identifiers start with a
    letter and continue with letters, digits, `_` or `'`; symbolic identifiers are
    runs of printable non-alphanumeric characters other than `"#();[\\]{}`);
  * operator table, Parser.hs:198-208, high to low precedence:
        `.` (right)  >  `**` (left)  >  `*` `/` (left)  >  unary `+` `-`
        >  binary `+` `-` (left)  >  `<?` `>?` `<%` (n-ary, left)  >  comparisons;
  * application by juxtaposition and grid indexing `a[i+1/2,j]`: Parser.hs:241-257;
  * `begin function ... end function` sugar: Parser.hs:304-336;
  * typed / multi-substitution statements: Parser.hs:339-366;
  * program, `dimension` / `axes` declarations and the injected `NX = ...`
    extents: Parser.hs:449-499.

Literals are exact rationals (trifecta's `naturalOrScientific` -> `toRational`),
so `0.05` is 1/20 and `2.3e-9` is 23/10^10.

The AST is made of plain tuples; the first element is the node tag:
  ('imm', Fraction) ('ident', name) ('tuple', [x]) ('uniop', op, x)
  ('binop', op, a, b) ('naryop', op, [x]) ('triop', 'ite', c, a, b)
  ('grid', [(idxname, Fraction)], x) ('vector', name, x) ('apply', f, x)
  ('lambda', lhs, body) ('let', [stmt], body)
Statements: ('typedecl', (modifiers, type), lhs) and ('subst', lhs, rhs).
Types: ('elemtype', name) ('gridtype', [Fraction], t) ('functype',)
  ('tuple', [t]) ('toptype',).
"""
from __future__ import annotations

import unicodedata
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Optional, Tuple

KEYWORDS = {
    "begin", "end", "function", "returns", "let", "in", "fun", "dimension", "axes",
    "if", "then", "else", "const", "extern", "manifest",
    "+", "-", "*", "/", ".", "**", "::", "=", ",",
}
MINMAX_OPS = {"<?", ">?", "<%"}
COMPARISON_OPS = {"<=", "==", "!=", ">=", "<", ">"}
KEYWORDS |= MINMAX_OPS | COMPARISON_OPS
ELEMENT_TYPENAMES = ["void", "bool", "int", "Rational", "float", "double", "real", "Real"]


class FormuraSyntaxError(Exception):
    pass


def _is_alpha0(c: str) -> bool:
    return c.isalpha()


def _is_alpha1(c: str) -> bool:
    return c.isalnum() or c in "_'"


def _is_symbol(c: str) -> bool:
    if c.isspace() or _is_alpha1(c) or c in "\"#();[\\]{},":
        return False
    return c.isprintable() and unicodedata.category(c)[0] != "C"


@dataclass
class Token:
    kind: str       # 'num', 'alpha', 'sym', 'punct', 'nl', 'eof'
    text: str
    line: int
    value: Optional[Fraction] = None


def tokenize(src: str) -> List[Token]:
    toks: List[Token] = []
    i, line, n = 0, 1, len(src)
    while i < n:
        c = src[i]
        if c == "\n":
            toks.append(Token("nl", "\n", line))
            line += 1
            i += 1
        elif c == "\r" or (c.isspace()):
            i += 1
        elif c == "#":
            while i < n and src[i] != "\n":
                i += 1
        elif c == "\\":
            # line continuation: backslash, optional blanks, newline
            j = i + 1
            while j < n and src[j] in " \t\r":
                j += 1
            if j < n and src[j] == "\n":
                line += 1
                i = j + 1
            else:
                raise FormuraSyntaxError(f"line {line}: stray backslash")
        elif c in "();[],{}":
            toks.append(Token("punct", c, line))
            i += 1
        elif c.isdigit():
            j = i
            while j < n and src[j].isdigit():
                j += 1
            if j + 1 < n and src[j] == "." and src[j + 1].isdigit():
                j += 1
                while j < n and src[j].isdigit():
                    j += 1
            if j < n and src[j] in "eE":
                k = j + 1
                if k < n and src[k] in "+-":
                    k += 1
                if k < n and src[k].isdigit():
                    while k < n and src[k].isdigit():
                        k += 1
                    j = k
            text = src[i:j]
            toks.append(Token("num", text, line, Fraction(text)))
            i = j
        elif _is_alpha0(c):
            j = i + 1
            while j < n and _is_alpha1(src[j]):
                j += 1
            toks.append(Token("alpha", src[i:j], line))
            i = j
        elif _is_symbol(c):
            j = i + 1
            while j < n and _is_symbol(src[j]):
                j += 1
            toks.append(Token("sym", src[i:j], line))
            i = j
        else:
            raise FormuraSyntaxError(f"line {line}: unexpected character {c!r}")
    toks.append(Token("eof", "", line))
    return toks


@dataclass
class Program:
    """A parsed `.fmr` program (reference: Syntax.hs:292-297)."""
    dimension: int
    axes: List[str]
    statements: List[tuple]
    numerical_config: object = None
    extents: List[Tuple[str, int]] = field(default_factory=list)


class _Parser:
    def __init__(self, toks: List[Token]):
        self.toks = toks
        self.p = 0

    # -- token helpers -------------------------------------------------
    def peek(self, k: int = 0) -> Token:
        return self.toks[min(self.p + k, len(self.toks) - 1)]

    def next(self) -> Token:
        t = self.toks[self.p]
        self.p += 1
        return t

    def err(self, msg: str):
        t = self.peek()
        raise FormuraSyntaxError(f"line {t.line}: {msg} (at {t.text!r})")

    def at_kw(self, k: str) -> bool:
        t = self.peek()
        return t.kind in ("alpha", "sym") and t.text == k

    def at_punct(self, c: str) -> bool:
        t = self.peek()
        return t.kind == "punct" and t.text == c

    def expect_kw(self, k: str):
        if not self.at_kw(k):
            self.err(f"expected {k!r}")
        self.next()

    def expect_punct(self, c: str):
        if not self.at_punct(c):
            self.err(f"expected {c!r}")
        self.next()

    def skip_nl(self):
        while self.peek().kind == "nl":
            self.next()

    def at_ident(self) -> bool:
        t = self.peek()
        return t.kind in ("alpha", "sym") and t.text not in KEYWORDS

    # -- program -------------------------------------------------------
    def program(self):
        decls, stmts = [], []
        self.skip_delims()
        while self.peek().kind != "eof":
            if self.at_kw("dimension"):
                self.next()
                self.expect_kw("::")
                t = self.next()
                if t.kind != "num" or t.value.denominator != 1:
                    self.err("dimension must be a natural number")
                decls.append(("dimension", int(t.value)))
            elif self.at_kw("axes"):
                self.next()
                self.expect_kw("::")
                names = [self.ident_name()]
                while self.at_punct(","):
                    self.next()
                    names.append(self.ident_name())
                decls.append(("axes", names))
            else:
                stmts.extend(self.statement_compound())
            if self.peek().kind == "eof":
                break
            if not self.at_delim():
                self.err("expected end of statement")
            self.skip_delims()
        return decls, stmts

    def at_delim(self) -> bool:
        return self.peek().kind == "nl" or self.at_punct(";")

    def skip_delims(self):
        while self.at_delim():
            self.next()

    def ident_name(self) -> str:
        if not self.at_ident():
            self.err("expected identifier")
        return self.next().text

    # -- statements ----------------------------------------------------
    def binding(self, terminators) -> List[tuple]:
        stmts = []
        self.skip_delims()
        while not any(self.at_kw(k) for k in terminators) and self.peek().kind != "eof":
            stmts.extend(self.statement_compound())
            if any(self.at_kw(k) for k in terminators):
                break
            if not self.at_delim():
                self.err("expected end of statement")
            self.skip_delims()
        return stmts

    def statement_compound(self) -> List[tuple]:
        if self.at_kw("begin"):
            return self.function_sugar()
        return self.type_value_statements()

    def function_sugar(self) -> List[tuple]:
        # Parser.hs:304-336
        self.expect_kw("begin")
        self.expect_kw("function")
        save = self.p
        try:
            fn = ("ident", self.ident_name())
            inx = self.tuple_of(self.lexpr)
            self.expect_kw("returns")
            outx = self.rexpr()
        except FormuraSyntaxError:
            self.p = save
            outx = self.rexpr()
            self.expect_kw("=")
            fn = ("ident", self.ident_name())
            inx = self.tuple_of(self.lexpr)
        if not self.at_delim():
            self.err("expected newline after function header")
        body = self.binding(("end",))
        self.expect_kw("end")
        self.expect_kw("function")
        return [("subst", fn, ("lambda", inx, ("let", body, outx)))]

    def type_value_statements(self) -> List[tuple]:
        # Parser.hs:339-366
        save = self.p
        mtype = None
        try:
            mt = self.mod_type_expr()
            if self.at_kw("::"):
                self.next()
                mtype = mt
            else:
                self.p = save
        except FormuraSyntaxError:
            self.p = save
        if mtype is not None:
            pairs = [self.lhs_and_maybe_rhs()]
            while self.at_punct(","):
                self.next()
                pairs.append(self.lhs_and_maybe_rhs())
            out = [("typedecl", mtype, l) for l, _ in pairs]
            out += [("subst", l, r) for l, r in pairs if r is not None]
            return out
        lhs = self.lexpr()
        self.expect_kw("=")
        rhs = self.rexpr()
        return [("subst", lhs, rhs)]

    def lhs_and_maybe_rhs(self):
        lhs = self.lexpr()
        rhs = None
        if self.at_kw("="):
            self.next()
            rhs = self.rexpr()
        return lhs, rhs

    # -- types ----------------------------------------------------------
    def mod_type_expr(self):
        mods = []
        while self.peek().text in ("const", "extern", "manifest") and self.peek().kind == "alpha":
            mods.append(self.next().text)
        t = None
        if self.peek().kind == "alpha" and (self.peek().text in ELEMENT_TYPENAMES or self.peek().text == "function") \
                or self.at_punct("("):
            t = self.type_fexpr()
        while self.peek().text in ("const", "extern", "manifest") and self.peek().kind == "alpha":
            mods.append(self.next().text)
        if not mods and t is None:
            self.err("expected a type")
        return (tuple(mods), t if t is not None else ("toptype",))

    def type_aexpr(self):
        if self.at_punct("("):
            return self.tuple_of(self.type_fexpr)
        t = self.peek()
        if t.kind == "alpha" and t.text in ELEMENT_TYPENAMES:
            self.next()
            return ("elemtype", t.text)
        if t.kind == "alpha" and t.text == "function":
            self.next()
            return ("functype",)
        self.err("expected a type")

    def type_fexpr(self):
        t = self.type_aexpr()
        while True:
            if self.at_punct("["):
                self.next()
                offs = []
                if not self.at_punct("]"):
                    offs.append(self.const_rational_expr())
                    while self.at_punct(","):
                        self.next()
                        offs.append(self.const_rational_expr())
                self.expect_punct("]")
                t = ("gridtype", offs, t)
            elif self.at_punct("(") and self.peek(1).kind == "num":
                self.next()
                n = self.next()
                self.expect_punct(")")
                t = ("vectortype", int(n.value), t)
            else:
                return t

    # -- l-expressions -------------------------------------------------
    def lexpr(self):
        if self.at_punct("("):
            f = self.tuple_of(self.lexpr)
        else:
            f = ("ident", self.ident_name())
        while True:
            if self.at_punct("["):
                f = ("grid", self.grid_indices(), f)
            elif self.at_punct("(") and self.peek(1).kind == "alpha" and self.peek(2).text == ")":
                self.next()
                name = self.next().text
                self.next()
                f = ("vector", name, f)
            else:
                return f

    def tuple_of(self, p):
        # a one-element tuple is a parenthesised expression (Parser.hs:168-176)
        self.expect_punct("(")
        xs = []
        if not self.at_punct(")"):
            xs.append(p())
            while self.at_punct(","):
                self.next()
                xs.append(p())
        self.expect_punct(")")
        if len(xs) == 1:
            return xs[0]
        return ("tuple", xs)

    def grid_indices(self):
        # Parser.hs:177-192, n+k patterns such as `i+1/2`, `j`, `-1`
        self.expect_punct("[")
        xs = []
        if not self.at_punct("]"):
            xs.append(self.n_plus_k())
            while self.at_punct(","):
                self.next()
                xs.append(self.n_plus_k())
        self.expect_punct("]")
        return xs

    def n_plus_k(self):
        name = ""
        if self.peek().kind == "alpha" and self.peek().text not in KEYWORDS:
            name = self.next().text
        k = Fraction(0)
        if self.at_kw("+") or self.at_kw("-"):
            s = self.next().text
            k = self.const_rational_expr()
            if s == "-":
                k = -k
        elif not name and not (self.at_punct(",") or self.at_punct("]")):
            k = self.const_rational_expr()
        # an empty slot (`a[,j+1]`) is the pattern "" + 0, as in Parser.hs:186-192
        if not (self.at_punct(",") or self.at_punct("]")):
            self.err("malformed grid index")
        return (name, k)

    def const_rational_expr(self) -> Fraction:
        # Parser.hs:413-426: +, -, *, / over literals
        e = self.expr_of(self.imm_only)
        return _eval_const(e)

    def imm_only(self):
        t = self.peek()
        if t.kind == "num":
            self.next()
            return ("imm", t.value)
        if self.at_punct("("):
            return self.tuple_of(lambda: self.expr_of(self.imm_only))
        self.err("expected a rational literal")

    # -- r-expressions ---------------------------------------------------
    def rexpr(self):
        return self.expr_of(self.fexpr)

    def expr_of(self, term):
        return self._level7(term)

    def _level7(self, term):
        a = self._level6(term)
        if self.peek().kind == "sym" and self.peek().text in COMPARISON_OPS:
            op = self.next().text
            b = self._level6(term)
            a = ("binop", op, a, b)
        return a

    def _level6(self, term):
        a = self._level5(term)
        while self.peek().kind == "sym" and self.peek().text in MINMAX_OPS:
            op = self.next().text
            b = self._level5(term)
            aparts = a[2] if a[0] == "naryop" and a[1] == op else [a]
            bparts = b[2] if b[0] == "naryop" and b[1] == op else [b]
            a = ("naryop", op, aparts + bparts)
        return a

    def _level5(self, term):
        a = self._level4(term)
        while self.at_kw("+") or self.at_kw("-"):
            op = self.next().text
            b = self._level4(term)
            a = ("binop", op, a, b)
        return a

    def _level4(self, term):
        if self.at_kw("+") or self.at_kw("-"):
            op = self.next().text
            return ("uniop", op, self._level3(term))
        return self._level3(term)

    def _level3(self, term):
        a = self._level2(term)
        while self.at_kw("*") or self.at_kw("/"):
            op = self.next().text
            b = self._level2(term)
            a = ("binop", op, a, b)
        return a

    def _level2(self, term):
        a = self._level1(term)
        while self.at_kw("**"):
            self.next()
            b = self._level1(term)
            a = ("binop", "**", a, b)
        return a

    def _level1(self, term):
        a = term()
        if self.at_kw("."):
            self.next()
            b = self._level1(term)
            a = ("binop", ".", a, b)
        return a

    def fexpr(self):
        # function application chain, Parser.hs:241-257
        f = self.aexpr()
        while True:
            if self.at_punct("["):
                f = ("grid", self.grid_indices(), f)
                continue
            if self.starts_aexpr():
                x = self.aexpr()
                f = ("apply", f, x)
                continue
            return f

    def starts_aexpr(self) -> bool:
        t = self.peek()
        if t.kind == "punct":
            return t.text == "("
        if t.kind == "num":
            return True
        if t.kind in ("alpha", "sym"):
            if t.text in ("let", "fun", "if"):
                return True
            return t.text not in KEYWORDS
        return False

    def aexpr(self):
        t = self.peek()
        if self.at_punct("("):
            return self.tuple_of(self.rexpr)
        if self.at_kw("let"):
            self.next()
            b = self.binding(("in",))
            self.expect_kw("in")
            return ("let", b, self.rexpr())
        if self.at_kw("fun"):
            self.next()
            lhs = self.tuple_of(self.lexpr)
            return ("lambda", lhs, self.rexpr())
        if self.at_kw("if"):
            self.next()
            c = self.rexpr()
            self.expect_kw("then")
            a = self.rexpr()
            self.expect_kw("else")
            b = self.rexpr()
            return ("triop", "ite", c, a, b)
        if t.kind == "num":
            self.next()
            return ("imm", t.value)
        if self.at_ident():
            return ("ident", self.next().text)
        self.err("expected an expression")


def _eval_const(e) -> Fraction:
    tag = e[0]
    if tag == "imm":
        return e[1]
    if tag == "uniop":
        v = _eval_const(e[2])
        return v if e[1] == "+" else -v
    if tag == "binop":
        a, b = _eval_const(e[2]), _eval_const(e[3])
        return {"+": a + b, "-": a - b, "*": a * b, "/": a / b}[e[1]] if e[1] in "+-*/" else _bad_const(e[1])
    raise FormuraSyntaxError("unsupported operator in const rational expression")


def _bad_const(op):
    raise FormuraSyntaxError(f"unsupported operator {op!r} in const rational expression")


def desugar_grid_dims(node, dim: int):
    """Pad every grid index / grid type to the declared dimension with zeros
    (reference: Desugar.hs:36-57)."""
    if isinstance(node, tuple):
        if node and node[0] == "grid":
            idx = list(node[1])[:dim] + [("", Fraction(0))] * max(0, dim - len(node[1]))
            return ("grid", idx, desugar_grid_dims(node[2], dim))
        if node and node[0] == "gridtype":
            offs = list(node[1])[:dim] + [Fraction(0)] * max(0, dim - len(node[1]))
            return ("gridtype", offs, desugar_grid_dims(node[2], dim))
        return tuple(desugar_grid_dims(x, dim) for x in node)
    if isinstance(node, list):
        return [desugar_grid_dims(x, dim) for x in node]
    return node


def parse_program(src: str, numerical_config=None) -> Program:
    """Parse `.fmr` source text; inject `NX = intra*mpi` extents as the reference
    does (Parser.hs:486-499) when a numerical config is given."""
    decls, stmts = _Parser(tokenize(src)).program()
    dims = [d[1] for d in decls if d[0] == "dimension"]
    axes = [d[1] for d in decls if d[0] == "axes"]
    if len(dims) != 1:
        raise FormuraSyntaxError("exactly one dimension declaration is required")
    if len(axes) != 1:
        raise FormuraSyntaxError("exactly one axes declaration is required")
    dim, ax = dims[0], axes[0]
    if len(ax) != dim:
        raise FormuraSyntaxError("number of declared axes does not match the declared dimension.")
    extents = []
    if numerical_config is not None:
        g = numerical_config.global_extents()
        extents = [("N" + a.upper(), n) for a, n in zip(ax, g)]
    inject = [("subst", ("ident", name), ("imm", Fraction(n))) for name, n in extents]
    prog = Program(dim, ax, desugar_grid_dims(inject + stmts, dim), numerical_config, extents)
    return prog
