"""Front end: parser precedence, literals, grid offsets, and the OM / MM graphs
of the reference's own example against the graphs its compiler logged
(examples-generated/pearson.log:13-84)."""
from __future__ import annotations

import os
import re
from fractions import Fraction

import pytest

from formura_amd.compiler.manifest import gen_mm_program, temporal_blocking
from formura_amd.compiler.numerical_config import NumericalConfig
from formura_amd.compiler.omgraph import FormuraCompileError, gen_om_program
from formura_amd.compiler.parser import FormuraSyntaxError, _Parser, parse_program, tokenize


def rexpr(src):
    return _Parser(tokenize(src)).rexpr()


def test_precedence_unary_minus_binds_looser_than_product():
    # Parser.hs:201-208: unary +/- sit between `* /` and binary `+ -`
    assert rexpr("-a*b*c") == ("uniop", "-", ("binop", "*", ("binop", "*", ("ident", "a"), ("ident", "b")),
                                               ("ident", "c")))


def test_precedence_table_order():
    assert rexpr("a - b - c") == ("binop", "-", ("binop", "-", ("ident", "a"), ("ident", "b")), ("ident", "c"))
    assert rexpr("a / b * c") == ("binop", "*", ("binop", "/", ("ident", "a"), ("ident", "b")), ("ident", "c"))
    assert rexpr("a ** b ** c")[2][0] == "binop"          # `**` is left associative
    assert rexpr("f . g . h")[3] == ("binop", ".", ("ident", "g"), ("ident", "h"))   # `.` is right associative
    assert rexpr("a + b * c") == ("binop", "+", ("ident", "a"), ("binop", "*", ("ident", "b"), ("ident", "c")))
    assert rexpr("a <? b <? c") == ("naryop", "<?", [("ident", "a"), ("ident", "b"), ("ident", "c")])
    # application binds tighter than every operator
    assert rexpr("f x + g y") == ("binop", "+", ("apply", ("ident", "f"), ("ident", "x")),
                                  ("apply", ("ident", "g"), ("ident", "y")))


def test_literals_are_exact_decimals():
    assert rexpr("0.05") == ("imm", Fraction(1, 20))
    assert rexpr("2.3e-9") == ("imm", Fraction(23, 10 ** 10))
    assert rexpr("1e-5") == ("imm", Fraction(1, 100000))


def test_grid_half_offsets_become_shifts():
    src = """dimension :: 1
axes :: x
begin function init() returns u
  double [] :: u = 0
end function
begin function u_next = step(u)
  h = u[i+1/2] - u[i-1/2]
  u_next = h[i+1/2] - h[i-1/2]
end function
"""
    om = gen_om_program(parse_program(src))
    g = om.step_graph
    shifts = [n.inst for n in g.values() if n.inst[0] == "Shift"]
    # u[i+1/2] -> Shift(+1) u at offset 1/2; h[i-1/2] -> Shift(-1) h at offset 0
    assert ("Shift", (1,), 0) in [(s[0], s[1], g[s[2]].inst[0] == "Load" and 0) for s in shifts] or \
        any(s[1] == (1,) and g[s[2]].inst == ("Load", "u") for s in shifts)
    assert any(s[1] == (-1,) for s in shifts)
    offs = {n.typ[1] for n in g.values() if n.typ[0] == "grid"}
    assert (Fraction(1, 2),) in offs and (Fraction(0),) in offs


def test_adding_grids_at_different_offsets_is_a_type_error():
    src = """dimension :: 1
axes :: x
begin function init() returns u
  double [] :: u = 0
end function
begin function u_next = step(u)
  u_next = u + u[i+1/2]
end function
"""
    with pytest.raises(FormuraCompileError):
        gen_om_program(parse_program(src))


def test_syntax_error_is_reported():
    with pytest.raises(FormuraSyntaxError):
        parse_program("dimension :: 2\naxes :: x, y\nfoo = (1 + \n")


def test_monitor_interval_rounds_up_to_multiple_of_2T():
    nc = NumericalConfig(intra_node_shape=[8], mpi_grid_shape=[1], temporal_blocking_interval=3,
                         monitor_interval=100)
    assert nc.monitor_interval2() == 102       # MPICxx/Translate.hs:1111-1113
    nc.temporal_blocking_interval = 2
    assert nc.monitor_interval2() == 100


# --------------------------------------------------------------------------
# the reference's logged graphs (only in the build container)

_LOG_NODE = re.compile(r"^(M| ) (\S*)\s+\((\d+),(\w+)F (.*) :: (.*)\)\s*$")


def _parse_log_step_graph(text):
    sec = text[text.index("## Debug print: step graph"):text.index("manifest node ID")]
    nodes = {}
    for line in sec.splitlines()[1:]:
        if not line.strip():
            continue
        m = _LOG_NODE.match(line)
        assert m, line
        nodes[int(m.group(3))] = (m.group(1) == "M", m.group(4), m.group(5), m.group(6))
    ids = re.findall(r"manifest node ID: \[([\d,]+)\]", text)
    manifest = [int(x) for x in ids[1].split(",")]     # the second list is the step graph
    return nodes, manifest


def _frac(s):
    a, b = s.split("%")
    return Fraction(int(a), int(b))


def _log_type(t):
    t = t.strip()
    if t in ("Rational", "void", "double", "float"):
        return ("elem", t)
    m = re.match(r"(\w+)\((.*)\)", t)
    offs = tuple(_frac(x) for x in m.group(2).split(","))
    return ("grid", offs, m.group(1))


def _mine_as_log(inst, shift):
    tag = inst[0]
    if tag == "Imm":
        return ("Imm", inst[1])
    if tag == "Load":
        return ("Load", inst[1])
    if tag == "Store":
        return ("Store", inst[1], inst[2] - shift)
    if tag == "Uniop":
        return ("Uniop", inst[1], inst[2] - shift)
    if tag == "Binop":
        return ("Binop", inst[1], inst[2] - shift, inst[3] - shift)
    if tag == "Shift":
        return ("Shift", inst[1], inst[2] - shift)
    raise AssertionError(tag)


def _log_as_tuple(kind, args):
    if kind == "Imm":
        return ("Imm", _frac(args.strip("()")))
    if kind == "Load":
        return ("Load", args.strip('"'))
    if kind == "Store":
        name, x = args.rsplit(" ", 1)
        return ("Store", name.strip('"'), int(x))
    if kind == "Uniop":
        op, x = args.split(" ")
        return ("Uniop", op.strip('"'), int(x))
    if kind == "Binop":
        op, a, b = args.split(" ")
        return ("Binop", op.strip('"'), int(a), int(b))
    if kind == "Shift":
        v, x = args.rsplit(" ", 1)
        return ("Shift", tuple(int(c) for c in v.strip("()").split(",")), int(x))
    raise AssertionError(kind)


@pytest.fixture(scope="module")
def pearson(reference_tree):
    d = os.path.join(reference_tree, "examples-generated")
    with open(os.path.join(d, "pearson.fmr"), encoding="utf-8") as f:
        src = f.read()
    nc = NumericalConfig.from_file(os.path.join(d, "pearson.nc"))
    with open(os.path.join(d, "pearson.log"), encoding="utf-8") as f:
        log = f.read()
    return src, nc, log


def test_step_graph_matches_reference_log_node_for_node(pearson):
    src, nc, log = pearson
    om = gen_om_program(parse_program(src, nc))
    # the reference build that wrote the log did not inject NX/NY literals
    # (Parser.hs:488-497 of the current compiler does); skip them
    shift = len(nc.intra_node_shape)
    mine = {k - shift: n for k, n in om.step_graph.items() if k >= shift}
    ref, manifest = _parse_log_step_graph(log)
    assert sorted(mine) == sorted(ref)
    for k, (is_m, kind, args, typ) in ref.items():
        n = mine[k]
        assert _mine_as_log(n.inst, shift) == _log_as_tuple(kind, args), k
        assert (n.typ if n.typ[0] == "grid" else n.typ) == _log_type(typ), k
        assert n.manifest == is_m, k
    assert sorted(k - shift for k, n in om.step_graph.items() if n.manifest) == manifest


def _parse_log_mm(log):
    sec = log[log.index("## Debug print: manifested step graph"):log.index("#### Allocation List")]
    out = {}
    for line in sec.splitlines()[1:]:
        if not line.strip():
            continue
        m = re.match(r"^(\S+)\s+\((\d+),fromList \[(.*)\] :: ", line)
        items = re.findall(r"\((\d+),(\w+)F ([^:]*?) :: (\w+)\)", m.group(3))
        out[int(m.group(2))] = [(int(i), k, a.strip(), t) for i, k, a, t in items]
    return out


def _cursor(s, dim):
    s = s.strip()
    if s == "(0..)":
        return (0,) * dim
    return tuple(int(c) for c in s.strip("()").split(","))


def test_manifested_graph_matches_reference_log(pearson):
    src, nc, log = pearson
    om = gen_om_program(parse_program(src, nc))
    mm = gen_mm_program(om)
    shift = len(nc.intra_node_shape)
    ref = _parse_log_mm(log)
    assert sorted(ref) == sorted(k - shift for k in mm.step)
    for k, items in ref.items():
        mine = mm.step[k + shift].micro
        assert sorted(mine) == [i for i, *_ in items]
        for i, kind, args, typ in items:
            inst = mine[i].inst
            if kind == "LoadCursorStatic":
                cur, name = args.rsplit(" ", 1)
                assert inst == ("LoadCursorStatic", _cursor(cur, 2), name.strip('"'))
            elif kind == "LoadCursor":
                cur, nid = args.rsplit(" ", 1)
                assert inst == ("LoadCursor", _cursor(cur, 2), int(nid) + shift)
            elif kind == "Imm":
                assert inst == ("Imm", _frac(args.strip("()")))
            elif kind == "Binop":
                op, a, b = args.split(" ")
                assert inst == ("Binop", op.strip('"'), int(a), int(b))
            elif kind == "Uniop":
                op, a = args.split(" ")
                assert inst == ("Uniop", op.strip('"'), int(a))
            elif kind == "Store":
                name, a = args.rsplit(" ", 1)
                assert inst == ("Store", name.strip('"'), int(a))
            else:
                raise AssertionError(kind)
            assert mine[i].elem == typ


def test_temporal_blocking_fold_structure(pearson):
    src, nc, _ = pearson
    mm = gen_mm_program(gen_om_program(parse_program(src, nc)))
    folded = temporal_blocking(mm, 3)
    assert len(folded) == 3 * len(mm.step)
    stores = [mn.inst for n in folded.values() for mn in n.micro.values() if mn.inst[0] == "Store"]
    assert len(stores) == 2            # only the last copy stores (TemporalBlocking.hs:129-134)
    statics = [mn.inst for n in folded.values() for mn in n.micro.values() if mn.inst[0] == "LoadCursorStatic"]
    assert len(statics) == sum(1 for n in mm.step.values() for mn in n.micro.values()
                               if mn.inst[0] == "LoadCursorStatic")   # only the first copy loads state


def test_empty_index_slot_is_offset_zero():
    """`a[,i+1]` leaves the first slot empty: offset 0 on that axis
    (Parser.hs:186-192), the form the reference's MHD examples use."""
    from formura_amd.compiler.driver import compile_source
    src = """
dimension :: 2
axes :: x, y
begin function init() returns q
  double [] :: q = 0
end function
begin function q_next = step(q)
  q_next = q[,i+1] - q[i-1,]
end function
"""
    nc = NumericalConfig.from_yaml_text("intra_node_shape: [8, 8]\nmpi_grid_shape: [1, 1]\nmonitor_interval: 2\n")
    c = compile_source("slot", src, nc)
    loads = sorted(op.args[1] for op in c.ir.ops if op.kind == "load")
    assert loads == [(-1, 0), (0, 1)]


def test_reference_mhd_programs_compile_to_stage_kernels(reference_tree):
    """The reference's largest examples (2D and 3D MHD: Unicode names, tuple
    lambdas, empty index slots, half-integer grids, dozens of manifest nodes)
    go through the front end, the stage lowering and the HIP emitter.  For
    examples-generated/2d-mhd.fmr the manifest nodes are as many as the Om_
    arrays of the reference's generated examples-generated/2d-mhd.c (80), in
    the same order: its node ids run behind ours by an offset that only grows
    from block to block (that C was generated by an older compiler version
    whose step graph had fewer nodes per stage)."""
    from formura_amd.compiler.driver import compile_files
    from formura_amd.compiler.emit_hip import emit_hip
    from formura_amd.compiler.pointir import group_stages, lower_stages
    gen = os.path.join(reference_tree, "examples-generated")
    c = compile_files("m2", os.path.join(gen, "2d-mhd.fmr"), os.path.join(gen, "2d-mhd.nc"),
                      nc_override={"mpi_grid_shape": [1, 1], "intra_node_shape": [1200, 1200]})
    st = lower_stages(c.mm)
    mine = sorted(s.om_id for s in st if s.target == "temp")
    ref = sorted({int(x) for x in re.findall(r"Om_(\d+)", open(os.path.join(gen, "2d-mhd.c")).read())})
    assert len(mine) == len(ref) == 80
    gaps = [a - b for a, b in zip(mine, ref)]
    assert all(0 <= g for g in gaps) and gaps == sorted(gaps)
    ex = os.path.join(reference_tree, "examples")
    for fmr, yml, shape, nodes in (("3d-mhd.fmr", "3d-mhd.yaml", [16, 16, 16], 72),
                                  ("2d-mhd.fmr", "2d-mhd.yaml", [32, 32], 99)):
        c = compile_files("m", os.path.join(ex, fmr), os.path.join(ex, yml),
                          nc_override={"mpi_grid_shape": [1] * len(shape), "intra_node_shape": shape})
        st = lower_stages(c.mm)
        assert len(st) == nodes
        groups = group_stages(st)
        assert sum(len(g.members) for g in groups) == nodes
        src = emit_hip(c.layout, c.ir, "faithful", [], st)
        assert "static constexpr bool INLINED = false;" in src
        assert src.count("&fmr::launch_stage<Prog, Stage") == len(groups)


# examples of the reference tree and what the current grammar says of them
REFERENCE_EXAMPLES_OK = [
    "2d-mhd-dc.fmr", "2d-mhd.fmr", "3d-localsun.fmr", "3d-mhd-lessM.fmr", "3d-mhd-rk4.fmr", "av-mhd.fmr",
    "champion.fmr", "diffusion.fmr", "finalist.fmr", "lotka-volterra.fmr", "pearson-2d.fmr",
    "pearson-3d-0M-alt.fmr", "pearson-3d-0M.fmr", "pearson-3d-fast.fmr", "pearson-3d-lap.fmr",
    "pearson-3d.fmr", "pearson-3dx2.fmr", "pearson.fmr", "rule30.fmr",
]
# written for an older grammar: numerical config inside the .fmr
# (`intra_node_shape :: ...`) or `for(x)` lambdas; the current keyword set
# (Parser.hs:80-83) has `fun` and no such declarations
REFERENCE_EXAMPLES_OLD_GRAMMAR = ["av-mhd-3d.fmr", "hydro.fmr", "mhd.fmr", "diffusion-long.fmr"]


def _compile_reference_example(reference_tree, name):
    from formura_amd.compiler.driver import compile_source
    ex = os.path.join(reference_tree, "examples")
    src = open(os.path.join(ex, name), encoding="utf-8").read()
    dim = int(re.search(r"dimension\s*::\s*(\d+)", src).group(1))
    y = os.path.join(ex, name[:-4] + ".yaml")
    nc = NumericalConfig.from_file(y) if os.path.exists(y) else NumericalConfig(
        intra_node_shape=[32] * dim, mpi_grid_shape=[1] * dim, monitor_interval=2)
    nc.intra_node_shape = [min(64, n) for n in nc.intra_node_shape]
    nc.mpi_grid_shape = [1] * dim
    return compile_source("ex", src, nc)


def test_reference_examples_compile(reference_tree):
    """Every example of the reference tree written in the current grammar goes
    through the front end, the inlining to a point IR and the stage lowering
    (one rank, grids capped at 64 per axis)."""
    from formura_amd.compiler.pointir import lower_stages
    for name in REFERENCE_EXAMPLES_OK:
        c = _compile_reference_example(reference_tree, name)
        assert c.ir.ops and c.layout.field_names, name
        assert lower_stages(c.mm), name


def test_reference_examples_in_the_old_grammar_are_rejected(reference_tree):
    for name in REFERENCE_EXAMPLES_OLD_GRAMMAR:
        with pytest.raises((FormuraSyntaxError, FormuraCompileError)):
            _compile_reference_example(reference_tree, name)
