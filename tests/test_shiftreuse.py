"""The shift analysis behind the lane-axis reuse of the march kernel.

Its claim -- op B at site s has the bits of op A at site s - e -- is checked by
evaluating both ops over a small periodic grid with numpy (IEEE float64, the
same operations) and comparing B with A rolled by one site, bit for bit.  The
pairs for the Laplacian are pinned by hand: B = c - c[k-1] reuses
A = c[k+1] - c (the opposite direction would fail the roll check), and an op
that reads the site's coordinates is never paired."""
from __future__ import annotations

import numpy as np

from formura_amd import build
from formura_amd.compiler.driver import compile_files, compile_source
from formura_amd.compiler.numerical_config import NumericalConfig
from formura_amd.compiler.pointir import c_expr_of
from formura_amd.compiler.shiftreuse import ShiftAnalysis, op_children


def program(name):
    fmr, yaml_, precision, params, over, _ = build.PROGRAMS[name]
    return compile_files(name, fmr, yaml_, precision, params, over)


def evaluate(ir, fields, axis_sizes):
    """Every op of the point IR at every site of a periodic grid (numpy)."""
    vals = []
    for op in ir.ops:
        k = op.kind
        if k == "load":
            f, off = op.args
            a = fields[f]
            for ax, d in enumerate(off):
                if d:
                    a = np.roll(a, -d, axis=ax)        # a[s] = field[s + d]
            vals.append(a)
        elif k == "imm":
            vals.append(np.float64(float(op.args[0])))
        elif k == "index":
            vals.append(np.indices(axis_sizes)[op.args[0]].astype(np.float64))
        elif k == "uniop":
            o, a = op.args
            vals.append(-vals[a] if o == "-" else +vals[a])
        elif k == "binop":
            o, a, b = op.args
            x, y = vals[a], vals[b]
            vals.append({"+": x + y, "-": x - y, "*": x * y, "/": x / y}[o])
        else:
            raise NotImplementedError(k)
    return vals


def test_laplacian_pairs_are_pinned():
    c = program("heat512")
    ir = c.ir
    name = lambda j: f"a{j}"
    for axis in range(3):
        pairs = ShiftAnalysis(ir).pairs(axis)
        assert len(pairs) == 1
        (b, a), = pairs.items()
        lb = [ir.ops[j].args for j in op_children(ir.ops[b]) if ir.ops[j].kind == "load"]
        la = [ir.ops[j].args for j in op_children(ir.ops[a]) if ir.ops[j].kind == "load"]
        # B = c - c[-1] along the axis, A = c[+1] - c
        assert c_expr_of(ir.ops[b], name).count("-") == 1 and ir.ops[b].args[0] == "-"
        assert sorted(o[1][axis] for o in lb) == [-1, 0]
        assert sorted(o[1][axis] for o in la) == [0, 1]


def test_pairs_hold_bitwise_on_a_periodic_grid():
    rng = np.random.default_rng(5)
    for name in ("heat512", "gs3d512"):
        ir = program(name).ir
        shape = (6, 7, 8)
        fields = [rng.standard_normal(shape) for _ in ir.fields]
        vals = evaluate(ir, fields, shape)
        for axis in range(3):
            pairs = ShiftAnalysis(ir).pairs(axis)
            assert pairs, (name, axis)
            for b, a in pairs.items():
                moved = np.roll(np.broadcast_to(vals[a], shape), 1, axis=axis)   # A at s - e
                assert np.array_equal(moved.view(np.uint64), np.broadcast_to(vals[b], shape).view(np.uint64))


def test_index_dependent_ops_are_not_paired():
    src = """
dimension :: 1
axes :: x
begin function init() returns q
  double [] :: q = 0
end function
begin function q_next = step(q)
  q_next[i] = q[i] + (q[i+1] - q[i]) * i - (q[i] - q[i-1]) * i
end function
"""
    nc = NumericalConfig.from_yaml_text("intra_node_shape: [16]\nmpi_grid_shape: [1]\nmonitor_interval: 2\n")
    c = compile_source("idx", src, nc)
    sa = ShiftAnalysis(c.ir)
    pairs = sa.pairs(0)
    for b, a in pairs.items():
        assert sa.op_id[b] is not None and sa.op_id[a] is not None
    # the differences themselves pair; the products with the coordinate do not
    exprs = {c_expr_of(c.ir.ops[b], lambda j: f"a{j}") for b in pairs}
    assert all("*" not in e for e in exprs) and len(pairs) == 1
