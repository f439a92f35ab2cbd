"""Stage kernels that evaluate several manifest nodes at once.

`group_stages` puts manifest nodes that do not read each other's arrays into
one kernel and shares their equal loads and operations.  Both the per-node
sequence and the grouped kernels are evaluated here with numpy over a periodic
grid (IEEE float64, the same operations in the same order) and compared bit
for bit with each other and with the oracle's step."""
from __future__ import annotations

import numpy as np

from formura_amd.compiler.driver import compile_source
from formura_amd.compiler.numerical_config import NumericalConfig
from formura_amd.compiler.pointir import group_stages, lower_stages
from tests.common import normal
from tests.test_oracle import EX, build_oracle_variant

BIN = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide, "<": np.less, "<=": np.less_equal,
       ">": np.greater, ">=": np.greater_equal, "==": np.equal, "!=": np.not_equal}


def evaluate(ir, arrays, shape):
    """Values of the outputs of a point IR at every site of the grid."""
    vals = []
    for op in ir.ops:
        k, a = op.kind, op.args
        if k == "load":
            v = arrays[a[0]]
            for ax, d in enumerate(a[1]):
                if d:
                    v = np.roll(v, -d, axis=ax)
            vals.append(v)
        elif k == "imm":
            vals.append(np.float64(float(a[0])))
        elif k == "index":
            vals.append(np.indices(shape)[a[0]].astype(np.float64))
        elif k == "uniop":
            vals.append(-vals[a[1]] if a[0] == "-" else +vals[a[1]])
        elif k == "binop":
            vals.append(BIN[a[0]](vals[a[1]], vals[a[2]]))
        elif k == "triop":
            vals.append(np.where(vals[a[1]], vals[a[2]], vals[a[3]]))
        elif k == "naryop":
            xs = [vals[x] for x in a[1]]
            hi = xs[-1]
            lo = xs[-1]
            for x in reversed(xs[:-1]):
                hi, lo = np.fmax(x, hi), np.fmin(x, lo)
            vals.append({">?": hi, "<?": lo}.get(a[0], None) if a[0] != "<%" else
                        np.fmin(0.0, hi) + np.fmax(0.0, lo))
        else:
            raise NotImplementedError(k)
    return [np.broadcast_to(vals[o], shape) for o in ir.outputs]


def run_step(units, nf, ntemps, state, shape, members_of, ir_of):
    arrays = list(state) + [np.zeros(shape) for _ in range(ntemps)]
    new = [None] * nf
    for u in units:
        outs = evaluate(ir_of(u), arrays, shape)
        for st, v in zip(members_of(u), outs):
            if st.target == "temp":
                arrays[st.index] = np.array(v)
            else:
                new[st.index] = np.array(v)
    return new


def compiled(fmr, shape, monitor=1):
    text = open(f"{EX}/{fmr}", encoding="utf-8").read()
    nc = NumericalConfig(intra_node_shape=list(shape), mpi_grid_shape=[1] * len(shape), monitor_interval=monitor)
    return compile_source("grp", text, nc)


def check(fmr, shape, state, tmp_path, expect_kernels):
    c = compiled(fmr, shape)
    stages = lower_stages(c.mm)
    groups = group_stages(stages)
    nf = len(c.layout.field_names)
    ntemps = sum(1 for s in stages if s.target == "temp")
    assert sum(len(g.members) for g in groups) == len(stages)
    assert len(groups) == expect_kernels
    lib, _ = build_oracle_variant(tmp_path, "grp", fmr, shape, 1)
    lib.set_state(state)
    lib.forward()
    want = lib.get_state()
    seq = grp = [state[f] for f in c.layout.field_names]
    for _ in range(lib.layout.steps_per_forward):
        seq = run_step(stages, nf, ntemps, seq, shape, lambda s: [s], lambda s: s.ir)
        grp = run_step(groups, nf, ntemps, grp, shape, lambda g: g.members, lambda g: g.ir)
    for i, f in enumerate(c.layout.field_names):
        assert np.array_equal(seq[i], want[f]), f
        assert np.array_equal(grp[i], want[f]), f
    return stages, groups


def test_fv_euler_groups_match_per_node_stages_and_oracle(tmp_path):
    from tests.test_oracle import fv_state
    shape = (10, 12, 14)
    # per SSP-RK3 stage: the face fluxes of four fields along three axes, then the state
    stages, groups = check("fv-euler-3d.fmr", shape, fv_state(shape, 1), tmp_path, 6)
    assert len(stages) == 48
    assert [len(g.members) for g in groups] == [12, 4, 12, 4, 12, 4]
    # the twelve face fluxes of one stage read the four fields' neighbours once
    g0 = groups[0]
    loads = [op for op in g0.ir.ops if op.kind == "load"]
    assert len(loads) == len({(op.args[0], op.args[1]) for op in loads}) == 40
    assert 2 * len(loads) < sum(1 for s in g0.members for op in s.ir.ops if op.kind == "load")


def test_rk4_gray_scott_groups(tmp_path):
    from tests.common import gray_scott_ic
    shape = (8, 10, 12)
    u, v = gray_scott_ic(shape, 3)
    check("gray-scott-3d-rk4.fmr", shape, {"U": u, "V": v}, tmp_path, 4)


def test_group_limits_split_a_level():
    c = compiled("fv-euler-3d.fmr", (8, 8, 8))
    stages = lower_stages(c.mm)
    groups = group_stages(stages, max_members=5)
    assert all(len(g.members) <= 5 for g in groups)
    assert [len(g.members) for g in groups] == [5, 5, 2, 4] * 3
    small = group_stages(stages, max_loads=12)
    for g in small:
        assert len(g.members) == 1 or sum(1 for op in g.ir.ops if op.kind == "load") <= 12
