"""The region plan restated from Cut.hs (formura_amd/compiler/cut_plan.py) against the
reference's own verbose plan print for examples-generated/pearson.fmr
(examples-generated/pearson.log:85-466: allocation list, buffer sharing,
ridge list and distributed program of the normal and the wall-inverted plan;
walls x [200, 400], y [300], 600 x 600 per rank).  That log predates the MPI
facets of the current Cut.hs, so the facet instructions (wait, send-receive)
are left out of the comparison; everything else must match line for line."""
from __future__ import annotations

import os
import re

import pytest

from formura_amd.compiler.manifest import gen_mm_program
from formura_amd.compiler.numerical_config import NumericalConfig
from formura_amd.compiler.omgraph import gen_om_program
from formura_amd.compiler.parser import parse_program
from formura_amd.compiler.cut_plan import Box, IRank, Rsc, RidgeID, make_plan


def vec(v):
    return "(" + ",".join(str(x) for x in v) + ")"


def box(b: Box):
    return f"「{vec(b.lo)} {vec(b.hi)}」"


def irank(ir: IRank):
    return f"IRank {vec(ir.v)}"


def rsc(r: Rsc, shift):
    if r.static:
        return f'ResourceStatic "{r.name}" ()'
    nid = r.name - shift
    if r.extra is None:
        return f"ResourceOMNode {nid} ()"
    if isinstance(r.extra, IRank):
        return f"ResourceOMNode {nid} ({irank(r.extra)})"
    return f"ResourceOMNode {nid} ({irank(r.extra[0])},{irank(r.extra[1])})"


def ridge(r: RidgeID, shift):
    return f"RidgeID {{_ridgeDeltaMPI = MPIRank {vec(r.dmpi)}, _ridgeDelta = {rsc(r.delta, shift)}}}"


def plan_lines(p, shift):
    out = ["#### Allocation List ####"]
    for k in sorted(p.array_alloc):
        out += [rsc(k, shift), "  " + box(p.array_alloc[k])]
    out.append("#### Resource Sharing ####")
    groups = {}
    for k in sorted(p.sharing):
        groups.setdefault(p.sharing[k], []).append(k)
    for i in sorted(groups):
        out.append(str(i))
        out += ["  " + rsc(k, shift) for k in groups[i]]
    out.append("#### Ridge List ####")
    for r in sorted(p.ridge_alloc):
        out += [ridge(r, shift), "  " + box(p.ridge_alloc[r])]
    out.append("#### Program ####")
    for inst in p.program:
        if inst[0] == "stage":
            out.append(f"Stage ({ridge(inst[1], shift)})")
        elif inst[0] == "unstage":
            out.append(f"Unstage ({ridge(inst[1], shift)})")
        elif inst[0] == "compute":
            ir, nid = inst[1]
            out.append(f"Computation ({irank(ir)},{nid - shift}) ({rsc(inst[2], shift)})")
        elif inst[0] == "free":
            out.append(f"FreeResource ({rsc(inst[1], shift)})")
    return out


@pytest.fixture(scope="module")
def pearson_plans(reference_tree):
    d = os.path.join(reference_tree, "examples-generated")
    src = open(os.path.join(d, "pearson.fmr"), encoding="utf-8").read()
    nc = NumericalConfig.from_file(os.path.join(d, "pearson.nc"))
    nc.temporal_blocking_interval = 1      # the log's step graph is one step
    log = open(os.path.join(d, "pearson.log"), encoding="utf-8").read()
    mm = gen_mm_program(gen_om_program(parse_program(src, nc)))
    shift = len(nc.intra_node_shape)        # NX, NY literals injected by the current parser
    plans = [make_plan(mm.step, 2, ["x", "y"], nc.intra_node_shape, nc.initial_walls, mm.state_order, inv)
             for inv in (False, True)]
    return plans, shift, log


def log_plans(log):
    lines = log.splitlines()
    starts = [i for i, l in enumerate(lines) if l == "#### Allocation List ####"]
    out = []
    for a, b in zip(starts, starts[1:] + [len(lines)]):
        sec = [l for l in lines[a:b] if l.strip() and not l.startswith("##") or l.startswith("####")]
        out.append([l for l in sec if not l.startswith("## ")])
    return out


def test_plans_match_the_reference_plan_print(pearson_plans):
    plans, shift, log = pearson_plans
    want = log_plans(log)
    assert len(want) == 2
    for p, w in zip(plans, want):
        got = plan_lines(p, shift)
        assert len(got) == len(w) == 191
        assert got == w, next(f"line {i}: got {g!r} want {x!r}" for i, (g, x) in enumerate(zip(got, w)) if g != x)


def test_system_offsets_and_regions(pearson_plans):
    plans, shift, log = pearson_plans
    p0, p1 = plans
    # the store nodes' walls: x [1, 199, 399, 601] and y [1, 299, 601] in the
    # normal plan, mirrored in the inverted one
    assert p0.system_offset == (1, 1) and p1.system_offset == (-1, -1)
    assert [ir.v for ir in p0.iranks] == [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]
    assert [ir.v for ir in p1.iranks] == list(reversed([ir.v for ir in p0.iranks]))


# --- region-plan libraries against the plain oracle ------------------------

def _states(name, shape):
    from tests.common import gray_scott_ic, normal
    if name.startswith("heat"):
        return {"u": normal(31, shape)}
    u, v = gray_scott_ic(shape, 32)
    return {"U": u, "V": v}


@pytest.mark.parametrize("name", ["gs3d_plan", "heat_plan", "gs2d_plan", "heat_rk2_plan", "gs3d_rk4_plan",
                                  "gs3d_nbu_plan", "heat_nbu_plan"])
def test_region_plan_library_equals_oracle_bitwise(name, tmp_path):
    """The C generated on the restated region plan (regions, skewed in-place
    stores, ridges and self-message facets, normal and inverted plans, the
    T-fold step graph, buffer sharing) computes the same bits as the plain
    oracle of the same program."""
    import subprocess

    import numpy as np

    from formura_amd import build as product_build
    from formura_amd.runtime import StepLibrary
    from oracle.build_oracle import CFLAGS, planned_layout, planned_path
    from oracle.emit_cxx import write_oracle

    if not os.path.exists(planned_path(name)):
        pytest.skip("region-plan library not built (oracle/build_oracle.py)")
    c, lay = planned_layout(name)
    a = StepLibrary(planned_path(name), lay, device=False)
    a.init()
    d = str(tmp_path)
    src = write_oracle(d, c.layout, c.mm)
    so = os.path.join(d, "liboracle.so")
    subprocess.check_call(["gcc"] + CFLAGS + ["-O1", "-I", d, "-I", product_build.MPI_STUB, src, "-o", so, "-lm"])
    b = StepLibrary(so, c.layout, device=False)
    b.init()
    st = _states(name, tuple(lay.intra))
    a.set_state(st)
    b.set_state(st)
    for _ in range(2):
        a.forward()
        b.forward()
    assert a.time_step == b.time_step == 2 * lay.steps_per_forward
    ga, gb = a.get_state(), b.get_state()
    for f in ga:
        assert np.array_equal(ga[f], gb[f]), (name, f, float(np.abs(ga[f] - gb[f]).max()))


def test_walls_closer_than_the_reach_are_rejected():
    """Regions narrower than the reach of the T-fold step would make a region
    wait for a facet its own rank sends later (read stale in the reference's
    generated code); the planner refuses such walls."""
    from formura_amd.compiler.driver import compile_source
    from oracle.emit_planned import _Emitter

    src = open(os.path.join(os.path.dirname(__file__), "..", "examples", "heat-3d-rk2.fmr"), encoding="utf-8").read()
    walls = {"x": [6], "y": [8], "z": [10]}
    nc = NumericalConfig(intra_node_shape=[12, 16, 20], mpi_grid_shape=[1, 1, 1], temporal_blocking_interval=2,
                         monitor_interval=4, initial_walls=walls)
    c = compile_source("narrow", src, nc)
    with pytest.raises(ValueError, match="narrower than the stencil reach"):
        _Emitter(c.layout, c.mm, 2, walls, ["omp"])


def test_nanoblock_unroll_loops_and_margins():
    """Nanoblock unroll (Manifestation.hs:146-158, MPICxx/Translate.hs:698-719,
    :942): with nbux2 every loop nest steps x by 2 and stores two sites per
    iteration, an operation that both sites use is computed once, and the
    state arrays grow by nbu - 1 on the upper side of x only.  A region that
    is not a whole number of nanoblocks is refused."""
    from formura_amd.compiler.driver import compile_source
    from oracle.emit_planned import _Emitter

    src = open(os.path.join(os.path.dirname(__file__), "..", "examples", "heat-3d-512.fmr"), encoding="utf-8").read()
    walls = {"x": [10], "y": [12], "z": [14]}

    def emitter(opts, shape=(20, 24, 28)):
        nc = NumericalConfig(intra_node_shape=list(shape), mpi_grid_shape=[1, 1, 1], temporal_blocking_interval=2,
                             monitor_interval=4, initial_walls=walls, option_strings=opts)
        c = compile_source("nb", src, nc)
        return _Emitter(c.layout, c.mm, 2, walls, opts)

    plain = emitter(["omp"])
    nb = emitter(["omp", "nbux2"])
    assert [u - p for u, p in zip(nb.static_box.hi, plain.static_box.hi)] == [1, 0, 0]
    assert nb.static_box.lo == plain.static_box.lo
    p = plain.plans[0]
    inst = next(i for i in p.program if i[0] == "compute")
    a, b = plain.computation(p, *inst[1], inst[2]), nb.computation(nb.plans[0], *inst[1], inst[2])
    assert "ix += 2)" in "\n".join(b) and "ix += 1)" in "\n".join(a)
    stores = lambda code: [ln for ln in code if " = a_" in ln and not ln.startswith(("double", "float"))]
    ops = lambda code: [ln for ln in code if ln.startswith("double a_")]
    assert len(stores(b)) == 2 * len(stores(a)) == 2
    assert len(ops(a)) < len(ops(b)) < 2 * len(ops(a))
    with pytest.raises(ValueError, match="whole number of nanoblocks"):
        e = emitter(["omp", "nbux3"])
        q = e.plans[0]
        for i in q.program:
            if i[0] == "compute":
                e.computation(q, *i[1], i[2])
