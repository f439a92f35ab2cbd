"""Batched launches (many 2D Gray-Scott instances of uneven size and
parameters): FP64 batch against the CPU oracle bit for bit, every compiled
variant; hipGraph replays against eager launches bit for bit; the FP32 bulk
batch against the FP64 batch within the stated 1e-3 relative L2 bound."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from formura_amd.batched import InstanceSpec, initial_state, open_batch, relative_l2, size_buckets, sweep_instances
from formura_amd.runtime import StepLibrary

pytestmark = pytest.mark.gpu

def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def batch(name, specs):
    lib = open_batch(name)
    if lib.specs != list(specs):
        lib.init(specs)
    return lib


# small instances: 256^2 with the .fmr constants, two odd shapes with random
# parameters; each has a CPU oracle built with the same parameters
SMALL = [InstanceSpec(256, 256, 0.04, 0.06), InstanceSpec(96, 160, 0.023, 0.051),
         InstanceSpec(64, 32, 0.047, 0.064)]
ORACLES = ["gs2d_p256", "gs2d_p96x160", "gs2d_p64x32"]


def oracle_after(i, u, v, forwards):
    from oracle.build_oracle import compiled_extra, oracle_path
    name = ORACLES[i]
    c = compiled_extra(name)
    lib = StepLibrary(oracle_path(name), c.layout, device=False)
    lib.init()
    import ctypes
    p = (ctypes.c_double * 5)(*SMALL[i].params())
    assert lib.lib.Formura_Set_Params(p, 5) == 0
    lib.set_state({"U": u, "V": v})
    for _ in range(forwards):
        lib.forward()
    s = lib.get_state()
    return s["U"], s["V"], c.layout.steps_per_forward * forwards


def test_fp64_batch_every_variant_matches_oracle_bitwise():
    lib = batch("gs2d_batch_f64", SMALL)
    ics = [initial_state(s, 100 + i) for i, s in enumerate(SMALL)]
    want = [oracle_after(i, u, v, 1) for i, (u, v) in enumerate(ics)]
    steps = want[0][2]
    buckets = [0, 1, 1]
    for T, vy in lib.variants():
        for lx in (64, 7):
            _say(f"[batched] FP64 T={T} vy={vy} lx={lx} steps={steps}")
            lib.plan(buckets, [(T, vy, lx), (T, vy, lx)])
            for i, (u, v) in enumerate(ics):
                lib.set_state(i, [u, v])
            lib.advance(steps, graph=False)
            for i in range(len(SMALL)):
                gu, gv = lib.get_state(i)
                assert np.array_equal(gu, want[i][0]) and np.array_equal(gv, want[i][1]), \
                    f"instance {i} differs for T={T} vy={vy} lx={lx}"


def test_fused_launch_mixed_lane_widths_matches_oracle_bitwise():
    """A plan of one depth T with a different lane width and march chunk per
    bucket runs as one persistent launch per T-block over the tiles of every
    instance (Formura_Batch_Fused); the captured graph holds one kernel node
    per T-block; the FP64 result equals the oracle bit for bit."""
    lib = batch("gs2d_batch_f64", SMALL)
    ics = [initial_state(s, 200 + i) for i, s in enumerate(SMALL)]
    want = [oracle_after(i, u, v, 1) for i, (u, v) in enumerate(ics)]
    steps = want[0][2]
    widths = {}
    for T, vy in lib.variants():
        widths.setdefault(T, set()).add(vy)
    for T in (1, 2, 4, 8):
        if T not in widths or len(widths[T]) < 2 or steps % T:
            continue
        a, b = sorted(widths[T])[:2]
        lib.plan([0, 1, 1], [(T, a, 16), (T, b, 64)])
        assert lib.fused() == (T, (1 if a == 1 else 2 if a == 2 else 4) | (1 if b == 1 else 2 if b == 2 else 4))
        for graph in (False, True):
            _say(f"[batched] fused T={T} widths {a},{b} graph={graph} steps={steps}")
            for i, (u, v) in enumerate(ics):
                lib.set_state(i, [u, v])
            lib.advance(steps, graph=graph)
            if graph:
                assert lib.last_graph_nodes() == steps // T
            for i in range(len(SMALL)):
                gu, gv = lib.get_state(i)
                assert np.array_equal(gu, want[i][0]) and np.array_equal(gv, want[i][1]), \
                    f"instance {i} differs for fused T={T} widths {a},{b} graph={graph}"


def test_graph_replays_match_eager_launches():
    lib = batch("gs2d_batch_f64", SMALL)
    ics = [initial_state(s, 7 + i) for i, s in enumerate(SMALL)]
    vs = lib.variants()
    T, vy = max(v for v in vs if v[1] == 2)
    lib.plan([0, 1, 2], [(T, vy, 32), (1, vy, 16), (2, vy, 64)])
    for i, (u, v) in enumerate(ics):
        lib.set_state(i, [u, v])
    lib.advance(10, graph=False)
    lib.advance(10, graph=False)
    eager = [lib.get_state(i) for i in range(3)]
    for rep in range(3):
        _say(f"[batched] graph replay round {rep}")
        for i, (u, v) in enumerate(ics):
            lib.set_state(i, [u, v])
        lib.advance(10, graph=True)     # captured on first use, replayed after
        lib.get_state(0)                # eager work between replays
        lib.advance(10, graph=True)
        for i in range(3):
            g = lib.get_state(i)
            assert np.array_equal(g[0], eager[i][0]) and np.array_equal(g[1], eager[i][1])


def test_plain_method_matches_oracle_bitwise():
    lib = batch("gs2d_batch_f64", SMALL)
    ics = [initial_state(s, 300 + i) for i, s in enumerate(SMALL)]
    want = [oracle_after(i, u, v, 1) for i, (u, v) in enumerate(ics)]
    for i, (u, v) in enumerate(ics):
        lib.set_state(i, [u, v])
    lib.advance_plain(want[0][2])
    for i in range(3):
        gu, gv = lib.get_state(i)
        assert np.array_equal(gu, want[i][0]) and np.array_equal(gv, want[i][1])


def test_fp32_sweep_within_1e3_of_fp64():
    """The full sweep (256 instances, 5000 steps): FP32 bulk batch against the
    FP64 batch (itself bit-equal to the oracle above), per instance and field."""
    specs = sweep_instances()
    f64 = batch("gs2d_batch_f64", specs)
    f32 = batch("gs2d_batch_f32", specs)
    b = size_buckets(specs)
    nb = max(b) + 1
    f64.plan(b, [(2, 2, 32)] * nb)
    f32.plan(b, [(4, 2, 32)] * nb)
    for i, s in enumerate(specs):
        u, v = initial_state(s, 1000 + i)
        f64.set_state(i, [u, v])
        f32.set_state(i, [u, v])
    _say("[batched] sweep: 256 instances x 5000 steps, FP64 and FP32")
    f64.advance(5000, graph=True)
    f32.advance(5000, graph=True)
    worst = 0.0
    for i, s in enumerate(specs):
        for f in range(2):
            # relative L2; a field that decayed below RMS 1e-6 is measured on
            # the absolute scale of an O(1) field (see relative_l2)
            worst = max(worst, relative_l2(f32.diff(i, f, f64), s.nx * s.ny))
    _say(f"[batched] worst relative L2 error FP32 vs FP64 over 256 instances x 2 fields: {worst:.3e}")
    assert worst <= 1e-3
