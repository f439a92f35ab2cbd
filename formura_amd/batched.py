"""Batched launches of a 2D program: Python view of the batch C ABI
(include/formura_amd_batch.h), the instance generator of the batched
Gray-Scott parameter sweep, and the autotuner that chooses the fusion depth T,
the lane width and the march chunk per size bucket.

Sweep (one MI355X): 256 independent periodic 2D Gray-Scott instances, side
log-uniform in [128, 1024] rounded to a multiple of 32, F ~ U[0.01, 0.06],
k ~ U[0.045, 0.07], Du = 2e-5, Dv = 1e-5, dx = 1/side, dt = 0.2 dx^2/Du;
u = 1, v = 0 except a centred square of side/8 with u = 0.5, v = 0.25, plus
uniform noise of +-0.01 on u and v.
"""
from __future__ import annotations

import ctypes
import math
import os
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import build as _build
from .compiler.emit_batched import FUSED_T

PARAMS = ("F", "k", "dt", "cu", "cv")   # named scalars of examples/gray-scott-2d.fmr
DU, DV = 2e-5, 1e-5


@dataclass
class InstanceSpec:
    nx: int
    ny: int
    F: float
    k: float

    def params(self) -> List[float]:
        """(F, k, dt, cu, cv) with the arithmetic of the .fmr program:
        dt = ((0.2*dx)*dx)/Du, cu = Du/(dx*dx), cv = Dv/(dx*dx), dx = 1/side."""
        dx = 1.0 / self.nx
        return [self.F, self.k, ((0.2 * dx) * dx) / DU, DU / (dx * dx), DV / (dx * dx)]


def sweep_instances(n: int = 256, seed: int = 2026, lo: int = 128, hi: int = 1024) -> List[InstanceSpec]:
    """The batched parameter sweep, from a SplitMix64 stream (portable)."""
    from .rng import uniform
    u = uniform(seed, (n, 3))
    out = []
    for i in range(n):
        side = math.exp(math.log(lo) + u[i, 0] * (math.log(hi) - math.log(lo)))
        side = int(min(hi, max(lo, 32 * round(side / 32))))
        F = 0.01 + 0.05 * u[i, 1]
        k = 0.045 + 0.025 * u[i, 2]
        out.append(InstanceSpec(side, side, F, k))
    return out


def initial_state(spec: InstanceSpec, seed: int):
    from .rng import gray_scott_ic
    return gray_scott_ic((spec.nx, spec.ny), seed)


_LOADED: Dict[str, "BatchLibrary"] = {}


def open_batch(name: str) -> "BatchLibrary":
    """The process-wide handle of a batch library (its state is global)."""
    if name not in _LOADED:
        _LOADED[name] = BatchLibrary(name)
    return _LOADED[name]


class BatchLibrary:
    def __init__(self, name: str):
        path = _build.lib_path(name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"batched library {path} is not built (no CPU fallback exists)")
        self.name = name
        L = self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.Formura_Batch_Init.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int]
        L.Formura_Batch_List_Variants.argtypes = [ip, ctypes.c_int]
        L.Formura_Batch_Plan.argtypes = [ctypes.c_int, ip, ip]
        L.Formura_Batch_Set_State.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.Formura_Batch_Get_State.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.Formura_Batch_Advance.argtypes = [ctypes.c_int, ctypes.c_int]
        L.Formura_Batch_Advance.restype = ctypes.c_float
        L.Formura_Batch_Advance_Bucket.argtypes = [ctypes.c_int, ctypes.c_int]
        L.Formura_Batch_Advance_Bucket.restype = ctypes.c_float
        L.Formura_Batch_Advance_Plain.argtypes = [ctypes.c_int]
        L.Formura_Batch_Advance_Plain.restype = ctypes.c_float
        L.Formura_Batch_Device_Field.argtypes = [ctypes.c_int, ctypes.c_int]
        L.Formura_Batch_Device_Field.restype = ctypes.c_void_p
        L.Formura_Batch_Diff.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, dp]
        L.Formura_Batch_Fused.argtypes = [ip, ip]
        L.Formura_Batch_Last_Graph_Nodes.restype = ctypes.c_int
        self.nf = L.Formura_Batch_Num_Fields()
        self.np = L.Formura_Batch_Num_Params()
        self.elem = L.Formura_Batch_Elem_Size()
        self.specs: List[InstanceSpec] = []

    def init(self, specs: Sequence[InstanceSpec]) -> None:
        """Allocate the instances (releasing any earlier set)."""
        if self.specs:
            self.lib.Formura_Batch_Destroy()
        self.specs = list(specs)
        n = len(specs)
        nx = (ctypes.c_int * n)(*[s.nx for s in specs])
        ny = (ctypes.c_int * n)(*[s.ny for s in specs])
        prm = (ctypes.c_double * (n * self.np))(*[p for s in specs for p in s.params()])
        if self.lib.Formura_Batch_Init(n, nx, ny, prm, self.np) != 0:
            raise RuntimeError("Formura_Batch_Init failed")

    def variants(self) -> List[Tuple[int, int]]:
        buf = (ctypes.c_int * 256)()
        k = self.lib.Formura_Batch_List_Variants(buf, 128)
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(k)]

    def plan(self, bucket_of: Sequence[int], variants: Sequence[Tuple[int, int, int]]) -> None:
        n = len(self.specs)
        b = (ctypes.c_int * n)(*bucket_of)
        v = (ctypes.c_int * (3 * len(variants)))(*[x for t in variants for x in t])
        if self.lib.Formura_Batch_Plan(len(variants), b, v) != 0:
            raise ValueError(f"bad batch plan {variants}")

    def set_state(self, i: int, fields: Sequence[np.ndarray]) -> None:
        for f, a in enumerate(fields):
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert a.shape == (self.specs[i].nx, self.specs[i].ny)
            self.lib.Formura_Batch_Set_State(i, f, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def get_state(self, i: int) -> List[np.ndarray]:
        out = []
        for f in range(self.nf):
            a = np.empty((self.specs[i].nx, self.specs[i].ny))
            self.lib.Formura_Batch_Get_State(i, f, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            out.append(a)
        return out

    def advance(self, steps: int, graph: bool = True) -> float:
        return float(self.lib.Formura_Batch_Advance(int(steps), 1 if graph else 0))

    def advance_bucket(self, b: int, steps: int) -> float:
        """Advance only the instances of bucket b (no graph); GPU ms."""
        ms = float(self.lib.Formura_Batch_Advance_Bucket(int(b), int(steps)))
        if ms < 0:
            raise ValueError("bad bucket")
        return ms

    def advance_plain(self, steps: int) -> float:
        return float(self.lib.Formura_Batch_Advance_Plain(int(steps)))

    def fused(self) -> Optional[Tuple[int, int]]:
        """(T, lane-width mask) of the fused plan -- every instance advanced
        by one launch per T-block -- or None (per-bucket launches)."""
        T, m = ctypes.c_int(), ctypes.c_int()
        return (T.value, m.value) if self.lib.Formura_Batch_Fused(ctypes.byref(T), ctypes.byref(m)) == 1 else None

    def last_graph_nodes(self) -> int:
        """Nodes of the hipGraph captured last (advance with graph=True)."""
        return int(self.lib.Formura_Batch_Last_Graph_Nodes())

    def diff(self, i: int, f: int, ref: "BatchLibrary") -> Tuple[float, float, float]:
        r = (ctypes.c_double * 3)()
        ptr = ref.lib.Formura_Batch_Device_Field(i, f)
        if self.lib.Formura_Batch_Diff(i, f, ctypes.c_void_p(ptr), ref.elem, r) != 0:
            raise ValueError("bad diff")
        return math.sqrt(r[0]), math.sqrt(r[1]), r[2]


def relative_l2(diff_result, n: int, floor: float = 1e-6) -> float:
    """||a - b|| / max(||b||, floor * sqrt(n)) from a device diff (d, r, max).

    A field whose RMS has decayed below `floor` (a Gray-Scott pattern that died
    out: v -> 0, whose FP64 squares underflow) is compared on the absolute
    scale of an O(1) field instead of its own vanishing norm."""
    d, r, _ = diff_result
    return d / max(r, floor * math.sqrt(n))


def size_buckets(specs: Sequence[InstanceSpec], edges=(256, 512)) -> List[int]:
    """Bucket of each instance by side length: [128,256), [256,512), [512,1024]."""
    return [sum(1 for e in edges if s.ny >= e) for s in specs]


def autotune(lib: BatchLibrary, bucket_of: Sequence[int], probe_steps: int = 24,
             chunks=(16, 32, 64), depths=FUSED_T) -> Tuple[List[Tuple[int, int, int]], Dict]:
    """Choose one fusion depth T for all size buckets and, per bucket, the
    lane width and march chunk, by timing every compiled (T, lane width) with
    every chunk on that bucket's instances alone: for each T of `depths`,
    each bucket takes its fastest (lane width, chunk); the T whose buckets
    take the least time in all wins.  One depth over all buckets lets one
    fused kernel advance every instance per T-block; at most two lane widths
    are used (the fused kernels hold one or two), the two that save the most.
    Timing advances the state, so callers reload it afterwards.  Returns the
    plan and the table of rates (GLUP/s)."""
    nb = max(bucket_of) + 1
    vs = [v for v in lib.variants() if v[0] in depths]
    table: Dict[int, List] = {}
    plan = [(vs[0][0], vs[0][1], 64)] * nb
    cells = [sum(s.nx * s.ny for s, bb in zip(lib.specs, bucket_of) if bb == b) for b in range(nb)]
    rate = {}
    for b in range(nb):
        if cells[b] == 0:
            continue
        for T, vy in vs:
            for lx in chunks:
                trial = list(plan)
                trial[b] = (T, vy, lx)
                lib.plan(bucket_of, trial)
                steps = max(T, (probe_steps // T) * T)
                lib.advance_bucket(b, steps)
                ms = lib.advance_bucket(b, steps)
                r = steps * cells[b] / (ms * 1e-3) / 1e9
                rate[(b, T, vy, lx)] = r
                table.setdefault(b, []).append(((T, vy, lx), round(r, 1)))
    best = None
    for T in sorted({t for t, _ in vs}):
        widths = sorted({vy for t, vy in vs if t == T})
        for pair in {(a, c) for a in widths for c in widths if a <= c}:
            cand, total = [], 0.0
            for b in range(nb):
                if cells[b] == 0:
                    cand.append((T, pair[0], 64))
                    continue
                opts = [(rate[(b, T, vy, lx)], (T, vy, lx)) for vy in set(pair) for lx in chunks
                        if (b, T, vy, lx) in rate]
                r, choice = max(opts)
                cand.append(choice)
                total += cells[b] / r           # time per step of this bucket, up to a constant
            if best is None or total < best[0]:
                best = (total, cand)
    plan = best[1]
    lib.plan(bucket_of, plan)
    return plan, table


def sweep_benchmark(steps: int = 5000, plain_steps: int = 200, seed: int = 1000, log=print) -> Dict:
    """The batched sweep end to end on one GPU.

    FP32 bulk batch (autotuned plan, hipGraph) and its FP64 twin advance all
    256 instances by `steps`; the FP64 run is the accuracy reference (bit-equal
    to the CPU oracle per instance, see the GPU tests). The plain method --
    one FP64 single-site-per-thread launch per instance per step -- is timed
    over `plain_steps` steps and scaled linearly to `steps` (each step of the
    plain method is the same set of launches)."""
    specs = sweep_instances()
    buckets = size_buckets(specs)
    cells = sum(s.nx * s.ny for s in specs)
    ics = [initial_state(s, seed + i) for i, s in enumerate(specs)]
    res: Dict = {"instances": len(specs), "cells": cells, "steps": steps}
    libs = {}
    for name in ("gs2d_batch_f32", "gs2d_batch_f64"):
        lib = open_batch(name)
        lib.init(specs)
        t0 = time.perf_counter()
        plan, table = autotune(lib, buckets)
        log(f"{name}: plan {plan} (autotune {time.perf_counter() - t0:.1f} s)")
        for i, (u, v) in enumerate(ics):
            lib.set_state(i, [u, v])
        lib.advance(8, graph=True)          # capture a small graph once (warm-up) ...
        for i, (u, v) in enumerate(ics):    # ... and restart from the initial state
            lib.set_state(i, [u, v])
        t0 = time.perf_counter()
        ms = lib.advance(steps, graph=True)
        wall = time.perf_counter() - t0
        libs[name] = lib
        glups = steps * cells / (ms * 1e-3) / 1e9
        fz = lib.fused()
        # roofline: a T-fused launch reads and writes each instance's two
        # fields once, 2 fields x 2 x element size / T bytes per update
        T = fz[0] if fz else max(t for t, _, _ in plan)
        bpu = 2 * 2 * lib.elem / T
        res[name] = {"plan": plan, "gpu_ms": round(ms, 2), "wall_s_incl_capture": round(wall, 3),
                     "glups": round(glups, 1), "fused": fz, "graph_nodes": lib.last_graph_nodes(),
                     "launches_per_T_block": 1 if fz else len(plan),
                     "roofline": {"bound": "hbm", "bytes_per_update": bpu, "achieved_GBs": round(glups * bpu, 1),
                                  "peak_GBs": 8000.0, "frac": round(glups * bpu / 8000.0, 4)}}
        log(f"{name}: {res[name]}")
    f32, f64 = libs["gs2d_batch_f32"], libs["gs2d_batch_f64"]
    worst, mean = 0.0, 0.0
    for i, s in enumerate(specs):
        for f in range(2):
            worst_i = relative_l2(f32.diff(i, f, f64), s.nx * s.ny)
            worst = max(worst, worst_i)
            mean += worst_i
    res["fp32_vs_fp64_rel_l2_max"] = worst
    res["fp32_vs_fp64_rel_l2_mean"] = mean / (2 * len(specs))
    # plain FP64 method on the same instances
    for i, (u, v) in enumerate(ics):
        f64.set_state(i, [u, v])
    f64.advance_plain(2)
    pms = f64.advance_plain(plain_steps)
    plain_ms = pms * steps / plain_steps
    res["plain_fp64"] = {"gpu_ms_measured": round(pms, 2), "steps_measured": plain_steps,
                         "gpu_ms_scaled": round(plain_ms, 1),
                         "glups": round(plain_steps * cells / (pms * 1e-3) / 1e9, 1)}
    res["speedup_batched_fp32_vs_plain_fp64"] = round(plain_ms / res["gs2d_batch_f32"]["gpu_ms"], 2)
    res["speedup_batched_fp64_vs_plain_fp64"] = round(plain_ms / res["gs2d_batch_f64"]["gpu_ms"], 2)
    return res
