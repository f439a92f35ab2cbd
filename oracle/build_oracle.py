"""ORACLE build recipe (test infrastructure only).

Emits the reference-faithful C/OpenMP step library of each program with
oracle/emit_cxx.py and compiles it with gcc -O3 -fopenmp -ffp-contract=off into
oracle/_build/lib<name>_cpu.so.  The same header as the HIP library is used, so
one driver or one Python harness runs either.

The reference's own generated C in the reference tree
(examples-generated/pearson.c) is NOT built: it needs its generated header
`pearson.h` (not in the reference) and <mpi.h> (no MPI in this image), and
writing stand-ins for either is not allowed; see the oracle notes in the design
document.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from formura_amd import build as product_build  # noqa: E402
from formura_amd.compiler.driver import compile_files, copy_extern_header  # noqa: E402
from oracle.emit_cxx import write_oracle  # noqa: E402
from oracle.emit_planned import write_planned  # noqa: E402

OUT = os.path.join(HERE, "_build")
CFLAGS = ["-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-std=gnu99", "-mcmodel=medium"]


# oracle-only programs: (registered program, numerical-config overrides,
# parameter names)
_GS2D_PARAMS = ("F", "k", "dt", "cu", "cv")
ORACLE_ONLY = {
    # the 512^3 heat program with 10 steps per Formura_Forward: the CPU-baseline sample
    "heat512_cpu": ("heat512", {"monitor_interval": 10}, ()),
    # Gray-Scott 512^3 with 12 steps per Formura_Forward (10 rounded up to a
    # multiple of 2T, T = 2): full-size GPU parity
    "gs3d512_cpu": ("gs3d512", {"monitor_interval": 12}, ()),
    # 2D Gray-Scott with per-run parameters, for the batched-launch parity tests
    "gs2d_p256": ("gs2d", {"monitor_interval": 40}, _GS2D_PARAMS),
    "gs2d_p96x160": ("gs2d", {"intra_node_shape": [96, 160], "monitor_interval": 40}, _GS2D_PARAMS),
    "gs2d_p64x32": ("gs2d", {"intra_node_shape": [64, 32], "monitor_interval": 40}, _GS2D_PARAMS),
}


# region-plan libraries (the reference's loop and region structure, Cut.hs):
# (registered program, numerical-config overrides, temporal blocking T).
# T >= 2: the plan stores the state in place at a skewed position, which is
# safe when the state updates read the state only through manifest arrays --
# always the case for the T-fold graph with T >= 2.
PLANNED = {
    # the CPU-baseline sample: heat 512^3, 2 x 2 x 2 regions, T = 2, 10 steps per call
    "heat512_plan": ("heat512", {"monitor_interval": 10}, 2),
    # small programs for the bitwise checks against the plain oracle (regions
    # wider than the stencil reach of the T-fold step)
    "gs3d_plan": ("gs3d_t", {"monitor_interval": 4, "initial_walls": {"x": [12], "y": [40], "z": [64, 100]}}, 2),
    "heat_plan": ("heat_t48", {"monitor_interval": 6, "initial_walls": {"x": [10, 30], "y": [24], "z": [36]}}, 3),
    "gs2d_plan": ("gs2d", {"monitor_interval": 4, "initial_walls": {"x": [100], "y": [60, 180]}}, 2),
    "heat_rk2_plan": ("heat_rk2_t", {"intra_node_shape": [32, 40, 36], "monitor_interval": 4,
                                     "initial_walls": {"x": [16], "y": [20], "z": [18]}}, 2),
    "gs3d_rk4_plan": ("gs3d_rk4_t", {"intra_node_shape": [40, 48, 56], "monitor_interval": 4,
                                     "initial_walls": {"x": [20], "y": [24], "z": [28]}}, 2),
    # nanoblock unroll (nbu<axis>N): two programs with 2-site nanoblocks
    "gs3d_nbu_plan": ("gs3d_t", {"monitor_interval": 4, "initial_walls": {"x": [12], "y": [40], "z": [64, 100]},
                                 "option_strings": ["omp", "omp-collapse", "nbux2", "nbuz2"]}, 2),
    "heat_nbu_plan": ("heat_t48", {"monitor_interval": 16, "initial_walls": {"x": [20], "y": [24], "z": [36]},
                                   "option_strings": ["omp", "omp-collapse", "nbux2"]}, 8),
    # the option shape of the reference's examples/pearson-3d.yaml (T = 8,
    # nbux2, one wall per axis) at the bench's heat 512^3: a second CPU baseline
    "heat512_t8_plan": ("heat512", {"monitor_interval": 16, "initial_walls": {"x": [272], "y": [272], "z": [272]},
                                    "option_strings": ["omp", "omp-collapse", "nbux2"]}, 8),
}


def planned_compiled(name: str):
    base, over, T = PLANNED[name]
    fmr, yaml_, _precision, _params, over0, _ = product_build.PROGRAMS[base]
    ov = dict(over0 or {})
    ov.update(over)
    ov["temporal_blocking_interval"] = T
    return compile_files(name, fmr, yaml_, "faithful", (), ov), T


def planned_path(name: str) -> str:
    return os.path.join(OUT, f"lib{name}.so")


def planned_layout(name: str):
    """Compiled program and the layout of its region-plan library (array
    shapes and margins from the plan)."""
    from oracle.emit_planned import _Emitter
    c, T = planned_compiled(name)
    nc = c.mm.numerical_config
    e = _Emitter(c.layout, c.mm, T, nc.initial_walls, nc.option_strings or ["omp", "omp-collapse"])
    return c, e.layout()


def build_planned(name: str) -> str:
    c, T = planned_compiled(name)
    src, _ = write_planned(OUT, c.layout, c.mm, T)
    out = planned_path(name)
    cmd = ["gcc"] + CFLAGS + ["-I", OUT, "-I", product_build.MPI_STUB, src, "-o", out + ".tmp", "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"region-plan build failed: {' '.join(cmd)}\n{r.stdout[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


def oracle_path(name: str) -> str:
    return os.path.join(OUT, f"lib{name}_cpu.so")


def compiled(name: str):
    fmr, yaml_, _precision, params, over, _ = product_build.PROGRAMS[name]
    # the oracle always computes in the program's declared (FP64) types
    return compile_files(name, fmr, yaml_, "faithful", params, over)


def compiled_extra(name: str):
    base, over, params = ORACLE_ONLY[name]
    fmr, yaml_, _precision, _params, over0, _ = product_build.PROGRAMS[base]
    ov = dict(over0 or {})
    ov.update(over)
    return compile_files(name, fmr, yaml_, "faithful", (), ov)


def build_oracle(name: str) -> str:
    c = compiled_extra(name) if name in ORACLE_ONLY else compiled(name)
    params = ORACLE_ONLY[name][2] if name in ORACLE_ONLY else ()
    src = write_oracle(OUT, c.layout, c.mm, params, copy_extern_header(c, OUT))
    out = oracle_path(name)
    cmd = ["gcc"] + CFLAGS + ["-I", OUT, "-I", product_build.MPI_STUB, src, "-o", out + ".tmp", "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stdout[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


def build_all(names=None):
    names = names or (list(product_build.PROGRAMS) + list(ORACLE_ONLY))
    out = {n: build_oracle(n) for n in names}
    for n in PLANNED:
        out[n] = build_planned(n)
    if "gs3d64" in out:
        # the example driver, linked against the CPU library of the same program
        out["gs3d64_main"] = product_build.build_driver("gs3d64", oracle_lib=out["gs3d64"])
    return out


if __name__ == "__main__":
    for k, v in build_all(sys.argv[1:] or None).items():
        print(k, v)
