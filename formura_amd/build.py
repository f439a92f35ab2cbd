"""Build the step libraries in-tree (formura_amd/_build/).

Every program in `PROGRAMS` is compiled by the formura_amd compiler into a
header and a `.hip` source, then by hipcc (gfx950) into `lib<name>.so`.
Flags: -O3, -ffp-contract=off (no fused multiply-add, so the arithmetic is the
reference C's statement for statement), denormals kept, packed-FP32 off.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import re
import shutil
import subprocess
import sys
from typing import Dict, List, Optional, Tuple

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "formura_amd")
BUILD = os.path.join(PKG, "_build")
CSRC = os.path.join(PKG, "csrc")
EXAMPLES = os.path.join(ROOT, "examples")
MPI_STUB = os.path.join(ROOT, "include", "formura_serial_mpi")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Xarch_host", "-mcmodel=medium"]


def _ex(name: str) -> str:
    return os.path.join(EXAMPLES, name)


# name -> (fmr, yaml, precision, params, nc overrides, march variants or None)
PROGRAMS: Dict[str, tuple] = {
    # Gray-Scott 3D, 64^3, FP64 (the reference's CPU-runnable case)
    "gs3d64": (_ex("gray-scott-3d-64.fmr"), _ex("gray-scott-3d-64.yaml"), "faithful", (), None, None),
    # heat 3D, 512^3, FP64 and FP32 bulk
    "heat512": (_ex("heat-3d-512.fmr"), _ex("heat-3d-512.yaml"), "faithful", (), None, None),
    "heat512_f32": (_ex("heat-3d-512.fmr"), _ex("heat-3d-512.yaml"), "fp32", (), None, None),
    # Gray-Scott 3D, 512^3: FP64 oracle run on the GPU and FP32 bulk
    "gs3d512": (_ex("gray-scott-3d-512.fmr"), _ex("gray-scott-3d-512.yaml"), "faithful", (), None, None),
    "gs3d512_f32": (_ex("gray-scott-3d-512.fmr"), _ex("gray-scott-3d-512.yaml"), "fp32", (), None, None),
    # small odd-shaped variants for parity tests against the CPU oracle
    "heat_t48": (_ex("heat-3d-512.fmr"), _ex("heat-3d-512.yaml"), "faithful", (),
                 {"intra_node_shape": [40, 48, 72], "monitor_interval": 12}, None),
    "heat_t48_f32": (_ex("heat-3d-512.fmr"), _ex("heat-3d-512.yaml"), "fp32", (),
                     {"intra_node_shape": [40, 48, 72], "monitor_interval": 12}, None),
    "gs3d_t": (_ex("gray-scott-3d-64.fmr"), _ex("gray-scott-3d-64.yaml"), "faithful", (),
               {"intra_node_shape": [36, 70, 130], "monitor_interval": 20}, None),
    "gs2d": (_ex("gray-scott-2d.fmr"), _ex("gray-scott-2d.yaml"), "faithful", (), None, None),
    # multi-stage programs (manifest nodes): explicit midpoint heat, RK4 Gray-Scott
    "heat_rk2": (_ex("heat-3d-rk2.fmr"), _ex("heat-3d-rk2.yaml"), "faithful", (), None, None),
    "heat_rk2_t": (_ex("heat-3d-rk2.fmr"), _ex("heat-3d-rk2.yaml"), "faithful", (),
                   {"intra_node_shape": [24, 40, 36], "monitor_interval": 4}, None),
    "gs3d_rk4": (_ex("gray-scott-3d-rk4.fmr"), _ex("gray-scott-3d-rk4.yaml"), "faithful", (), None, None),
    "gs3d_rk4_t": (_ex("gray-scott-3d-rk4.fmr"), _ex("gray-scott-3d-rk4.yaml"), "faithful", (),
                   {"intra_node_shape": [20, 34, 40], "monitor_interval": 4}, None),
    # isothermal gas dynamics, finite volumes (MUSCL + Rusanov), SSP-RK3: 48
    # manifest nodes, a point function too large to inline (stage kernels only)
    "fveuler": (_ex("fv-euler-3d.fmr"), _ex("fv-euler-3d.yaml"), "faithful", (), None, None),
    "fveuler_t": (_ex("fv-euler-3d.fmr"), _ex("fv-euler-3d.yaml"), "faithful", (),
                  {"intra_node_shape": [22, 26, 34], "monitor_interval": 2}, None),
    # isothermal ideal MHD with GLM divergence cleaning and a passive scalar:
    # nine fields, 108 manifest nodes (SSP-RK3 stages), stage kernels only
    "mhd": (_ex("mhd-glm-3d.fmr"), _ex("mhd-glm-3d.yaml"), "faithful", (), None, None),
    "mhd_t": (_ex("mhd-glm-3d.fmr"), _ex("mhd-glm-3d.yaml"), "faithful", (),
              {"intra_node_shape": [20, 24, 28], "monitor_interval": 2}, None),
    # an extern function call and index variables (global coordinates) in
    # the step; and the power operator (pow)
    "forced_heat": (_ex("forced-heat-3d.fmr"), _ex("forced-heat-3d.yaml"), "faithful", (), None, None),
    "power_heat": (_ex("power-heat-3d.fmr"), _ex("power-heat-3d.yaml"), "faithful", (), None, None),
}


def lib_path(name: str) -> str:
    return os.path.join(BUILD, f"lib{name}.so")


def generate(name: str):
    sys.path.insert(0, ROOT)
    from formura_amd.compiler.driver import compile_files, write_hip
    fmr, yaml_, precision, params, over, variants = PROGRAMS[name]
    c = compile_files(name, fmr, yaml_, precision, params, over)
    h, src = write_hip(c, BUILD, variants)
    return c, h, src


def _run(cmd: List[str]) -> str:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout[-6000:]}")
    return r.stdout


RESOURCE_FLAG = "-Rpass-analysis=kernel-resource-usage"
_MARCH_NAME = re.compile(r"march3d_kernel\w*?ProgELi(\d+)ELi(\d+)ELi(\d+)ELi(\d+)ELi(\d+)E")
_MARCHV_NAME = re.compile(r"marchv_kernel\w*?ProgELi(\d+)ELi(\d+)ELi(\d+)ELi(\d+)ELb(\d)ELi(\d+)ELi(\d+)ELi(\d+)E")


def parse_resources(text: str) -> Dict[str, Dict[str, int]]:
    """Per-kernel VGPRs, AGPRs, ScratchSize (bytes per lane) and Occupancy
    (waves per SIMD) from the compiler's kernel-resource-usage remarks."""
    out: Dict[str, Dict[str, int]] = {}
    cur = None
    for line in text.splitlines():
        m = re.search(r"remark: Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|AGPRs|ScratchSize|Occupancy)(?: \[[^\]]*\])?: (\d+)", line)
        if m and cur:
            out[cur][m.group(1)] = int(m.group(2))
    return out


def march_resources(res: Dict[str, Dict[str, int]]) -> Dict[tuple, Dict[str, int]]:
    """(kind, T, rows, waves, prefetch, occupancy) -> resources, for the march
    kernels (kind 1: march3d_kernel; 3 / 4: marchv_kernel with one / two
    z-sites per lane; 5 / 6: the same with edge rows)."""
    out = {}
    for name, r in res.items():
        m = _MARCH_NAME.search(name)
        if m:
            out[(1,) + tuple(int(x) for x in m.groups())] = r
            continue
        m = _MARCHV_NAME.search(name)
        if m:
            T, by, W, vz, _xc, pf, occ, ey = (int(x) for x in m.groups())
            out[(2 + vz + (2 if ey else 0), T, by, W, pf, occ)] = r
    return out


def compile_hip(src: str, out: str, extra: Optional[List[str]] = None) -> str:
    tmp = out + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", CSRC, "-I", MPI_STUB, "-I", BUILD, "-I", INCLUDE] + (extra or []) + \
        [RESOURCE_FLAG, src, "-o", tmp]
    text = _run(cmd)
    os.replace(tmp, out)
    res = march_resources(parse_resources(text))
    with open(out[:-3] + ".resources.json", "w", encoding="utf-8") as f:
        json.dump({",".join(map(str, k)): v for k, v in sorted(res.items())}, f, indent=0)
    return out


def resources(name: str) -> Dict[tuple, Dict[str, int]]:
    """Resources of the march variants compiled into lib<name>.so, by
    (kind, T, rows, waves, prefetch, occupancy)."""
    with open(lib_path(name)[:-3] + ".resources.json", encoding="utf-8") as f:
        return {tuple(int(x) for x in k.split(",")): v for k, v in json.load(f).items()}


def build_program(name: str) -> str:
    """Generate and compile one program.  A march variant whose registers
    spill to scratch memory is compiled again with one row per wave fewer;
    if that spills too it is dropped."""
    c, _, src = generate(name)
    out = compile_hip(src, lib_path(name))
    spilled = {k for k, r in resources(name).items() if r.get("ScratchSize", 0) > 0}
    if spilled:
        from formura_amd.compiler.driver import write_hip
        from formura_amd.compiler.emit_hip import (KIND_MARCHV1E, KIND_MARCHV2E, _with_kind, march_edge_rows,
                                                   variants_for)
        base = [_with_kind(v) for v in (PROGRAMS[name][5] or variants_for(c.layout, c.precision))]
        trial = []
        for v in base:
            kind, T, by, W = v[:4]
            fewer = (kind, T, by - 1) + v[3:]
            edge = kind in (KIND_MARCHV1E, KIND_MARCHV2E)
            if v not in spilled:
                trial.append(v)
            elif (by - 1) * W > 2 * T and fewer not in base and (not edge or by - 1 + march_edge_rows(T) > T):
                trial.append(fewer)
        _, src = write_hip(c, BUILD, trial)
        out = compile_hip(src, lib_path(name))
        spilled = {k for k, r in resources(name).items() if r.get("ScratchSize", 0) > 0}
        if spilled:
            _, src = write_hip(c, BUILD, [v for v in trial if v not in spilled])
            out = compile_hip(src, lib_path(name))
    return out


def build_copy_probe() -> str:
    """The HBM copy calibration binary (formura_amd/_build/hbm_copy)."""
    out = os.path.join(BUILD, "hbm_copy")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", os.path.join(CSRC, "hbm_copy.hip"), "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_driver(name: str = "gs3d64", oracle_lib: Optional[str] = None) -> str:
    """examples/gray-scott-3d-main.cpp linked against the HIP library of a 3D
    Gray-Scott program (or, with oracle_lib, against that CPU library)."""
    if oracle_lib:
        libdir, lib, out = os.path.dirname(oracle_lib), os.path.basename(oracle_lib)[3:-3], oracle_lib[:-3] + "_main"
        extra = ["-fopenmp"]
        incdir = libdir
    else:
        libdir, lib, out = BUILD, name, os.path.join(BUILD, f"{name}_main")
        extra = []
        incdir = BUILD
    cmd = ["g++", "-O2", f'-DGS3D_HEADER="{name}.h"', "-I", incdir, "-I", MPI_STUB,
           os.path.join(EXAMPLES, "gray-scott-3d-main.cpp"), "-L", libdir, f"-l{lib}",
           f"-Wl,-rpath,{libdir}"] + extra + ["-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all(names=None, jobs: int = 8) -> Dict[str, str]:
    os.makedirs(BUILD, exist_ok=True)
    if names is None:
        build_copy_probe()
    names = list(names or PROGRAMS)
    for n in names:      # code generation is quick and serial
        generate(n)
    out = {}
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(build_program, n): n for n in names}
        for f in cf.as_completed(futs):
            out[futs[f]] = f.result()
    if "gs3d64" in names:
        out["gs3d64_main"] = build_driver("gs3d64")
    return out


if __name__ == "__main__":
    names = sys.argv[1:] or None
    for k, v in build_all(names).items():
        print(k, v)
