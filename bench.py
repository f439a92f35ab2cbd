#!/usr/bin/env python3
"""Benchmark of the generated explicit step on one MI355X.

Headline workload (one line of JSON on stdout): the 3D heat equation
u_t = D lap(u), 7-point stencil, 512^3 periodic grid, FP64, D=1, dx=1, dt=0.1,
initial u i.i.d. N(0,1) -- compiled from examples/heat-3d-512.fmr into a HIP
step library.  A "step" is one time step over the whole grid; `value` is
lattice-site updates per second (GLUP/s) with the state resident in HBM.
The kernel variant (fusion depth T, tile shape, work split) is chosen by a
short timed search before the timed region; the best rate for each compiled
fusion depth T is reported as well (t_sweep_glups), with its fraction of the
HBM roofline (t_sweep_frac: 16 / T bytes per update).

Also reported (same line, key "extra"): the FP32 bulk heat rate, 3D Gray-Scott
512^3 in FP64 and FP32, 3D Gray-Scott 64^3 FP64 (1000 steps), the multi-stage
programs (explicit midpoint heat 256^3, RK4 Gray-Scott 64^3, finite-volume
isothermal gas dynamics with SSP-RK3 256^3: stage kernels, with their own
roofline, and the inlined point function where there is one), the batched 2D
Gray-Scott sweep (256 uneven instances x 5000 steps, FP32 vs FP64 vs the plain
method) and the checked surrogate on held-out 256^2 trajectories.

cpu_baseline: the same program as C/OpenMP on all host cores, built with gcc
-O3 -fopenmp -ffp-contract=off, on a bounded sample (about 96 steps of the
full 512^3 grid for the region-plan library, 20 for the plain oracle): the
library generated on the reference's own region plan (Cut.hs restated) and
the plain oracle are both timed; the faster is reported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, published


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def seeded_state(lib, seed):
    import numpy as np
    from formura_amd.rng import gray_scott_ic, normal
    rng = np.random.default_rng(seed)
    shape = tuple(lib.layout.intra)
    if lib.layout.field_names == ["u"]:
        return {"u": rng.standard_normal(shape)}
    if len(lib.layout.field_names) == 9:
        from formura_amd.rng import mhd_state
        return mhd_state(shape, seed)
    if lib.layout.field_names == ["r", "mx", "my", "mz"]:
        r = 1.0 + 0.2 * normal(seed, shape)
        return {"r": r, "mx": r * 0.1 * normal(seed + 1, shape), "my": r * 0.1 * normal(seed + 2, shape),
                "mz": r * 0.1 * normal(seed + 3, shape)}
    u, v = gray_scott_ic(shape, seed)
    return {"U": u, "V": v}


def has_direct(lib):
    """Whether the library has the inlined point function (direct kernel)."""
    from formura_amd.runtime import KIND_DIRECT
    try:
        lib.set_kernel(KIND_DIRECT)
        return True
    except ValueError:
        return False


def staged_bytes_per_site(name):
    """Algorithmic bytes per site and step of the stage kernels: each kernel
    reads each array it loads once and writes each of its outputs once."""
    from formura_amd import build
    from formura_amd.compiler.driver import compile_files
    from formura_amd.compiler.pointir import group_stages, lower_stages
    fmr, yaml_, precision, params, over, _ = build.PROGRAMS[name]
    c = compile_files(name, fmr, yaml_, precision, params, over)
    elem = 8 if precision == "faithful" else 4
    total = 0
    for g in group_stages(lower_stages(c.mm, precision, params)):
        arrays = {op.args[0] for op in g.ir.ops if op.kind == "load"}
        total += (len(arrays) + len(g.members)) * elem
    return total


def candidates(lib):
    """Every (kernel, work split) the search may pick: the direct kernel, the
    stage kernels, and each compiled march variant with each split of
    runtime.march_chunks -- the set the full-size parity tests run."""
    from formura_amd.runtime import KIND_DIRECT, KIND_STAGED, march_chunks
    out = [((KIND_DIRECT, 1, 0, 1, 1, 1), 64)] if has_direct(lib) else []
    if lib.num_stages() > 0:
        for rows in (4, 16):     # wavefronts per workgroup of the stage kernels
            out.append(((KIND_STAGED, 1, rows, 1, 1, 1), 64))
    for v in lib.list_kernels():
        for lx in march_chunks(v[5]):   # run length, or -workgroups per SIMD
            out.append((tuple(v), lx))
    return out


def autotune(lib, probe_steps=24, final_steps=120, finalists=4):
    """Time every compiled variant briefly, then the best few again over more
    steps; return {T: (GLUP/s, variant, lx)} (best per fusion depth T, from the
    longer timing where there is one)."""
    import numpy as np
    cells = int(np.prod(lib.layout.intra))

    def rate(var, lx, steps):
        T = var[1]
        lib.set_kernel(*var)
        lib.set_chunk(lx)
        steps = max(T, (steps // T) * T)
        lib.forward_device_timed(T)
        lib.prepare(steps)
        ms = lib.forward_device_timed(steps)
        return steps * cells / (ms * 1e-3) / 1e9

    probes = sorted(((rate(var, lx, probe_steps), var, lx) for var, lx in candidates(lib)), reverse=True)
    best = {}
    for r, var, lx in probes:
        if var[1] not in best:
            best[var[1]] = (r, var, lx)
    for _, var, lx in probes[:finalists]:
        r = rate(var, lx, final_steps)
        T = var[1]
        if best[T][1:] == (var, lx) or r > best[T][0]:
            best[T] = (r, var, lx)
    return best


def spin_up(lib, var, lx, seconds=2.0):
    """Run the kernel (var, lx) for about `seconds` of wall time, untimed: the
    shader clock of an idle GPU takes a while to rise, which the search's
    timings otherwise do before the timed region.  For runs that skip the
    search (--variant)."""
    lib.set_kernel(*var)
    lib.set_chunk(lx)
    t0 = time.perf_counter()
    n = max(var[1], 8)
    while time.perf_counter() - t0 < seconds:
        lib.forward_device(n)
        n = min(2 * n, 4096)


def timed_run(lib, var, lx, steps, warmup):
    """Warm up, then time exactly `steps` steps (device synchronised on both
    sides).  Returns (wall seconds, GPU event milliseconds, milliseconds of the
    selected kernel's own launches, their count).  The hipGraph that the timed
    call replays is captured and instantiated before the clock starts
    (Formura_Prepare_Device: it launches nothing); every step's kernels run
    inside the timed region."""
    lib.set_kernel(*var)
    lib.set_chunk(lx)
    if warmup:
        lib.forward_device(warmup)          # returns after the GPU is idle
    lib.prepare(steps)
    t0 = time.perf_counter()
    ev_ms = lib.forward_device_timed(steps)  # synchronises on the library stream
    t1 = time.perf_counter()
    main_ms, main_launches, _ = lib.last_timing()
    return t1 - t0, ev_ms, main_ms, main_launches


def _time_cpu_library(lib, cells, per, sample_steps):
    import numpy as np
    rng = np.random.default_rng(1)
    lib.set_state({"u": rng.standard_normal(tuple(lib.layout.intra))})
    lib.forward()                       # warm-up (first touch, thread start)
    n = max(1, sample_steps // per)
    t0 = time.perf_counter()
    for _ in range(n):
        lib.forward()
    dt = time.perf_counter() - t0
    return n * per * cells / dt / 1e9, n * per, dt


def cpu_baseline(sample_steps=96, plain_steps=20):
    """Heat 512^3 FP64 on the host cores, about `sample_steps` steps of each
    region-plan library and `plain_steps` of the plain oracle: the C generated
    on the reference's own region plan (Cut.hs restated: skewed in-place
    stores, ridges and self-message facets) with 2 x 2 x 2 regions at T = 2,
    and with the option shape of the reference's examples/pearson-3d.yaml
    (T = 8, 2-site nanoblocks along x), and the plain oracle (one loop nest
    per manifest node over the whole grid).  The fastest one is the
    baseline."""
    import numpy as np
    from oracle.build_oracle import compiled_extra, oracle_path, planned_layout, planned_path
    from formura_amd.runtime import StepLibrary
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    runs = {}
    for name, label in (("heat512_plan", "region plan T=2"), ("heat512_t8_plan", "region plan T=8 nbux2")):
        c, lay = planned_layout(name)
        lib = StepLibrary(planned_path(name), lay, device=False)
        lib.init()
        cells = int(np.prod(lay.intra))
        runs[label] = _time_cpu_library(lib, cells, lay.steps_per_forward, sample_steps)
    c2 = compiled_extra("heat512_cpu")
    lib2 = StepLibrary(oracle_path("heat512_cpu"), c2.layout, device=False)
    lib2.init()
    runs["plain oracle"] = _time_cpu_library(lib2, cells, c2.layout.steps_per_forward, plain_steps)
    best = max(runs, key=lambda k: runs[k][0])
    rate, steps, dt = runs[best]
    other = [f"{k} {v[0]:.3f} GLUP/s" for k, v in runs.items() if k != best]
    what = {"region plan T=2": "C on the reference's region plan (Cut.hs restated: 2x2x2 regions, T=2)",
            "region plan T=8 nbux2": "C on the reference's region plan with the options of its pearson-3d.yaml "
                                     "(T=8, nbux2, 2x2x2 regions)",
            "plain oracle": "plain oracle C (one loop nest per manifest node)"}[best]
    return {"value": round(rate, 4), "unit": "GLUP/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of heat 512^3 FP64, {what}, gcc -O3 -fopenmp -ffp-contract=off, "
                      f"{dt:.1f} s; also timed: {'; '.join(other)}"}


# registered program -> label of its line under "extra"
EXTRA_PROGRAMS = (("heat512_f32", "heat512_fp32"), ("gs3d512", "grayscott512_fp64"),
                  ("gs3d512_f32", "grayscott512_fp32"), ("gs3d64", "grayscott64_fp64"),
                  ("heat_rk2", "heat256_midpoint_fp64"), ("gs3d_rk4", "grayscott64_rk4_fp64"),
                  ("fveuler", "fv_euler256_ssprk3_fp64"), ("mhd", "mhd_glm128_ssprk3_fp64"))


GRAPHS = True   # --no-graphs: eager launches (for tracers that cannot follow graph replays)


def program_extras(only=None, libs=None, variant=None):
    """Autotuned rate of each program of EXTRA_PROGRAMS (or only the one
    labelled `only`, and with `variant` = ((kind, T, rows, waves, pf, occ), lx)
    that kernel instead of the search's pick); multi-stage programs also
    report their stage kernels.  The loaded libraries are left in `libs`
    (name -> library), each set to the kernel it was timed with."""
    import numpy as np
    from formura_amd.runtime import load_program
    extra = {}
    libs = {} if libs is None else libs
    for name, label in EXTRA_PROGRAMS:
        if only and label != only:
            continue
        lg = load_program(name)
        lg.init()
        lg.set_graphs(GRAPHS)
        libs[name] = lg
        lg.set_state(seeded_state(lg, 1))
        lg.upload()
        n = int(np.prod(lg.layout.intra))
        big = n >= 2 ** 24 or lg.num_stages() > 0
        if variant:
            if variant not in candidates(lg):
                print(f"bench.py: variant {variant} of {name} is not one the parity tests run", file=sys.stderr)
                sys.exit(2)
            spin_up(lg, *variant)
            best = {variant[0][1]: (0.0,) + tuple(variant)}
        else:
            best = autotune(lg, probe_steps=24 if big else 200, final_steps=96 if big else 1000)
        r, v, l = max(best.values())
        steps = 100 if big else 1000
        w, _, _, _ = timed_run(lg, v, l, steps, 10)
        extra[label] = {"glups": round(steps * n / w / 1e9, 2), "kernel": list(v) + [l],
                        "best_per_T": {T: round(x[0], 1) for T, x in sorted(best.items())}}
        if lg.num_stages() > 0:   # multi-stage program: stage kernels vs the inlined point function
            from formura_amd.runtime import KIND_DIRECT, KIND_STAGED
            kinds = [(KIND_STAGED, "staged_glups")] + ([(KIND_DIRECT, "inlined_direct_glups")]
                                                       if has_direct(lg) else [])
            for kind, key in kinds:
                # stage kernels: workgroups of 4 or 16 wavefronts, the faster one
                runs = {rows: timed_run(lg, (kind, 1, rows, 1, 1, 1), 64, 20, 2)
                        for rows in ((4, 16) if kind == KIND_STAGED else (0,))}
                rows = min(runs, key=lambda r: runs[r][0])
                w2, ev2, _, _ = runs[rows]
                extra[label][key] = round(20 * n / w2 / 1e9, 2)
                if kind == KIND_STAGED:
                    b = staged_bytes_per_site(name) * n
                    gbs = b * 20 / (ev2 * 1e-3) / 1e9
                    extra[label]["staged"] = {
                        "manifest_nodes": lg.num_stages(), "kernels_per_step": lg.num_stage_kernels(),
                        "rows": rows, "glups_by_rows": {r: round(20 * n / v[0] / 1e9, 2) for r, v in runs.items()},
                        "algorithmic_bytes_per_step": b, "ms_per_step": round(ev2 / 20, 4),
                        "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        lg.set_kernel(*v)                  # leave the library on the kernel it was timed with
        lg.set_chunk(l)
        log(f"{label}: {extra[label]}")
    return extra


def precision_errors(pairs, seed=2):
    """FP32 bulk against its FP64 twin on the GPU, on the same seeded input:
    for each (label, FP32 library, FP64 library, steps), both advance `steps`
    steps with the kernels they were timed with, and every field is compared
    on the device with FP64 reductions (relative L2 and max abs)."""
    out = {}
    for label, f32, f64, steps in pairs:
        st = seeded_state(f64, seed)
        for lib in (f32, f64):
            lib.set_state(st)
            lib.upload()
            lib.forward_device(steps)
        per = {}
        for i, f in enumerate(f64.layout.field_names):
            d, r, m = f32.diff(i, f64)
            per[f] = {"rel_l2": d / r, "max_abs": m}
        out[label] = {"steps": steps, "fields": per, "reductions": "FP64",
                      "fp32_kernel": list(f32.get_kernel()), "fp64_kernel": list(f64.get_kernel())}
        log(f"{label} FP32 vs FP64: {out[label]}")
    return out


def copy_peak():
    """Measured HBM copy rate (bytes read + written per second) of a plain
    16-byte-access copy kernel (formura_amd/csrc/hbm_copy.hip), in GB/s."""
    import subprocess
    from formura_amd import build
    exe = os.path.join(build.BUILD, "hbm_copy")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    for line in r.stdout.splitlines():
        if line.startswith("copy_peak_GBs"):
            return float(line.split()[1])
    return None


def _pmc_per_dispatch(outdir, counter, kernel_part):
    """Mean over dispatches of `counter` (summed over the CSV's rows of one
    dispatch) for kernels whose name contains `kernel_part`."""
    import csv
    per = {}
    for root, _, files in os.walk(outdir):
        for fn in files:
            if fn.endswith("counter_collection.csv"):
                with open(os.path.join(root, fn), newline="") as f:
                    for r in csv.DictReader(f):
                        if kernel_part in r["Kernel_Name"] and r["Counter_Name"] == counter:
                            per[r["Dispatch_Id"]] = per.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    return sum(per.values()) / len(per) if per else None


def measure_traffic(var, lx):
    """HBM bytes per launch of march variant (var, lx), measured now: four
    rocprofv3 --pmc runs as child processes (one counter each, as rocprofv3
    collects one pass per run): FETCH_SIZE and WRITE_SIZE of 12 steps of this
    bench with the variant fixed, and of the 1 GiB copy kernel of
    formura_amd/csrc/hbm_copy.hip, whose known bytes calibrate the counters'
    units.  Returns (bytes, details) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile
    from formura_amd import build
    prof = shutil.which("rocprofv3")
    exe = os.path.join(build.BUILD, "hbm_copy")
    if not prof or not os.path.exists(exe):
        return None, "rocprofv3 or hbm_copy missing"
    tag = ",".join(str(x) for x in list(var) + [lx])
    bench = [sys.executable, os.path.abspath(__file__), "--variant", tag, "--steps", "12", "--warmup", "0",
             "--no-extra", "--no-cpu-baseline", "--no-traffic"]
    env = dict(os.environ, TMPDIR="/tmp")
    vals = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for key, counter, cmd, kpart in (("fetch", "FETCH_SIZE", bench, "march"), ("write", "WRITE_SIZE", bench, "march"),
                                         ("copy_fetch", "FETCH_SIZE", [exe], "copy"),
                                         ("copy_write", "WRITE_SIZE", [exe], "copy")):
            d = os.path.join(tmp, key)
            log(f"traffic: rocprofv3 --pmc {counter} ({key}) ...")
            try:
                r = subprocess.run(["timeout", "-s", "KILL", "120", prof, "--pmc", counter, "--output-format", "csv",
                                    "-d", d, "-o", key, "--"] + cmd, cwd="/tmp", env=env,
                                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=150)
            except (subprocess.TimeoutExpired, OSError) as e:
                return None, f"{key}: {e}"
            v = _pmc_per_dispatch(d, counter, kpart)
            if r.returncode != 0 and v is None:
                return None, f"{key}: rocprofv3 exit {r.returncode}"
            if v is None or v <= 0:
                return None, f"{key}: no {counter} rows"
            vals[key] = v
    copy_bytes = 1 << 30
    rf = copy_bytes / (vals["copy_fetch"] * 1024)      # bytes per FETCH_SIZE KiB unit
    wf = copy_bytes / (vals["copy_write"] * 1024)
    read_b, write_b = vals["fetch"] * 1024 * rf, vals["write"] * 1024 * wf
    return int(read_b + write_b), {"FETCH_SIZE_KiB": round(vals["fetch"], 1), "WRITE_SIZE_KiB": round(vals["write"], 1),
                                   "copy_FETCH_SIZE_KiB": round(vals["copy_fetch"], 1),
                                   "copy_WRITE_SIZE_KiB": round(vals["copy_write"], 1),
                                   "read_bytes": int(read_b), "write_bytes": int(write_b)}


def cpu_model():
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 traffic passes")
    ap.add_argument("--variant", default="", help="kind,T,rows,waves,pf,occ,lx: skip the search (of heat 512^3, "
                                                 "or with --only of that program)")
    ap.add_argument("--only", default="", help="run only this line of 'extra' (e.g. for a profile)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch eagerly instead of replaying hipGraphs (same kernels; for kernel tracers)")
    a = ap.parse_args()
    global GRAPHS
    GRAPHS = not a.no_graphs
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design note, "
              "not a measured path", file=sys.stderr)
        sys.exit(2)

    if a.only:
        pick = None
        if a.variant:
            v = [int(x) for x in a.variant.split(",")]
            pick = (tuple(v[:6]), v[6])
        print(json.dumps({"only": a.only, "extra": program_extras(a.only, variant=pick)}), flush=True)
        return

    import numpy as np
    from formura_amd.runtime import load_program

    peak_copy = copy_peak() if not a.variant else None   # before this process's own GPU work
    lib = load_program("heat512")
    lib.init()
    lib.set_graphs(GRAPHS)
    cells = int(np.prod(lib.layout.intra))
    lib.set_state(seeded_state(lib, 1))
    lib.upload()
    sweep = {}
    if a.variant:
        v = [int(x) for x in a.variant.split(",")]
        var, lx = tuple(v[:6]), v[6]
        if (var, lx) not in candidates(lib):
            print(f"bench.py: variant {a.variant} is not one the parity tests run "
                  f"(a compiled march variant with a split of runtime.march_chunks)", file=sys.stderr)
            sys.exit(2)
        spin_up(lib, var, lx)
    else:
        log("autotuning heat512 FP64 ...")
        best = autotune(lib)
        sweep = {T: round(r, 1) for T, (r, _, _) in sorted(best.items())}
        _, var, lx = max(best.values())
        log(f"best per T: {sweep}; chosen {var} lx={lx}")
    T = var[1]
    wall, ev_ms, main_ms, launches = timed_run(lib, var, lx, a.steps, a.warmup)
    glups = a.steps * cells / wall / 1e9
    # the T-fused launches of the selected kernel, timed by events on the
    # library stream (the steps left when steps % T != 0 run after them, apart)
    bytes_per_launch = 2 * 8 * cells          # read + write of the FP64 state once per fused launch
    launch_ms = main_ms / max(1, launches)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    s0, _ = lib.reduce(0)

    # HBM bytes per launch of this kernel variant: measured in this run by
    # rocprofv3 FETCH_SIZE / WRITE_SIZE passes (child processes, after the
    # timed region) scaled by the copy-kernel calibration; if that cannot run,
    # the committed measurement of the same variant (profiles/), else null
    traffic, traffic_source, traffic_detail = None, None, None
    if not a.no_traffic and not a.variant:
        traffic, traffic_detail = measure_traffic(var, lx)
        traffic_source = "rocprofv3 --pmc in this run" if traffic else None
        log(f"traffic: {traffic} {traffic_detail}")
    if traffic is None:
        tfile = os.path.join(ROOT, "profiles", "traffic_heat512.json")
        if os.path.exists(tfile):
            with open(tfile) as f:
                t = json.load(f)
            rec = t.get("variants", {}).get(",".join(str(x) for x in list(var) + [lx]))
            if rec:
                traffic = rec["hbm_bytes_per_launch"]
                traffic_source = "profiles/traffic_heat512.json (an earlier session)"

    out = {
        "metric": "GLUP/s & %HBM roofline; batched/surrogate speedup at <=1e-3 rel L2 vs FP64",
        "value": round(glups, 2),
        "unit": "GLUP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: u ~ N(0,1) i.i.d. (numpy PCG64 seed 1)",
        "config": {"workload": "heat3d 7-point, 512^3 periodic, FP64, D=1 dx=1 dt=0.1",
                   "grid": [512, 512, 512], "kernel": {"kind": var[0], "T": var[1], "rows_per_wave": var[2],
                                                       "waves": var[3], "prefetch": var[4], "waves_per_simd": var[5],
                                                       "march_chunk": lx, "march_chunk_note": "lx > 0: planes per workgroup run; -300 <= lx < -200: whole runs, k per tile, "
                                                                           "one resident round at (-lx - 200) waves per SIMD; "
                                                                           "lx < -300: whole runs at (-lx - 300) waves per SIMD, "
                                                                           "the idle resident workgroups taking the last planes of every tile; "
                                                                           "-200 <= lx < -100: balanced runs, "
                                                                           "(-lx - 100) waves per SIMD, one run per workgroup; "
                                                                           "other lx < 0: x whole, -lx waves per SIMD"}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_vs_algorithmic": round(traffic / bytes_per_launch, 3) if traffic else None,
                     # the bytes HBM actually moved per second (traffic / launch time) and their
                     # fraction of the 8 TB/s peak; `achieved` and `frac` count algorithmic bytes
                     "moved_GBs": round(traffic / (launch_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "moved_frac": round(traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "traffic_source": traffic_source, "traffic_counters": traffic_detail if traffic else None,
                     "measured_copy_peak": peak_copy,
                     "frac_of_measured_copy_peak": round(achieved / peak_copy, 4) if peak_copy else None},
        "t_sweep_glups": sweep,
        # fraction of the 8 TB/s roofline per fusion depth: a T-fused launch
        # reads and writes the FP64 state once, 16 / T bytes per update
        "t_sweep_frac": {T: round(r * 16 / T / HBM_PEAK_GBS, 4) for T, r in sweep.items()},   # GLUP/s x B = GB/s
        "kernel_ms_per_launch": round(launch_ms, 5),
        "kernel_launches": launches,
        "gpu_event_ms": round(ev_ms, 4),
        "sum_u_after": s0,
    }

    if not a.no_extra:
        libs = {}
        extra = program_extras(libs=libs)
        out["extra"] = extra
        # FP32 bulk against its FP64 twin, same input, the configs' step counts
        # (heat 512^3: 200 steps; Gray-Scott 512^3: 10000 steps)
        log("FP32 vs FP64 ...")
        lib.set_kernel(*var)
        lib.set_chunk(lx)
        extra["fp32_vs_fp64"] = precision_errors([
            ("heat512", libs["heat512_f32"], lib, 200),
            ("grayscott512", libs["gs3d512_f32"], libs["gs3d512"], 10000)])
        # batched Gray-Scott sweep: 256 uneven 2D instances x 5000 steps, FP32
        # bulk (hipGraph, autotuned plan) vs its FP64 twin and the plain method
        from formura_amd.batched import sweep_benchmark
        log("batched sweep ...")
        extra["batched_sweep"] = sweep_benchmark(log=log)
        # checked surrogate on held-out 256^2 trajectories (committed weights)
        from formura_amd import surrogate as S
        if os.path.exists(S.WEIGHTS):
            log("checked surrogate ...")
            theta, scale, _ = S.load_weights()
            extra["surrogate"] = S.evaluate(theta, scale, log=log)

    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_baseline"]["cpu_model"] = cpu_model()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
