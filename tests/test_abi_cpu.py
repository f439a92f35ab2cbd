"""The C ABI without a GPU: every built library exports every symbol the
headers declare, the libraries load, and the HIP library and the CPU oracle of
one program share one header (so one driver links against either)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from formura_amd import build
from formura_amd.compiler.abi import header_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header, prefix="Formura_"):
    text = open(os.path.join(ROOT, "include", header), encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def built(names):
    paths = [build.lib_path(n) for n in names]
    missing = [p for p in paths if not os.path.exists(p)]
    if missing:
        pytest.skip(f"libraries not built: {missing}")
    return paths


def test_step_libraries_export_the_abi():
    want = set(declared("formura_amd.h")) | {"Formura_Init", "Formura_Forward"}
    assert "Formura_Forward_Device" in want and "Formura_Diff_Device" in want
    for name, path in zip(build.PROGRAMS, built(build.PROGRAMS)):
        syms = exported(path)
        assert want <= syms, f"{name} lacks {sorted(want - syms)}"
        # state arrays and the navigator are exported data
        assert "navi" in syms


def test_batch_libraries_export_the_abi():
    from formura_amd.extensions import BATCH_PROGRAMS
    want = set(declared("formura_amd_batch.h"))
    assert "Formura_Batch_Advance" in want
    for name, path in zip(BATCH_PROGRAMS, built(BATCH_PROGRAMS)):
        syms = exported(path)
        assert want <= syms, f"{name} lacks {sorted(want - syms)}"


def test_surrogate_library_exports_the_abi():
    want = set(declared("formura_amd_surrogate.h", "Sur_"))
    assert {"Sur_Forward", "Sur_Train_Step", "Sur_Set_Conv"} <= want
    path = build.lib_path("surrogate")
    if not os.path.exists(path):
        pytest.skip("surrogate library not built")
    syms = exported(path)
    assert want <= syms, f"libsurrogate lacks {sorted(want - syms)}"


def test_libraries_load_without_a_gpu():
    for path in built(["gs3d_t", "heat_t48"]):
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        assert hasattr(lib, "Formura_Forward")


def test_oracle_and_hip_library_share_the_header():
    from oracle.build_oracle import compiled
    for name in ("gs3d64", "heat512", "gs2d"):
        c = compiled(name)
        gen = os.path.join(build.BUILD, f"{name}.h")
        if not os.path.exists(gen):
            pytest.skip("not generated")
        assert open(gen, encoding="utf-8").read() == header_text(c.layout)


def _reference_program(tree, fmr, nc_path, **over):
    from formura_amd.compiler.driver import compile_source
    from formura_amd.compiler.numerical_config import NumericalConfig
    with open(fmr, encoding="utf-8") as f:
        src = f.read()
    nc = NumericalConfig.from_file(nc_path)
    for k, v in over.items():
        setattr(nc, k, v)
    return compile_source("pearson", src, nc)


def test_header_extents_follow_the_reference_plan(reference_tree):
    """State-array extents come from the region plans as the reference's
    collaboratePlans makes them (MPICxx/Translate.hs:935-969): the hull of
    both plans' static allocations, with nbuMargin = nbu - 1 + 2 added to
    its upper vertex.  For the reference's own examples-generated/pearson.fmr
    + pearson.nc (600^2 per rank, walls x [200, 400], y [300], T = 4) the
    plans allocate U and V on (0,0)-(602,602) and (-2,-2)-(600,600)
    (pearson.log:85-88 and the inverted plan's list), whose hull is the
    604 x 604 of the committed pearson.c:9-10 -- a generator older than the
    nbuMargin term -- and the current generator's header adds the margin:
    U[606][606].  The config's 2 x 2 rank grid is taken as one rank: the
    plan is per rank, and each rank's arrays have the same extents."""
    from formura_amd.compiler.abi import plan_geometry
    d = os.path.join(reference_tree, "examples-generated")
    c = _reference_program(reference_tree, os.path.join(d, "pearson.fmr"), os.path.join(d, "pearson.nc"),
                           mpi_grid_shape=[1, 1])
    lo, hi, _ = plan_geometry(c.mm)
    hull = [h - 2 - l for h, l in zip(hi, lo)]        # without nbuMargin (nbu = 1: + 2)
    with open(os.path.join(d, "pearson.c"), encoding="utf-8") as f:
        ref = f.read()
    assert hull == [604, 604]
    assert "double U[604][604];" in ref and "double V[604][604];" in ref
    with open(os.path.join(d, "pearson.log"), encoding="utf-8") as f:
        log = f.read()
    assert "ResourceStatic \"U\" ()\n  \u300c(0,0) (602,602)\u300d" in log
    assert "ResourceStatic \"U\" ()\n  \u300c(-2,-2) (600,600)\u300d" in log
    h = header_text(c.layout)
    assert "extern double U[606][606];" in h and "extern double V[606][606];" in h
    assert c.layout.lower() == [2, 2] and c.layout.upper() == [602, 602]
    # the neighbour ranks are the plans' ridge deltas and their negations,
    # sorted, named as the reference names them (`toCName`); the log's
    # ridges cross ranks along (0,1) and (1,0) (pearson.log:191-197)
    fields = re.findall(r"int (mpi_rank_\w+);", h)
    assert fields == sorted(fields, key=lambda f: tuple(-int(x[1:]) if x.startswith("m") else int(x)
                                                           for x in f[len("mpi_rank__"):].split("_")))
    assert {"mpi_rank__0_1", "mpi_rank__1_0", "mpi_rank__0_m1", "mpi_rank__m1_0"} <= set(fields)
    assert all(re.fullmatch(r"mpi_rank__(m?1|0)_(m?1|0)", f) for f in fields)


def test_nanoblock_unroll_widens_the_upper_margin(reference_tree):
    """`nbux2` (examples/pearson-3d.yaml) adds nbu - 1 = 1 to the upper
    margin along x only (nbuMargin = nbu - 1 + 2, MPICxx/Translate.hs:942)."""
    ex = os.path.join(reference_tree, "examples")
    c2 = _reference_program(reference_tree, os.path.join(ex, "pearson-3d.fmr"), os.path.join(ex, "pearson-3d.yaml"))
    opts = [o for o in c2.mm.numerical_config.option_strings if not o.startswith("nbu")]
    c1 = _reference_program(reference_tree, os.path.join(ex, "pearson-3d.fmr"), os.path.join(ex, "pearson-3d.yaml"),
                            option_strings=opts)
    assert c2.layout.margin == c1.layout.margin
    assert [a - b for a, b in zip(c2.layout.margin_hi, c1.layout.margin_hi)] == [1, 0, 0]
