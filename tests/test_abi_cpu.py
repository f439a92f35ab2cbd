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
