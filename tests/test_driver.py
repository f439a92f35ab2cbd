"""The example driver (examples/gray-scott-3d-main.cpp), written against the
reference ABI only, linked against the CPU oracle library (CPU test) and
against the HIP library (GPU test): the monitor files parse, and the GPU run
writes the same bytes as the CPU run."""
from __future__ import annotations

import filecmp
import glob
import os
import subprocess

import numpy as np
import pytest

from formura_amd import build
from formura_amd.monitor import read_monitor

CPU_MAIN = os.path.join(build.ROOT, "oracle", "_build", "libgs3d64_cpu_main")
GPU_MAIN = os.path.join(build.BUILD, "gs3d64_main")


def run(exe, outdir, t_max=200, t_mon=100, seed=5):
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built")
    os.makedirs(outdir, exist_ok=True)
    r = subprocess.run([exe, str(t_max), str(t_mon), str(outdir), str(seed)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    return sorted(glob.glob(os.path.join(str(outdir), "monitor*.txt")))


def test_cpu_driver_monitor_files(tmp_path):
    files = run(CPU_MAIN, tmp_path / "cpu")
    assert [os.path.basename(f) for f in files] == [
        "monitorX-000000-0.txt", "monitorX-000100-0.txt", "monitorX-000200-0.txt",
        "monitorY-000000-0.txt", "monitorY-000100-0.txt", "monitorY-000200-0.txt"]
    m0 = read_monitor(files[0])
    assert m0["position"] == (0, 0, 0) and m0["extent"] == (64, 64, 64)
    assert m0["U"].shape == (64, 64)
    # the centred cube (u = 0.5 +- 0.01) crosses the mid plane
    assert abs(m0["U"][32, 32] - 0.5) <= 0.0100001
    m2 = read_monitor(files[2])
    assert np.isfinite(m2["U"]).all() and not np.array_equal(m2["U"], m0["U"])


@pytest.mark.gpu
def test_gpu_driver_writes_the_cpu_bytes(tmp_path):
    cpu = run(CPU_MAIN, tmp_path / "cpu")
    gpu = run(GPU_MAIN, tmp_path / "gpu")
    assert [os.path.basename(f) for f in cpu] == [os.path.basename(f) for f in gpu]
    for a, b in zip(cpu, gpu):
        assert filecmp.cmp(a, b, shallow=False), os.path.basename(a)
