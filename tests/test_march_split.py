"""The work split of the march kernels, on the host: the C++ functions the
kernels and the step library call (march_split_of, march_item_range,
march_runs in formura_amd/csrc/fmr_stencil.hpp) compiled with hipcc into a
small host program (tests/host/march_split_check.cpp).  For every split mode
of runtime.march_chunks and many shapes, the runs of all workgroups store
every (tile, plane) exactly once; and the spare-workgroup split of the heat
512^3 FP64 edge-row tile shortens the longest workgroup's iterations."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from formura_amd.build import CSRC, HIPCC
from formura_amd.runtime import balanced_runs, march_chunks, spare_runs, whole_runs

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("split") / "march_split_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "host", "march_split_check.cpp"), "-o", exe],
                   check=True, capture_output=True)

    def run(cases):
        text = "".join(" ".join(map(str, c)) + "\n" for c in cases)
        out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
        return [tuple(int(x) for x in line.split()) for line in out if line.strip()]
    return run


def test_every_split_stores_every_plane_of_every_tile_once(check):
    cases = []
    for ntile in (1, 4, 7, 110, 120, 130, 300):
        for nx in (1, 5, 40, 64, 512, 517):
            for W, occ in ((1, 1), (1, 2), (8, 2), (4, 2)):
                for T in (1, 2, 4):
                    for lx in march_chunks(occ) + (7, 600, spare_runs(occ)):
                        cases.append((ntile, nx, W, T, lx, 1024))
    res = check(cases)
    assert len(res) == len(cases)
    bad = [(c, r) for c, r in zip(cases, res) if r[3] != 1 or r[0] < 1]
    assert not bad, bad[:5]


def test_spare_split_shortens_the_longest_workgroup(check):
    """120 tiles (12 y-tiles x 10 z-tiles) of 512 planes, 8-wave workgroups at
    two waves per SIMD, T = 4, 1024 SIMDs: whole runs take two 256-plane runs
    per tile on 240 of the 256 resident workgroups (264 iterations each); the
    spare split cuts them to 244 planes and gives the 16 spare workgroups the
    last 24 planes of every tile (at most 252 iterations)."""
    (g0, c0, it0, ok0), (g1, c1, it1, ok1) = check([(120, 512, 8, 4, whole_runs(2), 1024),
                                                    (120, 512, 8, 4, spare_runs(2), 1024)])
    assert (g0, c0, it0, ok0) == (240, 256, 264, 1)
    assert ok1 == 1 and g1 == 256 and c1 < 0 and it1 < it0
    assert (c1, it1) == (-244, 252)


def test_balanced_and_plain_splits_are_unchanged(check):
    res = check([(120, 512, 8, 4, 256, 1024), (120, 512, 8, 4, balanced_runs(2), 1024),
                 (120, 512, 8, 4, -1, 1024)])
    assert res[0][:2] == (240, 256)
    assert res[1][:2] == (256, 240)
    assert res[2][:2] == (128, 512)
