"""The march kernel's variant generator (formura_amd/compiler/emit_hip.py):
every generated tile is one the kernel accepts and the launcher can tile,
and the budget rules hold -- checked on the host, no GPU."""
from __future__ import annotations

import pytest

from formura_amd.compiler.emit_hip import (AGPR_REG_BUDGET, MAX_EDGE_LDS, OCC2_REG_BUDGET, OCC4_REG_BUDGET,
                                           default_march_variants, edge_lds_bytes, march_regs)

CASES = [(1, "double", 512), (2, "double", 512), (1, "float", 512), (2, "float", 512),
         (1, "double", 48), (2, "double", 70), (1, "double", 256), (2, "double", None)]


@pytest.mark.parametrize("nf,real,ny", CASES)
def test_generated_tiles_are_valid(nf, real, ny):
    words = 2 if real == "double" else 1
    vs = default_march_variants(nf, real, ny=ny)
    assert len(vs) == len(set(vs)), "duplicate variants"
    budget = {1: AGPR_REG_BUDGET, 2: OCC2_REG_BUDGET, 4: OCC4_REG_BUDGET}
    for T, by, W, pf, occ in vs:
        rows = by * W
        # the kernel's static checks: a tile taller than its ghost rows,
        # prefetch depth 1 or 2, and at most 8 waves (512 threads) per workgroup
        assert rows > 2 * T and pf in (1, 2) and 1 <= W <= 8 and occ in (1, 2, 4)
        # the register model of the rows a wave keeps
        assert march_regs(T, by, nf, words, pf) <= budget[occ], (T, by, W, pf, occ)
        if W > 1:
            assert edge_lds_bytes(T, W, nf, words) <= MAX_EDGE_LDS
        if T == 1:
            # aligned 64-lane rows: the halo registers hold 32 rows each
            assert by <= 64


@pytest.mark.parametrize("nf,real,ny", [c for c in CASES if c[2]])
def test_exact_tiling_tiles_divide_the_y_extent(nf, real, ny):
    """Each T >= 2 that has a multi-wave tile storing at least 4T rows also
    has one whose stored rows divide ny, whenever such a tile fits."""
    vs = default_march_variants(nf, real, ny=ny)
    base = default_march_variants(nf, real)
    extra = [v for v in vs if v not in base]
    for T, by, W, pf, occ in extra:
        stored = W * by - 2 * T
        assert T >= 2 and stored >= 4 * T and ny % stored == 0, (T, by, W, pf, occ)
    # heat FP64 on a 512-row grid: the 4 x 10 tile stores 32 rows
    if (nf, real, ny) == (1, "double", 512):
        assert (4, 10, 4, 1, 1) in vs


def test_shallow_and_deep_fusion_are_offered():
    """FP64 offers T = 1..5 and 8 (the sweep over {1, 2, 4, 8}); T = 1 has
    a multi-wave tile at two waves per SIMD and T = 2..5 an 8-wave one."""
    vs = default_march_variants(1, "double", ny=512)
    assert sorted({v[0] for v in vs}) == [1, 2, 3, 4, 5, 8]
    assert any(v[0] == 1 and v[2] > 1 and v[4] == 2 for v in vs)
    for T in (2, 3, 4, 5):
        assert any(v[0] == T and v[2] == 8 and v[4] == 2 for v in vs)
