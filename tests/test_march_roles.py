"""The schedule of the march kernel's edge-row tiles (MarchRole in
formura_amd/csrc/fmr_march.hpp), restated on the host: every value a tile
stores is computed from valid inputs, and every stored (plane, row) of the
tile is stored exactly once -- checked on a validity model, no GPU.

Model: per wave, which (level, plane, row) values are valid.  Level 0 planes
are loaded from x0 - T on; an iteration with input plane xin computes level
L + 1 at plane xin - L - 1 on the wave's rows [lo, hi) of that level (the
first wave [L + 1, RB), the last [0, RB - L - 1), the others all rows); the
y-neighbours across a wave boundary come from the edge rows the neighbour
published at the start of the iteration (level l at plane xin - l - 1; the
first wave publishes only its last row, the last wave only its first); the
x-neighbour behind is the plane kept in the window or, with the x-carry, the
carry computed one plane earlier.  Level T is stored for planes [x0, xend).
"""
from __future__ import annotations

import itertools

import pytest

from formura_amd.compiler.emit_hip import march_edge_rows


def simulate(T, BY, W, EY, XC, x0=0, xend=12):
    NP = 2 if XC else 3
    roles = ["all"] * W if EY == 0 else ["top"] + ["inner"] * (W - 2) + ["bot"]
    RB = [BY + EY if r in ("top", "bot") else BY for r in roles]
    row0 = [0]
    for w in range(1, W):
        row0.append(row0[-1] + RB[w - 1])
    ROWS = sum(RB)
    valid = [dict() for _ in range(W)]
    carry = [dict() for _ in range(W)]

    def lo_hi(w, L):
        lo = L + 1 if (W == 1 or roles[w] == "top") else 0
        hi = RB[w] - L - 1 if (W == 1 or roles[w] == "bot") else RB[w]
        return lo, hi

    def get(w, L, p, r):
        return valid[w].get((L, p, r), False)

    iters = (xend - x0 + 2 * T + NP - 1) // NP * NP
    stored, bad = {}, []
    for s in range(iters):
        xin = x0 - T + s
        for w in range(W):
            for r in range(RB[w]):
                valid[w][(0, xin, r)] = True
        edge = {}
        for w in range(W):
            for l in range(T):
                if roles[w] != "top":
                    edge[(l, w, 0)] = get(w, l, xin - l - 1, 0)
                if roles[w] != "bot":
                    edge[(l, w, 1)] = get(w, l, xin - l - 1, RB[w] - 1)
        for L in range(T):
            q = xin - (L + 1)
            if L + 1 == T and not (x0 - (1 if XC else 0) <= q < xend):
                continue
            for w in range(W):
                lo, hi = lo_hi(w, L)
                for r in range(lo, hi):
                    ok = get(w, L, q, r) and get(w, L, q + 1, r)
                    ok = ok and (carry[w].get((L, q - 1, r), False) if XC else get(w, L, q - 1, r))
                    ok = ok and (get(w, L, q, r - 1) if r > 0 else (W == 1 or edge.get((L, (w - 1) % W, 1)) is True))
                    ok = ok and (get(w, L, q, r + 1) if r < RB[w] - 1
                                 else (W == 1 or edge.get((L, (w + 1) % W, 0)) is True))
                    if XC:
                        carry[w][(L, q, r)] = get(w, L, q, r) and get(w, L, q + 1, r)
                    if L + 1 < T:
                        valid[w][(L + 1, q, r)] = ok
                    else:
                        t = row0[w] + r
                        if q >= x0 and (roles[w] != "all" or T <= t < ROWS - T):
                            assert (q, t) not in stored, f"plane {q}, tile row {t} stored twice"
                            stored[(q, t)] = ok
                            if not ok:
                                bad.append((q, t))
    want = {(q, t) for q in range(x0, xend) for t in range(T, ROWS - T)}
    return bad, want - set(stored), set(stored) - want


CASES = [(T, BY, W, XC) for T, BY, W, XC in itertools.product((1, 2, 3, 4, 5), (3, 4, 6, 8), (2, 4, 8), (True, False))
         if BY * W > 2 * T]


@pytest.mark.parametrize("T,BY,W,XC", CASES)
def test_edge_row_tiles_store_every_row_once_from_valid_inputs(T, BY, W, XC):
    EY = march_edge_rows(T)
    if BY + EY <= T:
        pytest.skip("edge waves must hold more rows than T")
    for ey in (0, EY):
        bad, missing, extra = simulate(T, BY, W, ey, XC)
        assert not bad and not missing and not extra, (ey, bad[:3], sorted(missing)[:3], sorted(extra)[:3])


def test_edge_rows_balance_the_row_levels():
    """The edge waves' row-levels, T (BY + EY) - T (T + 1) / 2, match the
    inner waves' T BY to within T / 2."""
    for T in range(1, 9):
        EY = march_edge_rows(T)
        for BY in range(T, 12):
            edge = sum(min(r, T) for r in range(BY + EY))
            assert abs(edge - T * BY) <= T / 2, (T, BY, EY, edge)
