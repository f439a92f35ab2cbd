"""HIP backend: emits `<name>.hip`, a step library for one MI355X with the
reference's `Formura_Navigator` ABI, beside the reference's MPI C++ backend
(dispatch by output extension as in MPICxx/Language.hs:17-22 and
exe-src/formura.hs:81-84; `.hip` selects this emitter).

The per-site arithmetic is emitted as the body of `Prog::point`: one `const T
a_k = (x op y);` per op of the point IR, T being the C type the reference
declares for that node (MPICxx/Translate.hs:218-230, 481-629). Everything
else -- tiling, temporal blocking, periodic wrap, device memory, streams --
lives in the hand-written templates of formura_amd/csrc.

`Formura_Forward` (reference: MPICxx/Translate.hs:1106-1131) uploads the
interior of the host state arrays, advances monitorInterval2 steps on the GPU,
downloads the interior and adds monitorInterval2 to `time_step`; it returns 0
and aborts with a message on any HIP error.
"""
from __future__ import annotations

import heapq
from typing import Dict, List, Sequence, Tuple

from .abi import Layout, init_body
from .pointir import PointIR, Stage, c_expr_of, group_stages, imm_literal
from .shiftreuse import ShiftAnalysis, depends_on, op_children

# Registers (32-bit, per lane) of one march wavefront, as the compiler
# allocates them: per row and field, the two live planes of each of the T
# window levels, the prefetched planes and about two planes of temporaries;
# about 24 more for the rest (row offsets are scalar).  A wavefront alone on
# its SIMD may hold 256 VGPRs plus 256 AGPRs, but every value the compiler parks
# in an AGPR costs a move in and out, so the largest tiles stay a little above
# 256; within 256 registers two wavefronts share a SIMD.
# Variants that still spill to scratch memory are dropped by the build.
AGPR_REG_BUDGET = 340    # one wave per SIMD, some window values parked in AGPRs
# A multi-stage step whose inlined point function would exceed this many
# operations per site gets stage kernels only (3D MHD inlined: ~550 000).
INLINE_LIMIT = 6000
OCC2_REG_BUDGET = 236    # two waves per SIMD (256 registers each, some slack for the scheduler)
OCC4_REG_BUDGET = 112    # four waves per SIMD (128 registers each)
MAX_WAVES = 4            # <= 4 waves per workgroup keeps one wave per SIMD
MAX_EDGE_LDS = 96 * 1024   # LDS of a multi-wave tile's edge-row exchange (160 KB per CU)
MAX_FIELDS = 16            # state fields per program (fmr::MAX_FIELDS in csrc/fmr_stencil.hpp)


def edge_lds_bytes(T: int, W: int, nf: int, words: int) -> int:
    """LDS of the march kernel's edge exchange: [2][T][W][2][NF][64] values."""
    return 2 * T * W * 2 * nf * 64 * 4 * words


def march_regs(T: int, by: int, nf: int, words: int, pf: int = 1) -> int:
    return by * nf * words * (2 * T + pf + 2) + 24


def _max_rows(T: int, nf: int, words: int, pf: int, budget: int) -> int:
    by = 0
    while march_regs(T, by + 1, nf, words, pf) <= budget:
        by += 1
    return by


def default_march_variants(nf: int, real: str, max_t: int = None, ny: int = None):
    """(T, rows per wave, waves, prefetch depth, waves per SIMD) variants of
    the march kernel, for each fusion depth T:
      * one wave per tile (no LDS, no barrier) with as many rows as fit the
        AGPR-assisted budget, one wave per SIMD;
      * one wave per tile at two waves per SIMD, prefetching one and two
        planes ahead (and, for T <= 2, at four waves per SIMD);
      * a tile of up to 64 + 2T rows split over up to 4 waves that exchange
        their edge rows through LDS (one wave per SIMD; for T = 1, where the
        re-read ghost rows are the kernel's main excess traffic, at two waves
        per SIMD);
      * for T >= 2 also multi-wave tiles at two waves per SIMD: up to 4
        waves, and (T <= 5) 8 waves, a whole CU, with two rows per wave
        below the register limit;
      * given the grid's y extent ny, for T >= 2 also the multi-wave tiles
        (one and two waves per SIMD) whose stored rows (W * rows - 2T) are
        the largest divisor of ny that fits: no tile then computes and
        stores rows another tile owns."""
    words = 2 if real == "double" else 1
    if max_t is None:
        # FP64: 1..5, and 8 for the sweep over {1, 2, 4, 8} (only a
        # multi-wave tile fits the registers there)
        depths = list(range(1, 6)) + [8] if real == "double" else list(range(1, 9))
    else:
        depths = list(range(1, max_t + 1))
    out = []
    for T in depths:
        cand = []
        shapes = [(AGPR_REG_BUDGET, 1, 1), (OCC2_REG_BUDGET, 1, 2), (OCC2_REG_BUDGET, 2, 2)]
        if T <= 2:      # shallow fusion is latency-bound: more waves in flight
            shapes += [(OCC4_REG_BUDGET, 1, 4), (OCC4_REG_BUDGET, 2, 4)]
        for budget, pf, occ in shapes:
            by = min(_max_rows(T, nf, words, pf, budget), 64 + 2 * T)
            if by >= 2 * T + 2:
                cand.append((T, by, 1, pf, occ))
        if T == 1:
            by2 = _max_rows(T, nf, words, 1, OCC2_REG_BUDGET)
            rows = min(64 + 2 * T, by2 * MAX_WAVES)
            W = -(-rows // by2)
            if W > 1:
                cand.append((T, -(-rows // W), W, 1, 2))
        by_max = _max_rows(T, nf, words, 1, AGPR_REG_BUDGET)
        if T >= 2 and by_max * MAX_WAVES >= 2 * T + 2:
            rows = min(64 + 2 * T, by_max * MAX_WAVES)
            W = -(-rows // by_max)
            if W > 1:
                cand.append((T, -(-rows // W), W, 1, 1))
        by2 = _max_rows(T, nf, words, 1, OCC2_REG_BUDGET)
        if T >= 2 and MAX_WAVES * by2 - 2 * T >= 4 * T:
            # the same at two waves per SIMD: fewer rows, no AGPR traffic
            rows = min(64 + 2 * T, by2 * MAX_WAVES)
            W = -(-rows // by2)
            cand.append((T, -(-rows // W), W, 1, 2))
        by8 = by2 - 2     # register slack for the scheduler
        if 2 <= T <= 5 and 8 * by8 - 2 * T >= 4 * T and edge_lds_bytes(T, 8, nf, words) <= MAX_EDGE_LDS:
            # a whole workgroup of 8 waves per CU (two per SIMD): ghost rows
            # shared by twice as many stored rows
            cand.append((T, by8, 8, 1, 2))
        if T >= 2 and ny:
            # (stored rows, rows per wave, waves): stored rows divide ny and
            # are at least 4T (ghost rows at most a third of the tile)
            for budget_rows, occ in ((by_max, 1), (by2, 2)):
                fits = [(W * by - 2 * T, by, W) for W in range(2, MAX_WAVES + 1) for by in range(1, budget_rows + 1)
                        if W * by - 2 * T >= 4 * T and W * by <= 64 + 2 * T and ny % (W * by - 2 * T) == 0]
                if fits:
                    _, by, W = max(fits)
                    cand.append((T, by, W, 1, occ))
        for v in cand:
            if v not in out:
                out.append(v)
    return out


# kernel kinds (fmr::KernelKind in csrc/fmr_library.hpp)
KIND_MARCH = 1       # march3d_kernel
KIND_MARCHV1 = 3     # marchv_kernel, one z-site per lane
KIND_MARCHV2 = 4     # marchv_kernel, two z-sites per lane
KIND_MARCHV1E = 5    # marchv_kernel, one z-site per lane, edge rows on the first and last wave
KIND_MARCHV2E = 6    # marchv_kernel, two z-sites per lane, edge rows on the first and last wave
MARCH_KINDS = (KIND_MARCH, KIND_MARCHV1, KIND_MARCHV2, KIND_MARCHV1E, KIND_MARCHV2E)
def march_edge_rows(T: int) -> int:
    """Extra rows of the first and the last wave of an edge-row tile at fusion
    depth T (csrc/fmr_library.hpp::march_edge_rows): their row-levels then
    match the inner waves'."""
    return (T + 1) // 2


def march_vz(kind: int) -> int:
    """z-sites per lane of a marchv kind."""
    return 2 if kind in (KIND_MARCHV2, KIND_MARCHV2E) else 1


def march_rows(kind: int, T: int, by: int, waves: int) -> int:
    """Rows of a march tile (csrc/fmr_library.hpp::march_rows)."""
    return by * waves + (2 * march_edge_rows(T) if kind in (KIND_MARCHV1E, KIND_MARCHV2E) else 0)
MAX_MARCHV_LDS = 144 * 1024   # LDS of one workgroup on gfx950 (160 KB per CU)


MARCHV_REG_BUDGET = 250       # of 256 VGPRs (two waves per SIMD)


def marchv_regs_per_word(T: int, vz: int) -> float:
    """VGPRs marchv_kernel takes per z-site of a lane and 32-bit word of the
    state, as the compiler allocated them for heat and Gray-Scott at T = 3, 4
    (one site per lane, with the x-carry: 11.3 and 14.9; two sites, three
    window planes: 10.1 and 12.8); about 24 more for the rest."""
    return 3.6 * T + 0.5 if vz == 1 else 2.7 * T + 2.0


def marchv_lds_bytes(T: int, W: int, nf: int, words: int, vz: int) -> int:
    """LDS of marchv_kernel's edge exchange: [2][T][W][2][NF][VZ][64] values."""
    return 2 * T * W * 2 * nf * vz * 64 * 4 * words


def default_marchv_variants(nf: int, real: str, ny: int = None, nz: int = None):
    """(kind, T, rows per wave, waves, prefetch depth, waves per SIMD) of
    marchv_kernel for the fusion depths where it competes (FP64 T = 3..5;
    FP32 T = 3..6), at two waves per SIMD (256 VGPRs, no AGPR moves, a second
    wave to hide latency) in workgroups of 8 waves (a CU): one and two z-sites
    per lane (two only when nz is even), as many rows per wave as the
    register model allows (one site per lane: two rows fewer, as slack for
    the scheduler), also one row fewer than that (one site per lane:
    prefetching one or two planes; on heat 512^3 FP64 at T = 4, 6 rows ran
    5 % faster than 7, and at T = 3 8 rows 12 % faster than 9, in one
    session; two sites per lane at T = 3, 4 rows ran 948 GLUP/s against 923
    for 5), and the tile whose stored rows are the largest divisor of ny
    that fits."""
    words = 2 if real == "double" else 1
    depths = range(3, 6) if real == "double" else range(3, 7)
    out = []
    for T in depths:
        for vz, kind in ((1, KIND_MARCHV1), (2, KIND_MARCHV2)):
            if vz == 2 and nz is not None and nz % 2:
                continue
            sites = int((MARCHV_REG_BUDGET - 24) // (nf * words * marchv_regs_per_word(T, vz)))
            by8 = sites // vz
            W, by = 8, by8
            if by >= 1 and W * by - 2 * T >= 4 * T and marchv_lds_bytes(T, W, nf, words, vz) <= MAX_MARCHV_LDS:
                out.append((kind, T, by, W, 1, 2))
                if vz == 1 and W * (by - 1) - 2 * T >= 4 * T:
                    out += [(kind, T, by - 1, W, 1, 2), (kind, T, by - 1, W, 2, 2)]
                elif vz == 2 and W * (by - 1) - 2 * T >= 4 * T:
                    out.append((kind, T, by - 1, W, 1, 2))
                # the one-site tiles with edge rows on the first and the last
                # wave (the two-site ones, kind 6, gave wrong bits now and then
                # in the parity tests, cause not found: not generated)
                if vz == 1:
                    for b in (by, by - 1):
                        if b >= 1 and b + march_edge_rows(T) > T and W * b - 2 * T >= 4 * T:
                            out.append((KIND_MARCHV1E, T, b, W, 1, 2))
                            if b == by - 1:   # and prefetching two planes (+2 % at T = 4, within noise)
                                out.append((KIND_MARCHV1E, T, b, W, 2, 2))
            if ny and by8 >= 1 and marchv_lds_bytes(T, 8, nf, words, vz) <= MAX_MARCHV_LDS:
                fits = [by for by in range(1, by8 + 1) if 8 * by - 2 * T >= 4 * T and ny % (8 * by - 2 * T) == 0]
                if fits and (kind, T, max(fits), 8, 1, 2) not in out:
                    out.append((kind, T, max(fits), 8, 1, 2))
    if nz == 64:
        out += [v for v in whole_line_marchv_variants(nf, real) if v not in out]
    return out


def whole_line_marchv_variants(nf: int, real: str):
    """Small tiles of marchv_kernel for grids whose z-lines are one wavefront
    long (nz = 64, one site per lane): a wavefront row then holds a whole
    periodic z-line (the lane moves rotate), so the tile has no ghost lanes,
    and small grids such as 64^3 need many short tiles to fill the chip.  T = 1
    to 3 in tiles of 1, 2 or 4 waves whose stored rows (W * rows - 2T) are 2 to
    16, at two waves per SIMD, within the register model."""
    words = 2 if real == "double" else 1
    out = []
    for T, tiles in ((1, ((6, 1), (10, 1))), (2, ((6, 1), (4, 2), (5, 4))), (3, ((4, 2), (4, 4)))):
        sites = int((MARCHV_REG_BUDGET - 24) // (nf * words * marchv_regs_per_word(T, 1)))
        for by, W in tiles:
            if by <= sites and marchv_lds_bytes(T, W, nf, words, 1) <= MAX_MARCHV_LDS:
                out.append((KIND_MARCHV1, T, by, W, 1, 2))
                if T == 2 and W == 1:
                    out.append((KIND_MARCHV1, T, by, W, 2, 2))
    return out


def _march_ny(lay: Layout):
    """The y extent the march kernel tiles (3D programs), else None."""
    return lay.intra[1] if len(lay.intra) == 3 else None


def _march_nz(lay: Layout):
    return lay.intra[2] if len(lay.intra) == 3 else None


def device_real(lay: Layout, precision: str) -> str:
    return "float" if precision == "fp32" else lay.field_ctype[0]


def variants_for(lay: Layout, precision: str):
    """The march variants emit_hip compiles in by default, as (kind, T, rows,
    waves, pf, occ): march3d_kernel's -- every T = 1 and T = 2 tile (T = 1:
    aligned 64-lane rows, only this kernel has them; both: small cache-
    resident grids such as 64^3, where the latency-hiding single-wave tiles
    win), and for T >= 3 the tiles at two or more waves per SIMD prefetching
    one plane (the one-wave-per-SIMD tiles park window values in AGPRs, and a
    second prefetched plane takes registers; neither won a timed search at
    512^3), except where only a one-wave-per-SIMD tile fits (FP64 T = 8) --
    and marchv_kernel's."""
    nf, real = len(lay.field_names), device_real(lay, precision)
    base = default_march_variants(nf, real, ny=_march_ny(lay))
    only_occ1 = {v[0] for v in base} - {v[0] for v in base if v[4] >= 2}
    old = [(KIND_MARCH,) + tuple(v) for v in base
           if v[0] <= 2 or (v[4] >= 2 and v[3] == 1) or v[0] in only_occ1]
    return old + default_marchv_variants(nf, real, ny=_march_ny(lay), nz=_march_nz(lay))


def _with_kind(v) -> tuple:
    v = tuple(v)
    return v if len(v) == 6 else (KIND_MARCH,) + v


def lane_reuse(ir: PointIR, real: str):
    """Lane-axis shift reuse: ({B: A}, op order).  B(s) = A(s - e) along the
    contiguous axis (the march kernel's lanes); the order is the op order with
    each such A moved before its B (a stable topological sort)."""
    sa = ShiftAnalysis(ir)
    lane = {b: a for b, a in sorted(sa.pairs(ir.dim - 1).items())
            if ir.ops[b].ctype == real and not depends_on(ir, a, b)}
    n = len(ir.ops)
    deps = [set(op_children(op)) for op in ir.ops]
    for b, a in lane.items():
        deps[b].add(a)
    users: List[List[int]] = [[] for _ in range(n)]
    for i, ds in enumerate(deps):
        for d in ds:
            users[d].append(i)
    left = [len(ds) for ds in deps]
    ready = [i for i in range(n) if left[i] == 0]
    heapq.heapify(ready)
    order: List[int] = []
    while ready:
        i = heapq.heappop(ready)
        order.append(i)
        for u in users[i]:
            left[u] -= 1
            if left[u] == 0:
                heapq.heappush(ready, u)
    return lane, order


def point_body(ir: PointIR, real: str) -> str:
    """The per-site statements as C++ for `P::point(a)`: one `const` per op, in
    op order except that a lane-reuse pair's A precedes its B; B is written as
    `a.lane_prev_of(A, [&]{ return B; })` (the kernel returns the previous
    lane's A, or evaluates B)."""
    lane, order = lane_reuse(ir, real)
    lines: List[str] = []
    names: Dict[int, str] = {}
    for i in order:
        op = ir.ops[i]
        v = f"a{i}"
        k = op.kind
        if k == "load":
            f, off = op.args
            o = list(off) + [0] * (3 - len(off))
            lines.append(f"const {real} {v} = a.template ld<{f}, {o[0]}, {o[1]}, {o[2]}>();")
        elif k == "imm":
            names[i] = imm_literal(op.args[0], op.ctype)
            continue
        elif k == "param":
            lines.append(f"const {op.ctype} {v} = ({op.ctype})a.param({op.args[0]});")
        elif k == "index":
            lines.append(f"const {op.ctype} {v} = ({op.ctype})a.index({op.args[0]});")
        else:
            e = c_expr_of(op, lambda j: names[j])
            if i in lane:
                e = f"a.lane_prev_of({names[lane[i]]}, [&]() -> {op.ctype} {{ return {e}; }})"
            lines.append(f"const {op.ctype} {v} = {e};")
        names[i] = v
    for f, o in enumerate(ir.outputs):
        lines.append(f"a.template st<{f}>(({real}){names[o]});")
    return "\n    ".join(lines)


def is_star(ir: PointIR) -> bool:
    for _, op in ir.loads():
        if sum(1 for c in op.args[1] if c != 0) > 1:
            return False
    return True


def emit_hip(lay: Layout, ir: PointIR, precision: str = "faithful",
             march_variants: Sequence[Tuple[int, int]] = None, stages: Sequence[Stage] = (),
             extern_header: str = None) -> str:
    """HIP library source of one program: the inlined point function (direct
    and march kernels) and, for a multi-stage program, one stage kernel per
    manifest node (`stages`, from lower_stages)."""
    dim = lay.dim
    if dim not in (2, 3):
        raise ValueError("the HIP backend supports 2D and 3D programs")
    real = device_real(lay, precision)
    if len(lay.field_names) > MAX_FIELDS:
        raise ValueError(f"{len(lay.field_names)} state fields: the HIP library holds at most {MAX_FIELDS}")
    if precision != "fp32" and len(set(lay.field_ctype)) != 1:
        raise ValueError("all state variables must share one element type")
    r = ir.radius() + [0] * (3 - dim)
    star = is_star(ir)
    stages = [st for st in stages] if stages and any(st.target == "temp" for st in stages) else []
    if stages and any(t != real for t in stages[0].ir.field_ctypes):
        stages = []       # stage arrays of another type than the state: inlined kernels only
    has_stages = bool(stages)
    inlined = len(ir.ops) <= INLINE_LIMIT or not has_stages
    use_march = dim == 3 and star and max(r) <= 1 and inlined
    if march_variants is None:
        march_variants = variants_for(lay, precision)
    variants = [_with_kind(v) for v in march_variants] if use_march else []
    if any(v[0] != KIND_MARCH for v in variants):
        # the x-carry with one z-site per lane; with two, the plane behind is
        # kept (measured: the carry raised the register peak there)
        mbody = {1: march_point_body(ir, real, 1), 2: march_point_body(ir, real, 2, use_carry=False)}
        nc = mbody[1][1]
    n = list(lay.intra) + [1] * (3 - dim)
    shape = lay.shape() + [1] * (3 - dim)
    margin = lay.margin + [0] * (3 - dim)
    dims_decl = "".join(f"[{s}]" for s in lay.shape())
    host_t = lay.field_ctype

    out: List[str] = []
    w = out.append
    w(f"// {lay.name}.hip -- generated by formura_amd (HIP backend). Do not edit.")
    w(f"// precision: {precision}; device element type: {real}")
    w('#include "fmr_library.hpp"')
    w(f'#include "{lay.name}.h"')
    if extern_header:
        w(f'#include "{extern_header}"   // the program\'s extern functions (FMR_HD: host and device)')
    w("")
    w("namespace {")
    w("struct Prog {")
    w(f"  using real = {real};")
    w(f"  static constexpr int DIM = {dim};")
    w(f"  static constexpr int NF = {len(lay.field_names)};")
    w(f"  static constexpr int RX = {r[0]}, RY = {r[1]}, RZ = {r[2]};")
    w(f"  static constexpr bool STAR = {'true' if star else 'false'};")
    w(f"  static constexpr bool INLINED = {'true' if inlined else 'false'};")
    if inlined:
        w("  template <class A>")
        w("  __device__ __forceinline__ static void point(A& a) {")
        w("    " + point_body(ir, real))
        w("  }")
    if any(v[0] != KIND_MARCH for v in variants):
        w(f"  // marchv_kernel, one z-site per lane: {nc} value(s) per site carried from the previous plane"
          + ("" if nc else " (none: the plane behind stays in the window)") + "; two z-sites per lane: none")
        w(f"  static constexpr int NC = {nc};")
        w(f"  static constexpr bool XC = {'true' if nc else 'false'};")
        w("  template <int VZ, class A>")
        w("  __device__ __forceinline__ static void mpoint(A& a) {")
        w("    if constexpr (VZ == 1) {")
        w("      " + mbody[1][0])
        w("    } else {")
        w("      " + mbody[2][0])
        w("    }")
        w("  }")
    else:
        w(f"  // the step inlined into one point function would take {len(ir.ops)} operations per")
        w("  // site (every manifest node recomputed inside its users): stage kernels only")
    w("};")
    w("")
    ntemps = sum(1 for st in stages if st.target == "temp")
    groups = group_stages(stages) if stages else []
    for k, g in enumerate(groups):
        for st in g.members:
            dest = "state field " + lay.field_names[st.index] if st.target == "state" else "stage array " + str(st.index)
            w(f"// manifest node {st.om_id} -> {dest}")
        w(f"struct Stage{k} {{")
        w(f"  using real = {real};")
        w(f"  static constexpr int DIM = {dim};")
        w(f"  static constexpr int NOUT = {len(g.members)};")
        w("  template <class A>")
        w("  __device__ __forceinline__ static void point(A& a) {")
        w("    " + point_body(g.ir, real))
        w("  }")
        w("};")
        w("")
    if groups:
        w("const fmr::StageEntry<Prog> kStages[] = {")
        for k, g in enumerate(groups):
            ts = ", ".join("1" if st.target == "state" else "0" for st in g.members)
            ix = ", ".join(str(st.index) for st in g.members)
            w(f"  {{&fmr::launch_stage<Prog, Stage{k}>, {len(g.members)}, {{{ts}}}, {{{ix}}}}},")
        w("};")
        w("")
    w("const fmr::VariantEntry<Prog> kVariants[] = {")
    for kind, T, by, wv, pf, occ in variants:
        if kind == KIND_MARCH:
            w(f"  {{{{fmr::KIND_MARCH, {T}, {by}, {wv}, {pf}, {occ}}}, "
              f"&fmr::launch_march3d<Prog, {T}, {by}, {wv}, {pf}, {occ}>}},")
        else:
            vz = march_vz(kind)
            ey = march_edge_rows(T) if kind in (KIND_MARCHV1E, KIND_MARCHV2E) else 0
            w(f"  {{{{{kind}, {T}, {by}, {wv}, {pf}, {occ}}}, "
              f"&fmr::launch_marchv<Prog, {T}, {by}, {wv}, {vz}, {pf}, {occ}, {ey}>}},")
    if not variants:
        w("  {{fmr::KIND_DIRECT, 1, 0, 1, 1, 1}, nullptr},")
    w("};")
    nv = len(variants)
    w(f"constexpr int kNumVariants = {nv};")
    w("fmr::StepLibrary<Prog> lib;")
    w(f"const int kShape[3] = {{{shape[0]}, {shape[1]}, {shape[2]}}};")
    w(f"const int kOrigin[3] = {{{margin[0]}, {margin[1]}, {margin[2]}}};")
    w("}  // namespace")
    w("")
    for f, t in zip(lay.field_names, host_t):
        w(f"{t} {f}{dims_decl};")
    w("struct Formura_Navigator navi;")
    w("")
    w('extern "C" {')
    w("")
    w("int Formura_Init (struct Formura_Navigator *navi, MPI_Comm comm)")
    w("{")
    w("  " + init_body(lay))
    w(f"  lib.init({n[0]}, {n[1]}, {n[2]}, kVariants, kNumVariants);")
    if stages:
        w(f"  lib.init_stages(kStages, {len(groups)}, {ntemps});")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Upload (struct Formura_Navigator *navi)")
    w("{")
    w("  (void)navi;")
    for i, f in enumerate(lay.field_names):
        w(f"  lib.upload({i}, &{f}{''.join('[0]' for _ in range(dim))}, kShape, kOrigin);")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Download (struct Formura_Navigator *navi)")
    w("{")
    w("  (void)navi;")
    w("  lib.sync();")
    for i, f in enumerate(lay.field_names):
        w(f"  lib.download({i}, &{f}{''.join('[0]' for _ in range(dim))}, kShape, kOrigin);")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Forward_Device (struct Formura_Navigator *navi, int nsteps)")
    w("{")
    w("  if (nsteps < 0) { fprintf(stderr, \"formura_amd: negative step count\\n\"); abort(); }")
    w("  lib.advance(nsteps);")
    w("  lib.sync();")
    w("  navi->time_step += nsteps;")
    w("  return 0;")
    w("}")
    w("")
    w("float Formura_Forward_Device_Timed (struct Formura_Navigator *navi, int nsteps)")
    w("{")
    w("  if (nsteps < 0) { fprintf(stderr, \"formura_amd: negative step count\\n\"); abort(); }")
    w("  const float ms = lib.timed_advance(nsteps);")
    w("  navi->time_step += nsteps;")
    w("  return ms;")
    w("}")
    w("")
    w("int Formura_Last_Timing (float *main_ms, int *main_launches, float *total_ms)")
    w("{")
    w("  *main_ms = lib.last_main_ms; *main_launches = lib.last_main_launches; *total_ms = lib.last_total_ms;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Forward (struct Formura_Navigator *navi)")
    w("{")
    w("  Formura_Upload(navi);")
    w(f"  lib.advance({lay.steps_per_forward});")
    w("  Formura_Download(navi);")
    w(f"  navi->time_step += {lay.steps_per_forward};")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Set_Kernel (int kind, int T, int rows, int waves, int pf, int occ)")
    w("{")
    w("  if (kind == fmr::KIND_DIRECT) {")
    w("    if (!Prog::INLINED) return -1;")
    w("    lib.choice = fmr::Variant{fmr::KIND_DIRECT, 1, 0, 1, 1, 1};")
    w("    return 0;")
    w("  }")
    w("  if (kind == fmr::KIND_STAGED) {")
    w("    if (lib.nstages == 0 || (rows != 0 && rows != 4 && rows != 16)) return -1;")
    w("    lib.choice = fmr::Variant{fmr::KIND_STAGED, 1, rows, 1, 1, 1};")
    w("    return 0;")
    w("  }")
    w("  if (!lib.find(kind, T, rows, waves, pf, occ)) return -1;")
    w("  lib.choice = fmr::Variant{kind, T, rows, waves, pf, occ};")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Num_Stages (void)")
    w("{")
    w("  return lib.nstages;")
    w("}")
    w("")
    w("int Formura_Num_Stage_Kernels (void)")
    w("{")
    w("  return lib.nkernels;")
    w("}")
    w("")
    w("int Formura_Set_March_Chunk (int lx)")
    w("{")
    w("  if (lx == 0) return -1;")
    w("  lib.lx = lx;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Get_Kernel (int *kind, int *T, int *rows, int *waves, int *pf, int *occ, int *lx)")
    w("{")
    w("  *kind = lib.choice.kind; *T = lib.choice.T; *rows = lib.choice.by; *waves = lib.choice.waves;")
    w("  *pf = lib.choice.pf; *occ = lib.choice.occ;")
    w("  *lx = lib.lx;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_List_Kernels (int *out, int max_variants)")
    w("{")
    w("  int k = 0;")
    w("  for (int i = 0; i < kNumVariants && k < max_variants; ++i, ++k) {")
    w("    out[6 * k] = kVariants[i].v.kind; out[6 * k + 1] = kVariants[i].v.T; out[6 * k + 2] = kVariants[i].v.by;")
    w("    out[6 * k + 3] = kVariants[i].v.waves; out[6 * k + 4] = kVariants[i].v.pf; out[6 * k + 5] = kVariants[i].v.occ;")
    w("  }")
    w("  return k;")
    w("}")
    w("")
    w("int Formura_Set_Graphs (int on)")
    w("{")
    w("  lib.sync();")
    w("  lib.use_graphs = on != 0;")
    w("  if (!lib.use_graphs) lib.drop_graphs();")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Prepare_Device (int nsteps)")
    w("{")
    w("  if (nsteps < 0) return -1;")
    w("  lib.prepare(nsteps);")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Set_Params (const double *p, int np)")
    w("{")
    w(f"  if (np < 0 || np > fmr::MAX_PARAMS) return -1;")
    w("  lib.sync();")
    w("  lib.drop_graphs();   // captured launches hold the parameters by value")
    w("  for (int i = 0; i < np; ++i) lib.prm.v[i] = p[i];")
    w("  lib.nparams = np;")
    w("  return 0;")
    w("}")
    w("")
    w("void *Formura_Device_Field (int f)")
    w("{")
    w(f"  if (f < 0 || f >= {len(lay.field_names)}) return nullptr;")
    w("  return lib.buf[lib.cur].f[f];")
    w("}")
    w("")
    w("int Formura_Device_Elem_Size (void) { return (int)sizeof(Prog::real); }")
    w("")
    w("int Formura_Reduce_Device (int f, double *res)")
    w("{")
    w(f"  if (f < 0 || f >= {len(lay.field_names)}) return -1;")
    w("  lib.reduce(f, res);")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Diff_Device (int f, const void *ref, int ref_elem_size, double *res)")
    w("{")
    w(f"  if (f < 0 || f >= {len(lay.field_names)} || !ref) return -1;")
    w("  if (ref_elem_size == 8) lib.diff(f, (const double *)ref, res);")
    w("  else if (ref_elem_size == 4) lib.diff(f, (const float *)ref, res);")
    w("  else return -1;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Diff_Arrays_NoAlloc (const void *a, int a_elem_size, const void *ref, int ref_elem_size,")
    w("                                 long long n, double *res, void *stream)")
    w("{")
    w("  return fmr::diff_noalloc_entry(a, a_elem_size, ref, ref_elem_size, n, res,")
    w("                                 stream ? (hipStream_t)stream : lib.stream);")
    w("}")
    w("")
    w('}  // extern "C"')
    return "\n".join(out) + "\n"


def x_carry(ir: PointIR, real: str):
    """x-carry plan of the march kernel: ({B: A}, live ops) where op B at
    site s equals op A at s - e_x (so the march takes B from the value A had
    at the previous plane), or ({}, live ops) when some load of the plane
    behind (x offset -1) would stay live -- the window then keeps that plane
    and nothing is carried.  Lane-axis reuse pairs (lane_reuse) are applied
    first; the live set follows both substitutions."""
    sa = ShiftAnalysis(ir)
    lane, _ = lane_reuse(ir, real)
    xp = {b: a for b, a in sorted(sa.pairs(0).items())
          if b not in lane and ir.ops[b].ctype == real and not depends_on(ir, a, b)}

    live = _live_without_carry(ir, lane, xp)
    carry = {b: a for b, a in xp.items() if b in live}
    if not carry or any(ir.ops[i].kind == "load" and ir.ops[i].args[1][0] == -1 for i in live):
        return {}, _live_without_carry(ir, lane)
    return carry, live


def march_point_body(ir: PointIR, real: str, vz: int, use_carry: bool = True) -> Tuple[str, int]:
    """The per-site statements for `vz` consecutive z-sites of a lane, as the
    body of `P::mpoint<vz>(a)` of the march kernel of csrc/fmr_march.hpp, and
    the number of carried values per site.  Op i at site v is `a{i}_{v}`;
    ops are emitted in op order (each lane-reuse A before its B), all sites of
    one op together, so that every operand is defined before use:
      * a lane-reuse B at site v > 0 is A at site v - 1 (same register); at
        site 0 it is A at the previous lane's last site (`a.lane_prev`);
      * an x-carried B is the carried value of A from the previous plane,
        read at the top (`a.xprev<k, v>()`); A's new value is handed back
        right after it is computed (`a.xkeep<k, v>`)."""
    lane, order = lane_reuse(ir, real)
    if use_carry:
        carry, live = x_carry(ir, real)
    else:
        carry, live = {}, _live_without_carry(ir, lane)
    kidx = {b: k for k, b in enumerate(sorted(carry))}
    a_of = {carry[b]: kidx[b] for b in carry}
    lines: List[str] = []
    names: Dict[Tuple[int, int], str] = {}
    for b in sorted(carry):
        for v in range(vz):
            nm = f"x{kidx[b]}_{v}"
            lines.append(f"const {ir.ops[b].ctype} {nm} = a.template xprev<{kidx[b]}, {v}>();")
            names[(b, v)] = nm
    for i in order:
        if i not in live:
            continue
        op = ir.ops[i]
        k = op.kind
        for v in range(vz):
            nm = f"a{i}_{v}"
            if i in carry:
                continue                      # named by its xprev read
            if k == "load":
                f, off = op.args
                o = list(off) + [0] * (3 - len(off))
                lines.append(f"const {real} {nm} = a.template ld<{f}, {o[0]}, {o[1]}, {o[2]}, {v}>();")
            elif k == "imm":
                names[(i, v)] = imm_literal(op.args[0], op.ctype)
                continue
            elif k == "param":
                lines.append(f"const {op.ctype} {nm} = ({op.ctype})a.param({op.args[0]});")
            elif k == "index":
                lines.append(f"const {op.ctype} {nm} = ({op.ctype})a.template index<{op.args[0]}, {v}>();")
            elif i in lane:
                src = lane[i]
                if v > 0:
                    names[(i, v)] = names[(src, v - 1)]
                    continue
                lines.append(f"const {op.ctype} {nm} = a.lane_prev({names[(src, vz - 1)]});")
            else:
                lines.append(f"const {op.ctype} {nm} = {c_expr_of(op, lambda j: names[(j, v)])};")
            names[(i, v)] = nm
            if i in a_of:
                lines.append(f"a.template xkeep<{a_of[i]}, {v}>({nm});")
    for v in range(vz):
        for f, o in enumerate(ir.outputs):
            lines.append(f"a.template st<{f}, {v}>(({real}){names[(o, v)]});")
    return "\n      ".join(lines), len(carry)


def _live_without_carry(ir: PointIR, lane, carry=None) -> set:
    """Ops the outputs need when each lane-reuse B (and each carried B) is
    replaced by its A."""
    carry = carry or {}
    live, stack = set(), list(ir.outputs)
    while stack:
        i = stack.pop()
        if i in live:
            continue
        live.add(i)
        if i in carry:
            stack.append(carry[i])
        elif i in lane:
            stack.append(lane[i])
        else:
            stack.extend(op_children(ir.ops[i]))
    return live
