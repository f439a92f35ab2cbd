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

from typing import Dict, List, Sequence, Tuple

from .abi import Layout, init_body
from .pointir import PointIR, c_expr_of, imm_literal

# Register budget (32-bit registers per lane) of one march wavefront: the plane
# windows (T levels x 3 planes), the prefetched plane and the row offsets.  A
# wavefront at one wave per SIMD owns 256 VGPRs plus 256 AGPRs; the compiler
# parks the overflow of the windows in AGPRs, and beyond this budget it would
# spill to scratch memory.
MARCH_REG_BUDGET = 340
MAX_WAVES = 4        # <= 4 waves per workgroup keeps one wave per SIMD (512 registers)
OCC2_REG_BUDGET = 200   # two waves per SIMD: 256 registers per wave


def march_regs(T: int, by: int, nf: int, words: int, pf: int = 1) -> int:
    return (3 * T + pf) * by * nf * words + by


def default_march_variants(nf: int, real: str, max_t: int = None):
    """(T, rows per wave, waves, prefetch depth) variants of the march kernel.

    For each fusion depth T: a tile of up to 64 + 2T rows (64 valid rows),
    split over the fewest waves (<= 4) whose registers fit the budget; and
    single-wave tiles (no LDS, no barrier) with as many rows as one wave can
    hold, prefetching one or two planes ahead."""
    words = 2 if real == "double" else 1
    if max_t is None:
        max_t = 4 if real == "double" else 8
    out = []
    for T in range(1, max_t + 1):
        by_max = 1
        while march_regs(T, by_max + 1, nf, words) <= MARCH_REG_BUDGET:
            by_max += 1
        if by_max * MAX_WAVES < 2 * T + 2:
            continue
        rows = min(64 + 2 * T, by_max * MAX_WAVES)
        W = -(-rows // by_max)
        by = -(-rows // W)
        if W > 1:
            out.append((T, by, W, 1, 1))
        by1 = min(by_max, 64 + 2 * T)
        if by1 >= 2 * T + 2:
            out.append((T, by1, 1, 1, 1))
        by2 = 1
        while march_regs(T, by2 + 1, nf, words, 2) <= MARCH_REG_BUDGET:
            by2 += 1
        if by2 >= 2 * T + 2:
            out.append((T, min(by2, 64 + 2 * T), 1, 2, 1))
        # two waves per SIMD: half the registers per wave, twice the waves
        # to hide memory and arithmetic latency
        by3 = 1
        while march_regs(T, by3 + 1, nf, words, 2) <= OCC2_REG_BUDGET:
            by3 += 1
        if by3 >= 2 * T + 4:
            out.append((T, by3, 1, 2, 2))
    return out


def point_body(ir: PointIR, real: str) -> str:
    lines: List[str] = []
    names: Dict[int, str] = {}
    for i, op in enumerate(ir.ops):
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
            lines.append(f"const {op.ctype} {v} = {c_expr_of(op, lambda j: names[j])};")
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
             march_variants: Sequence[Tuple[int, int]] = None) -> str:
    dim = lay.dim
    if dim not in (2, 3):
        raise ValueError("the HIP backend supports 2D and 3D programs")
    real = "float" if precision == "fp32" else lay.field_ctype[0]
    if precision != "fp32" and len(set(lay.field_ctype)) != 1:
        raise ValueError("all state variables must share one element type")
    r = ir.radius() + [0] * (3 - dim)
    star = is_star(ir)
    use_march = dim == 3 and star and max(r) <= 1
    if march_variants is None:
        march_variants = default_march_variants(len(lay.field_names), real)
    variants = list(march_variants) if use_march else []
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
    w("")
    w("namespace {")
    w("struct Prog {")
    w(f"  using real = {real};")
    w(f"  static constexpr int DIM = {dim};")
    w(f"  static constexpr int NF = {len(lay.field_names)};")
    w(f"  static constexpr int RX = {r[0]}, RY = {r[1]}, RZ = {r[2]};")
    w(f"  static constexpr bool STAR = {'true' if star else 'false'};")
    w("  template <class A>")
    w("  __device__ __forceinline__ static void point(A& a) {")
    w("    " + point_body(ir, real))
    w("  }")
    w("};")
    w("")
    w("const fmr::VariantEntry<Prog> kVariants[] = {")
    for T, by, wv, pf, occ in variants:
        w(f"  {{{{fmr::KIND_MARCH, {T}, {by}, {wv}, {pf}, {occ}}}, "
          f"&fmr::launch_march3d<Prog, {T}, {by}, {wv}, {pf}, {occ}>}},")
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
    w("  if (kind == fmr::KIND_DIRECT) { lib.choice = fmr::Variant{fmr::KIND_DIRECT, 1, 0, 1, 1, 1}; return 0; }")
    w("  if (!lib.find(kind, T, rows, waves, pf, occ)) return -1;")
    w("  lib.choice = fmr::Variant{kind, T, rows, waves, pf, occ};")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Set_March_Chunk (int lx)")
    w("{")
    w("  if (lx < 1) return -1;")
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
    w("    out[5 * k] = kVariants[i].v.T; out[5 * k + 1] = kVariants[i].v.by; out[5 * k + 2] = kVariants[i].v.waves;")
    w("    out[5 * k + 3] = kVariants[i].v.pf; out[5 * k + 4] = kVariants[i].v.occ;")
    w("  }")
    w("  return k;")
    w("}")
    w("")
    w("int Formura_Set_Params (const double *p, int np)")
    w("{")
    w(f"  if (np < 0 || np > fmr::MAX_PARAMS) return -1;")
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
    w('}  // extern "C"')
    return "\n".join(out) + "\n"
