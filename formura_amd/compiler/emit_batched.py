"""HIP backend for batched launches of a 2D program: many independent
instances, each with its own grid size and its own values of the program's
named scalar parameters (e.g. F, k, dt), advanced together.

The per-site arithmetic is the same `Prog::point` the single-instance backend
emits (formura_amd/compiler/emit_hip.py), with the parameter nodes read from
the instance's parameter vector. The kernels and the host runtime live in
formura_amd/csrc/fmr_batched.hpp. The entry points (declared in
include/formura_amd_batch.h) have no counterpart in the reference, which
generates one fixed-size program per compile.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

from .abi import Layout
from .emit_hip import MAX_FIELDS, is_star, point_body
from .pointir import PointIR

# (T, points per lane) variants; every T has variants for every lane width so
# that a remainder of fewer than T steps keeps the tile shape
BATCH_T = (1, 2, 3, 4, 6, 8)
BATCH_VY = (1, 2, 4)


FUSED_T = (1, 2, 4, 8)


def fused_kernels(variants: Sequence[Tuple[int, int]]) -> List[Tuple[int, int]]:
    """(T, lane-width mask) of the fused kernels: for each depth T of
    FUSED_T, every set of one or two of the lane widths compiled at T (bit 1:
    one point per lane, 2: two, 4: four), so that a plan of one depth with
    at most two lane widths over its size buckets runs as one launch per
    T-block (a remainder of fewer than T steps: 4 + 2 + 1 ...)."""
    out = []
    for T in sorted({t for t, _ in variants} & set(FUSED_T)):
        bits = 0
        for t, vy in variants:
            if t == T:
                bits |= {1: 1, 2: 2, 4: 4}[vy]
        for mask in range(1, 8):
            if mask & bits == mask and bin(mask).count("1") <= 2:
                out.append((T, mask))
    return out


def default_batch_variants(nf: int, real: str) -> List[Tuple[int, int]]:
    words = 2 if real == "double" else 1
    out = []
    for vy in BATCH_VY:
        for T in range(1, max(BATCH_T) + 1):
            if T not in BATCH_T and T not in (5, 7):
                continue
            if (3 * T + 1) * vy * nf * words <= 200 and 64 * vy > 2 * T + 1:
                out.append((T, vy))
    return out


def emit_batched(lay: Layout, ir: PointIR, precision: str = "fp32",
                 variants: Sequence[Tuple[int, int]] = None) -> str:
    if lay.dim != 2:
        raise ValueError("batched launches are implemented for 2D programs")
    if not is_star(ir) or max(ir.radius()) > 1:
        raise ValueError("batched launches need a radius-1 star stencil")
    real = "float" if precision == "fp32" else lay.field_ctype[0]
    nf = len(lay.field_names)
    if nf > MAX_FIELDS:
        raise ValueError(f"{nf} state fields: the batched library holds at most {MAX_FIELDS}")
    if variants is None:
        variants = default_batch_variants(nf, real)
    out: List[str] = []
    w = out.append
    w(f"// {lay.name}.hip -- batched step library generated by formura_amd. Do not edit.")
    w(f"// precision: {precision}; device element type: {real}; parameters: {', '.join(ir.params)}")
    w('#include "fmr_batched.hpp"')
    w("")
    w("namespace {")
    w("struct Prog {")
    w(f"  using real = {real};")
    w("  static constexpr int DIM = 2;")
    w(f"  static constexpr int NF = {nf};")
    w("  template <class A>")
    w("  __device__ __forceinline__ static void point(A& a) {")
    w("    " + point_body(ir, real))
    w("  }")
    w("};")
    w("const fmr::BatchVariantEntry<Prog> kVariants[] = {")
    for T, vy in variants:
        w(f"  {{{{{T}, {vy}}}, &fmr::launch_batched<Prog, {T}, {vy}>}},")
    w("};")
    w(f"constexpr int kNumVariants = {len(variants)};")
    fused = fused_kernels(variants)
    w("// fused kernels: depth T, lane widths of the mask (1: one point per lane, 2: two, 4: four)")
    w("const fmr::FusedEntry<Prog> kFused[] = {")
    for T, mask in fused:
        w(f"  {{{T}, {mask}, &fmr::launch_fused<Prog, {T}, {mask}>}},")
    w("};")
    w(f"constexpr int kNumFused = {len(fused)};")
    w("fmr::BatchLibrary<Prog> lib;")
    w("bool ready = false;")
    w("}  // namespace")
    w("")
    w('extern "C" {')
    w("")
    w(f"int Formura_Batch_Num_Fields (void) {{ return {nf}; }}")
    w(f"int Formura_Batch_Num_Params (void) {{ return {len(ir.params)}; }}")
    w("int Formura_Batch_Elem_Size (void) { return (int)sizeof(Prog::real); }")
    w("")
    w("int Formura_Batch_Init (int n, const int *nx, const int *ny, const double *params, int np)")
    w("{")
    w(f"  if (ready || n < 1 || np != {len(ir.params)}) return -1;")
    w("  for (int i = 0; i < n; ++i) if (nx[i] < 1 || ny[i] < 1) return -1;")
    w("  lib.init(n, nx, ny, params, np, kVariants, kNumVariants, kFused, kNumFused);")
    w("  ready = true;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Batch_Destroy (void)")
    w("{")
    w("  if (!ready) return -1;")
    w("  lib.destroy();")
    w("  ready = false;")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Batch_List_Variants (int *out, int max_variants)")
    w("{")
    w("  int k = 0;")
    w("  for (; k < kNumVariants && k < max_variants; ++k) { out[2 * k] = kVariants[k].v.T; out[2 * k + 1] = kVariants[k].v.vy; }")
    w("  return k;")
    w("}")
    w("")
    w("int Formura_Batch_Plan (int nbuckets, const int *bucket_of_instance, const int *variant)")
    w("{")
    w("  if (!ready || nbuckets < 1) return -1;")
    w("  return lib.plan(nbuckets, bucket_of_instance, variant);")
    w("}")
    w("")
    w("int Formura_Batch_Set_State (int i, int f, const double *src)")
    w("{")
    w(f"  if (!ready || i < 0 || i >= lib.n || f < 0 || f >= {nf}) return -1;")
    w("  lib.set_state(i, f, src);")
    w("  return 0;")
    w("}")
    w("")
    w("int Formura_Batch_Get_State (int i, int f, double *dst)")
    w("{")
    w(f"  if (!ready || i < 0 || i >= lib.n || f < 0 || f >= {nf}) return -1;")
    w("  lib.get_state(i, f, dst);")
    w("  return 0;")
    w("}")
    w("")
    w("float Formura_Batch_Advance (int nsteps, int use_graph)")
    w("{")
    w("  if (!ready || nsteps < 0) return -1.0f;")
    w("  return lib.advance(nsteps, use_graph);")
    w("}")
    w("")
    w("float Formura_Batch_Advance_Bucket (int b, int nsteps)")
    w("{")
    w("  if (!ready || b < 0 || b >= (int)lib.buckets.size() || nsteps < 0) return -1.0f;")
    w("  return lib.advance_bucket(b, nsteps);")
    w("}")
    w("")
    w("float Formura_Batch_Advance_Plain (int nsteps)")
    w("{")
    w("  if (!ready || nsteps < 0) return -1.0f;")
    w("  return lib.advance_plain(nsteps);")
    w("}")
    w("")
    w("void *Formura_Batch_Device_Field (int i, int f)")
    w("{")
    w(f"  if (!ready || i < 0 || i >= lib.n || f < 0 || f >= {nf}) return nullptr;")
    w("  return lib.device_field(i, f);")
    w("}")
    w("")
    w("int Formura_Batch_Fused (int *T, int *mask)")
    w("{")
    w("  if (!ready) return -1;")
    w("  *T = lib.fused_ntiles ? lib.fused_T : 0;")
    w("  *mask = lib.fused_ntiles ? lib.fused_mask : 0;")
    w("  return lib.fused_ntiles ? 1 : 0;")
    w("}")
    w("")
    w("int Formura_Batch_Last_Graph_Nodes (void)")
    w("{")
    w("  return ready ? lib.last_graph_nodes : -1;")
    w("}")
    w("")
    w("int Formura_Batch_Diff (int i, int f, const void *ref, int ref_elem_size, double *res)")
    w("{")
    w(f"  if (!ready || i < 0 || i >= lib.n || f < 0 || f >= {nf} || !ref) return -1;")
    w("  if (ref_elem_size == 8) lib.diff(i, f, (const double *)ref, res);")
    w("  else if (ref_elem_size == 4) lib.diff(i, f, (const float *)ref, res);")
    w("  else return -1;")
    w("  return 0;")
    w("}")
    w("")
    w('}  // extern "C"')
    return "\n".join(out) + "\n"
