"""Libraries beyond the per-program step libraries: the batched 2D
Gray-Scott library (FP32 bulk and its FP64 twin) built from
examples/gray-scott-2d.fmr with F, k, dt, cu, cv as per-instance parameters,
and the surrogate CNN library.  `build_all` compiles them in-tree."""
from __future__ import annotations

import os
from typing import Dict

from . import build as _b

BATCH_PROGRAMS = {
    # name -> (fmr, yaml, precision, parameter names)
    "gs2d_batch_f32": ("gray-scott-2d.fmr", "gray-scott-2d.yaml", "fp32", ("F", "k", "dt", "cu", "cv")),
    "gs2d_batch_f64": ("gray-scott-2d.fmr", "gray-scott-2d.yaml", "faithful", ("F", "k", "dt", "cu", "cv")),
}


def generate_batch(name: str) -> str:
    from .compiler.driver import compile_files
    from .compiler.emit_batched import emit_batched
    from .compiler.driver import _write_if_changed
    fmr, yaml_, precision, params = BATCH_PROGRAMS[name]
    c = compile_files(name, os.path.join(_b.EXAMPLES, fmr), os.path.join(_b.EXAMPLES, yaml_), precision, params)
    os.makedirs(_b.BUILD, exist_ok=True)
    src = os.path.join(_b.BUILD, f"{name}.hip")
    _write_if_changed(src, emit_batched(c.layout, c.ir, precision))
    return src


def build_surrogate() -> str:
    """The surrogate CNN library (hand-written, not generated)."""
    return _b.compile_hip(os.path.join(_b.CSRC, "surrogate.hip"), os.path.join(_b.BUILD, "libsurrogate.so"))


def build_all() -> Dict[str, str]:
    out = {"surrogate": build_surrogate()}
    for name in BATCH_PROGRAMS:
        src = generate_batch(name)
        out[name] = _b.compile_hip(src, _b.lib_path(name))
    return out
