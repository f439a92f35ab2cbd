"""Python view of a step library through its C ABI (ctypes, no torch types).

A step library -- the HIP one built by formura_amd.build, or any other
library with the same generated header -- exports the state arrays, the global
`navi`, `Formura_Init` and `Formura_Forward` (reference ABI,
src/Formura/MPICxx/Translate.hs:1000-1139) plus the device-resident entry points
of include/formura_amd.h.  `StepLibrary` maps the state arrays as numpy arrays
(interior view `[lower, upper)` per axis) and calls the entry points exactly as a
`main.cpp` would.  Loading fails loudly when the library file is missing.
"""
from __future__ import annotations

import ctypes
import itertools
import os
from typing import Dict, List, Sequence

import numpy as np

from .compiler.abi import Layout, delta_name


def navigator_type(layout: Layout):
    """struct Formura_Navigator of a library with this layout (field order
    of the generated header)."""
    fields = [("time_step", ctypes.c_int)]
    for a in layout.axes:
        fields += [(f"lower_{a}", ctypes.c_int), (f"upper_{a}", ctypes.c_int), (f"offset_{a}", ctypes.c_int)]
    fields += [("mpi_comm", ctypes.c_int), ("mpi_my_rank", ctypes.c_int)]
    for d in layout.deltas:
        fields.append((delta_name(d), ctypes.c_int))
    return type("Formura_Navigator", (ctypes.Structure,), {"_fields_": fields})


class StepLibrary:
    """One loaded step library (one program, one process-global state)."""

    def __init__(self, path: str, layout: Layout, device: bool = True):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"step library {path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        self.path = path
        self.layout = layout
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.Nav = navigator_type(layout)
        self.navi = self.Nav.in_dll(self.lib, "navi")
        shape = layout.shape()
        self.arrays: Dict[str, np.ndarray] = {}
        for f, t in zip(layout.field_names, layout.field_ctype):
            ct = ctypes.c_double if t == "double" else ctypes.c_float
            n = int(np.prod(shape))
            raw = (ct * n).in_dll(self.lib, f)
            self.arrays[f] = np.ctypeslib.as_array(raw).reshape(shape)
        self.device = device
        L = self.lib
        L.Formura_Init.argtypes = [ctypes.POINTER(self.Nav), ctypes.c_int]
        L.Formura_Forward.argtypes = [ctypes.POINTER(self.Nav)]
        if device:
            L.Formura_Upload.argtypes = [ctypes.POINTER(self.Nav)]
            L.Formura_Download.argtypes = [ctypes.POINTER(self.Nav)]
            L.Formura_Forward_Device.argtypes = [ctypes.POINTER(self.Nav), ctypes.c_int]
            L.Formura_Forward_Device_Timed.argtypes = [ctypes.POINTER(self.Nav), ctypes.c_int]
            L.Formura_Forward_Device_Timed.restype = ctypes.c_float
            L.Formura_Last_Timing.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_float)]
            L.Formura_Set_Kernel.argtypes = [ctypes.c_int] * 6
            L.Formura_Set_March_Chunk.argtypes = [ctypes.c_int]
            L.Formura_Get_Kernel.argtypes = [ctypes.POINTER(ctypes.c_int)] * 7
            L.Formura_List_Kernels.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
            L.Formura_Num_Stages.restype = ctypes.c_int
            L.Formura_Num_Stage_Kernels.restype = ctypes.c_int
            L.Formura_Set_Params.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
            L.Formura_Set_Graphs.argtypes = [ctypes.c_int]
            L.Formura_Prepare_Device.argtypes = [ctypes.c_int]
            L.Formura_Device_Field.argtypes = [ctypes.c_int]
            L.Formura_Device_Field.restype = ctypes.c_void_p
            L.Formura_Reduce_Device.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
            L.Formura_Diff_Device.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_double)]

    # -- reference ABI ------------------------------------------------------
    def init(self) -> None:
        rc = self.lib.Formura_Init(ctypes.byref(self.navi), 0)
        if rc != 0:
            raise RuntimeError(f"Formura_Init returned {rc}")

    def forward(self) -> None:
        rc = self.lib.Formura_Forward(ctypes.byref(self.navi))
        if rc != 0:
            raise RuntimeError(f"Formura_Forward returned {rc}")

    @property
    def time_step(self) -> int:
        return self.navi.time_step

    def interior(self, field: str) -> np.ndarray:
        lo, hi = self.layout.lower(), self.layout.upper()
        return self.arrays[field][tuple(slice(l, h) for l, h in zip(lo, hi))]

    def set_state(self, values: Dict[str, np.ndarray]) -> None:
        for f, v in values.items():
            self.interior(f)[...] = v

    def get_state(self) -> Dict[str, np.ndarray]:
        return {f: self.interior(f).copy() for f in self.layout.field_names}

    # -- device-resident entry points --------------------------------------
    def upload(self) -> None:
        self.lib.Formura_Upload(ctypes.byref(self.navi))

    def download(self) -> None:
        self.lib.Formura_Download(ctypes.byref(self.navi))

    def forward_device(self, nsteps: int) -> None:
        self.lib.Formura_Forward_Device(ctypes.byref(self.navi), int(nsteps))

    def forward_device_timed(self, nsteps: int) -> float:
        return float(self.lib.Formura_Forward_Device_Timed(ctypes.byref(self.navi), int(nsteps)))

    def num_stages(self) -> int:
        """Manifest nodes of a multi-stage program evaluated by stage kernels
        (0: single stage)."""
        return int(self.lib.Formura_Num_Stages())

    def num_stage_kernels(self) -> int:
        """Stage kernels launched per step (independent nodes share one)."""
        return int(self.lib.Formura_Num_Stage_Kernels())

    def last_timing(self):
        """(ms of the selected kernel's launches, their count, total ms) of the
        last forward_device_timed call."""
        m, k, t = ctypes.c_float(), ctypes.c_int(), ctypes.c_float()
        self.lib.Formura_Last_Timing(ctypes.byref(m), ctypes.byref(k), ctypes.byref(t))
        return float(m.value), int(k.value), float(t.value)

    def set_kernel(self, kind: int, T: int = 1, rows: int = 0, waves: int = 1, pf: int = 1, occ: int = 1) -> None:
        if self.lib.Formura_Set_Kernel(kind, T, rows, waves, pf, occ) != 0:
            raise ValueError(f"kernel variant kind={kind} T={T} rows={rows} waves={waves} pf={pf} occ={occ} "
                             "is not compiled in")

    def set_chunk(self, lx: int) -> None:
        if self.lib.Formura_Set_March_Chunk(int(lx)) != 0:
            raise ValueError("bad march chunk")

    def get_kernel(self):
        """(kind, T, rows per wave, waves, prefetch depth, waves per SIMD, march chunk)"""
        v = [ctypes.c_int() for _ in range(7)]
        self.lib.Formura_Get_Kernel(*[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def list_kernels(self) -> List[tuple]:
        """Compiled march variants as (kind, T, rows per wave, waves, prefetch
        depth, waves per SIMD), ready for set_kernel(*v)."""
        buf = (ctypes.c_int * (6 * 256))()
        n = self.lib.Formura_List_Kernels(buf, 256)
        return [tuple(buf[6 * i + j] for j in range(6)) for i in range(n)]

    def set_graphs(self, on: bool) -> None:
        """Capture and replay the launch sequences as hipGraphs (default), or
        launch eagerly."""
        self.lib.Formura_Set_Graphs(1 if on else 0)

    def prepare(self, steps: int) -> None:
        """Capture the graphs a following forward_device*(steps) will replay,
        without running them (no effect with graphs off)."""
        if self.lib.Formura_Prepare_Device(int(steps)) != 0:
            raise ValueError("negative step count")

    def set_params(self, p: Sequence[float]) -> None:
        arr = (ctypes.c_double * len(p))(*p)
        if self.lib.Formura_Set_Params(arr, len(p)) != 0:
            raise ValueError("too many parameters")

    def device_field(self, f: int) -> int:
        return int(self.lib.Formura_Device_Field(f))

    def elem_size(self) -> int:
        return int(self.lib.Formura_Device_Elem_Size())

    def reduce(self, f: int):
        r = (ctypes.c_double * 3)()
        if self.lib.Formura_Reduce_Device(f, r) != 0:
            raise ValueError("bad field")
        return r[0], r[1]

    def diff(self, f: int, ref: "StepLibrary"):
        """(L2 of difference, L2 of ref, max abs difference) of field f against
        the same field of another library's device state."""
        return self.diff_device(f, ref.device_field(f), ref.elem_size())

    def diff_device(self, f: int, ptr: int, elem_size: int):
        """(L2 of difference, L2 of ref, max abs difference) of field f against
        a device array at `ptr` of the same shape (elements of 8 or 4 bytes)."""
        r = (ctypes.c_double * 3)()
        if self.lib.Formura_Diff_Device(f, ctypes.c_void_p(ptr), int(elem_size), r) != 0:
            raise ValueError("bad diff arguments")
        return float(np.sqrt(r[0])), float(np.sqrt(r[1])), float(r[2])


KIND_DIRECT = 0
KIND_MARCH = 1       # temporally blocked march, one z-site per lane (march3d_kernel)
KIND_STAGED = 2      # multi-stage programs: one kernel per manifest node
KIND_MARCHV1 = 3     # march with the x-carry, one z-site per lane (marchv_kernel)
KIND_MARCHV2 = 4     # march, two z-sites per lane (marchv_kernel)
KIND_MARCHV1E = 5    # as 3, with two more rows on the first and the last wave of a tile
KIND_MARCHV2E = 6    # as 4, with two more rows on the first and the last wave of a tile
MARCH_KINDS = (KIND_MARCH, KIND_MARCHV1, KIND_MARCHV2, KIND_MARCHV1E, KIND_MARCHV2E)


def balanced_runs(occ: int) -> int:
    """The Formura_Set_March_Chunk value of the balanced-run split for a
    variant compiled for `occ` waves per SIMD: as many workgroups as are
    resident at once, x cut so that each has one run of its own."""
    return -100 - occ


def whole_runs(occ: int) -> int:
    """The Formura_Set_March_Chunk value of the whole-run split for a variant
    compiled for `occ` waves per SIMD: x cut into the most chunks k per tile
    whose k x tiles workgroups are resident at once, one run each."""
    return -200 - occ


def spare_runs(occ: int) -> int:
    """The Formura_Set_March_Chunk value of the whole-run split with spare
    workgroups: whole runs, and when they leave resident workgroups idle, the
    runs shortened and the planes left at the end of every tile shared by the
    idle ones (fmr::march_split_of)."""
    return -300 - occ


def march_chunks(occ: int) -> tuple:
    """The work splits (Formura_Set_March_Chunk values) offered for a march
    variant compiled for `occ` waves per SIMD: runs of 32, 64, 128 or 256
    planes (one workgroup per run; the tiles of one run length start
    together, so neighbouring tiles read their shared ghost rows while they
    are in L2), balanced runs (one round of resident workgroups, one run
    each), whole runs (k runs per tile, k x tiles workgroups resident at
    once, all in step), whole runs with the idle resident workgroups taking
    the planes at the end of every tile, or x whole with one or two (four,
    for occ >= 4) waves per SIMD.
    The benchmark's search chooses among exactly these, and the full-size
    parity tests run every compiled variant with each of them."""
    return (32, 64, 128, 256, balanced_runs(occ), whole_runs(occ), spare_runs(occ), -1, -2) + \
        ((-4,) if occ >= 4 else ())


def load_program(name: str) -> StepLibrary:
    """Load the built HIP step library of a registered program."""
    from . import build
    from .compiler.driver import compile_files
    fmr, yaml_, precision, params, over, _ = build.PROGRAMS[name]
    c = compile_files(name, fmr, yaml_, precision, params, over)
    return StepLibrary(build.lib_path(name), c.layout, device=True)
