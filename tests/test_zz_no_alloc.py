"""Formura_Diff_Arrays_NoAlloc -- the comparison of two device arrays with its
partial sums in the library's static scratch array instead of memory the
library allocates -- against Formura_Diff_Device, the entry point whose
kernels it shares, through the C ABI: two calls in a row, the second on a
smaller input, each bit-equal to Formura_Diff_Device on the same data.
Named to be collected last: it passes alone in a fresh process and after
the other GPU tests."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from formura_amd.runtime import load_program
from tests.common import gray_scott_ic, normal
from tests.gpu_mem import DeviceArray, download

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(lib):
    fn = lib.lib.Formura_Diff_Arrays_NoAlloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong,
                   ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def _diff_device(lib, f, ref):
    r = (ctypes.c_double * 3)()
    assert lib.lib.Formura_Diff_Device(f, ctypes.c_void_p(ref.address), ref.elem_size, r) == 0
    return np.array(r[:], dtype=np.float64)


def test_noalloc_diff_equals_diff_device_twice_second_smaller():
    lib = load_program("gs2d")
    lib.init()
    fn = _noalloc(lib)
    n = tuple(lib.layout.intra)
    count = int(np.prod(n))
    res = DeviceArray(np.full(3, -1.0))

    # call 1: the whole field U against a perturbed copy
    u, v = gray_scott_ic(n, 5)
    lib.set_state({"U": u, "V": v})
    lib.upload()
    ref1 = DeviceArray(u + 1e-3 * normal(6, n))
    want1 = _diff_device(lib, 0, ref1)
    _say(f"[noalloc] call 1: {count} elements")
    assert fn(lib.device_field(0), 8, ref1.address, 8, count, res.address, None) == 0
    lib.lib.Formura_Download(ctypes.byref(lib.navi))       # synchronises the library stream
    got1 = download(res.address, 3)
    assert np.array_equal(got1, want1), (got1, want1)

    # call 2: a third of the elements; the field and the reference are zero
    # beyond them, so Formura_Diff_Device over the whole field sums the same
    # values in the same order (adding zeros changes no partial sum)
    m = count // 3
    a = u.ravel().copy()
    a[m:] = 0.0
    b = a + 1e-3 * normal(7, (count,))
    b[m:] = 0.0
    lib.set_state({"U": a.reshape(n), "V": v})
    lib.upload()
    ref2 = DeviceArray(b)
    want2 = _diff_device(lib, 0, ref2)
    _say(f"[noalloc] call 2: {m} elements")
    assert fn(lib.device_field(0), 8, ref2.address, 8, m, res.address, None) == 0
    lib.lib.Formura_Download(ctypes.byref(lib.navi))
    got2 = download(res.address, 3)
    assert np.array_equal(got2, want2), (got2, want2)
    d = a[:m] - b[:m]
    assert abs(got2[0] - (d * d).sum()) <= 1e-12 * (d * d).sum()
    assert abs(got2[1] - (b[:m] ** 2).sum()) <= 1e-12 * (b[:m] ** 2).sum()
    assert got2[2] == np.abs(d).max()

    # sizes beyond the maximum and unknown element sizes are refused
    assert fn(lib.device_field(0), 8, ref2.address, 8, (1 << 36) + 1, res.address, None) == -1
    assert fn(lib.device_field(0), 2, ref2.address, 8, m, res.address, None) == -1
    for x in (res, ref1, ref2):
        x.free()
