"""GPU parity at the benchmark's sizes, for every kernel the benchmark's search
can pick: each march variant compiled into the 512^3 libraries, with each
work split of runtime.march_chunks, and the 64^3 Gray-Scott program over its
full 1000 steps.

* FP64 (heat 512^3, Gray-Scott 512^3): 10 (heat) or 12 (Gray-Scott, the
  monitor interval rounded up to a multiple of 2T) steps of each (variant,
  split) against the CPU oracle's same steps.  The oracle's result is uploaded once
  and compared on the device: max |GPU - oracle| must be 0 and the sum of
  squared differences 0, i.e. every site equal.  The steps left when 10 is
  not a multiple of T run on a variant of the deepest depth that fits, as in
  any Formura_Forward_Device call.
* FP32 bulk (heat 512^3, Gray-Scott 512^3): every (variant, split) against
  the FP32 direct kernel (one site per thread, neighbours from global
  memory) over the same steps, bit for bit -- the march kernels evaluate
  the same float operations in the same order -- and the direct kernel's
  result against the FP64 oracle within the stated tolerance.
* Gray-Scott 64^3 FP64 (the reference's CPU-runnable case): 1000 steps of
  every (variant, split), the whole state downloaded and compared with the
  oracle's 1000 steps with numpy, bit for bit.

Tiling must not change a bit: every kernel evaluates the per-site statement
list of the reference's generated C (src/Formura/MPICxx/Translate.hs:481-629)
on the same operands.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

from formura_amd import build
from formura_amd.runtime import KIND_DIRECT, StepLibrary, load_program, march_chunks
from tests.common import gray_scott_ic, normal

pytestmark = pytest.mark.gpu



def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def compiled_variants(name):
    """(kind, T, rows, waves, pf, occ) of the march variants compiled into
    lib<name>.so, from the build's resource file (no GPU needed to list)."""
    path = build.lib_path(name)[:-3] + ".resources.json"
    if not os.path.exists(path):
        return []
    with open(path, encoding="utf-8") as f:
        return sorted(tuple(int(x) for x in k.split(",")) for k in json.load(f))


_cache = {}


def hip(name):
    if name not in _cache:
        lib = load_program(name)
        lib.init()
        _cache[name] = lib
    return _cache[name]


def state_for(lib, seed):
    n = tuple(lib.layout.intra)
    if lib.layout.field_names == ["u"]:
        return {"u": normal(seed, n)}
    u, v = gray_scott_ic(n, seed)
    return {"U": u, "V": v}


def oracle_steps(oracle_name):
    from oracle.build_oracle import compiled_extra
    return compiled_extra(oracle_name).layout.steps_per_forward


def oracle_state(oracle_name, st):
    """The CPU oracle's result after one Formura_Forward (oracle_steps)."""
    from oracle.build_oracle import compiled_extra, oracle_path
    c = compiled_extra(oracle_name)
    o = StepLibrary(oracle_path(oracle_name), c.layout, device=False)
    o.init()
    o.set_state(st)
    o.forward()
    return o.get_state()


def reference(name, make):
    """Device copies of a reference result, computed once per module."""
    key = name + "/ref"
    if key not in _cache:
        from tests.gpu_mem import DeviceArray
        st, want = make()
        _cache[key] = (st, {f: DeviceArray(v) for f, v in want.items()}, want)
    return _cache[key]


def run_and_compare(lib, st, dev_ref, label, steps):
    lib.set_state(st)
    lib.upload()
    lib.forward_device(steps)
    for i, f in enumerate(lib.layout.field_names):
        d, r, m = lib.diff_device(i, dev_ref[f].address, dev_ref[f].elem_size)
        assert m == 0.0 and d == 0.0, f"{label}: field {f} differs (max abs {m:.3e}, rel L2 {d / r:.3e})"


FP64_CASES = [("heat512", "heat512_cpu", 512), ("gs3d512", "gs3d512_cpu", 99)]


def _fp64_params():
    out = []
    for name, oname, seed in FP64_CASES:
        out.append(pytest.param(name, oname, seed, None, id=f"{name}-direct"))
        for v in compiled_variants(name):
            out.append(pytest.param(name, oname, seed, v, id=f"{name}-{'-'.join(map(str, v))}"))
    return out


def _fp64_reference(lib, oname, seed):
    st = state_for(lib, seed)
    return st, oracle_state(oname, st)


def kernel_runs(variant):
    """(kernel, split) pairs of one variant: the direct kernel for None."""
    if variant is None:
        return [((KIND_DIRECT, 1, 0, 1, 1, 1), 64)]
    return [(tuple(variant), lx) for lx in march_chunks(variant[5])]


@pytest.mark.parametrize("name,oname,seed,variant", _fp64_params())
def test_fp64_variant_full_size_bitwise(name, oname, seed, variant):
    lib = hip(name)
    st, dev_ref, _ = reference(name, lambda: _fp64_reference(lib, oname, seed))
    steps = oracle_steps(oname)
    for var, lx in kernel_runs(variant):
        label = f"{name} kernel {var} lx={lx} {steps} steps"
        _say(f"[fullsize] {label}")
        lib.set_kernel(*var)
        lib.set_chunk(lx)
        run_and_compare(lib, st, dev_ref, label, steps)
    lib.set_chunk(64)


FP32_CASES = [("heat512_f32", "heat512_cpu", 513, 1e-6), ("gs3d512_f32", "gs3d512_cpu", 98, 1e-6)]


def _fp32_reference(name, oname, seed, tol):
    """The FP32 direct kernel's result after the oracle's steps; checked against the
    FP64 oracle within `tol` (relative L2 per field)."""
    lib = hip(name)
    st = state_for(lib, seed)
    lib.set_kernel(KIND_DIRECT)
    lib.set_state(st)
    lib.upload()
    lib.forward_device(oracle_steps(oname))
    lib.download()
    got = lib.get_state()
    want = oracle_state(oname, st)
    for f in want:
        err = np.linalg.norm(got[f] - want[f]) / np.linalg.norm(want[f])
        _say(f"[fullsize] {name} direct FP32 vs FP64 oracle, field {f}: rel L2 {err:.3e}")
        assert err <= tol, f"{name}: FP32 direct kernel rel L2 {err:.3e} > {tol}"
    return st, {f: got[f].astype(np.float32) for f in got}


def _fp32_params():
    out = []
    for name, oname, seed, tol in FP32_CASES:
        for v in compiled_variants(name):
            out.append(pytest.param(name, oname, seed, tol, v, id=f"{name}-{'-'.join(map(str, v))}"))
    return out


@pytest.mark.parametrize("name,oname,seed,tol,variant", _fp32_params())
def test_fp32_variant_full_size_equals_direct(name, oname, seed, tol, variant):
    lib = hip(name)
    st, dev_ref, _ = reference(name, lambda: _fp32_reference(name, oname, seed, tol))
    steps = oracle_steps(oname)
    for var, lx in kernel_runs(variant):
        label = f"{name} kernel {var} lx={lx} {steps} steps vs FP32 direct"
        _say(f"[fullsize] {label}")
        lib.set_kernel(*var)
        lib.set_chunk(lx)
        run_and_compare(lib, st, dev_ref, label, steps)
    lib.set_chunk(64)


def _gs64_params():
    return [pytest.param(v, id="gs3d64-" + "-".join(map(str, v))) for v in compiled_variants("gs3d64")]


def _gs64_reference():
    if "gs3d64/ref" not in _cache:
        from oracle.build_oracle import compiled, oracle_path
        c = compiled("gs3d64")
        o = StepLibrary(oracle_path("gs3d64"), c.layout, device=False)
        o.init()
        st = state_for(o, 64)
        o.set_state(st)
        for _ in range(1000 // c.layout.steps_per_forward):
            o.forward()
        assert o.time_step == 1000
        _cache["gs3d64/ref"] = (st, o.get_state())
    return _cache["gs3d64/ref"]


@pytest.mark.parametrize("variant", [None] + _gs64_params())
def test_grayscott64_1000_steps_full_state_bitwise(variant):
    """Gray-Scott 3D 64^3 FP64, 1000 steps (the whole run of the reference's
    CPU-runnable case): the full state, downloaded, equals the oracle's bit
    for bit, for the direct kernel (variant None) and every march variant
    with every split."""
    lib = hip("gs3d64")
    st, want = _gs64_reference()
    for var, lx in kernel_runs(variant):
        label = f"gs3d64 kernel {var} lx={lx} 1000 steps"
        _say(f"[fullsize] {label}")
        lib.set_kernel(*var)
        lib.set_chunk(lx)
        lib.set_state(st)
        lib.upload()
        lib.forward_device(1000)
        lib.download()
        got = lib.get_state()
        for f in want:
            assert np.array_equal(got[f], want[f]), \
                f"{label}: field {f} differs, max abs {np.abs(got[f] - want[f]).max():.3e}"
    lib.set_chunk(64)
