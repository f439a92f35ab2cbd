"""Pins for the CPU oracle (oracle/emit_cxx.py, built by oracle/build_oracle.py).

Each check is chosen so that a plausible mistake -- a dropped term, a wrong
sign or index, a swapped axis, a wrong association -- fails one of them:
  * the reference's own generated C (examples-generated/pearson.c), evaluated
    statement by statement, equals the oracle bit for bit (region plan, ridge
    exchange and skewed stores included);
  * an independent numpy restatement of the config formula, in the same
    association order, equals the oracle bit for bit (3D Gray-Scott, heat);
  * pure-Python loops on a tiny grid (brute force) equal the oracle bit for bit;
  * heat: a Fourier mode decays by the closed-form factor per step;
  * Gray-Scott: u=1, v=0 is an exact fixed point; a circular shift of the
    input gives the circularly shifted output bit for bit;
  * Formura_Forward advances time_step by monitorInterval2.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from formura_amd.compiler.driver import compile_source
from formura_amd.compiler.numerical_config import NumericalConfig
from formura_amd.runtime import StepLibrary
from oracle.emit_cxx import write_oracle
from tests.common import gray_scott_ic, normal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "examples")
MPI_STUB = os.path.join(ROOT, "include", "formura_serial_mpi")


def build_oracle_variant(tmp_path, name, fmr, intra, monitor, T=1, text_override=None):
    text = text_override or open(os.path.join(EX, fmr), encoding="utf-8").read()
    nc = NumericalConfig(intra_node_shape=list(intra), mpi_grid_shape=[1] * len(intra),
                         temporal_blocking_interval=T, monitor_interval=monitor)
    c = compile_source(name, text, nc)
    d = str(tmp_path / name)
    src = write_oracle(d, c.layout, c.mm)
    so = os.path.join(d, f"lib{name}.so")
    subprocess.check_call(["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-std=gnu99",
                           "-I", d, "-I", MPI_STUB, src, "-o", so, "-lm"])
    lib = StepLibrary(so, c.layout, device=False)
    lib.init()
    return lib, c


def lap3(q):
    out = None
    for ax in range(q.ndim):
        d = (np.roll(q, -1, ax) - q) - (q - np.roll(q, 1, ax))
        out = d if out is None else out + d
    return out


def gs_numpy(u, v, steps, Du, Dv, F, k, dx, dt):
    """Gray-Scott, forward Euler, in the order the program states it:
    dU = ((-((u*v)*v)) + (F*(1-u))) + ((Du/(dx*dx))*lap(u)); u' = u + dt*dU."""
    cu = Du / (dx * dx)
    cv = Dv / (dx * dx)
    for _ in range(steps):
        du = ((-((u * v) * v)) + (F * (1.0 - u))) + (cu * lap3(u))
        dv = (((u * v) * v) - ((F + k) * v)) + (cv * lap3(v))
        u, v = u + dt * du, v + dt * dv
    return u, v


def test_gs3d_oracle_equals_numpy_restatement(tmp_path):
    lib, c = build_oracle_variant(tmp_path, "gsa", "gray-scott-3d-64.fmr", (12, 10, 14), 8)
    u, v = gray_scott_ic((12, 10, 14), 3)
    lib.set_state({"U": u, "V": v})
    lib.forward()
    got = lib.get_state()
    wu, wv = gs_numpy(u, v, 8, 2e-5, 1e-5, 0.04, 0.06, 1.0 / 64.0, 1.0)
    assert np.array_equal(got["U"], wu) and np.array_equal(got["V"], wv)


def test_heat_oracle_equals_numpy_restatement(tmp_path):
    lib, _ = build_oracle_variant(tmp_path, "heata", "heat-3d-512.fmr", (9, 11, 13), 6)
    u = normal(9, (9, 11, 13))
    lib.set_state({"u": u})
    lib.forward()
    w = u.copy()
    coef = (0.1 * 1.0) / (1.0 * 1.0)
    for _ in range(6):
        w = w + coef * lap3(w)
    assert np.array_equal(lib.get_state()["u"], w)


def test_gs3d_oracle_equals_python_loops_on_tiny_grid(tmp_path):
    n = (4, 3, 5)
    lib, _ = build_oracle_variant(tmp_path, "gsb", "gray-scott-3d-64.fmr", n, 2)
    u0, v0 = gray_scott_ic(n, 17, cube_frac=0.5)
    lib.set_state({"U": u0, "V": v0})
    lib.forward()
    got = lib.get_state()
    u, v = u0.tolist(), v0.tolist()
    cu = 2e-5 / ((1.0 / 64.0) * (1.0 / 64.0))
    cv = 1e-5 / ((1.0 / 64.0) * (1.0 / 64.0))

    def lap(q, x, y, z):
        X, Y, Z = n
        s = (q[(x + 1) % X][y][z] - q[x][y][z]) - (q[x][y][z] - q[(x - 1) % X][y][z])
        s = s + ((q[x][(y + 1) % Y][z] - q[x][y][z]) - (q[x][y][z] - q[x][(y - 1) % Y][z]))
        s = s + ((q[x][y][(z + 1) % Z] - q[x][y][z]) - (q[x][y][z] - q[x][y][(z - 1) % Z]))
        return s
    for _ in range(2):
        nu = [[[0.0] * n[2] for _ in range(n[1])] for _ in range(n[0])]
        nv = [[[0.0] * n[2] for _ in range(n[1])] for _ in range(n[0])]
        for x in range(n[0]):
            for y in range(n[1]):
                for z in range(n[2]):
                    a, b = u[x][y][z], v[x][y][z]
                    du = ((-((a * b) * b)) + (0.04 * (1.0 - a))) + (cu * lap(u, x, y, z))
                    dv = (((a * b) * b) - ((0.04 + 0.06) * b)) + (cv * lap(v, x, y, z))
                    nu[x][y][z] = a + 1.0 * du
                    nv[x][y][z] = b + 1.0 * dv
        u, v = nu, nv
    assert np.array_equal(got["U"], np.array(u)) and np.array_equal(got["V"], np.array(v))


def test_heat_fourier_mode_closed_form(tmp_path):
    n = (16, 8, 12)
    steps = 10
    lib, _ = build_oracle_variant(tmp_path, "heatb", "heat-3d-512.fmr", n, steps)
    k = (1, 2, 3)
    grids = np.meshgrid(*[np.arange(s) for s in n], indexing="ij")
    phase = sum(2 * np.pi * kk * g / s for kk, g, s in zip(k, grids, n))
    u0 = np.cos(phase)
    lib.set_state({"u": u0})
    lib.forward()
    lam = 1.0 + 0.1 * sum(2.0 * np.cos(2 * np.pi * kk / s) - 2.0 for kk, s in zip(k, n))
    want = u0 * lam ** steps
    assert np.abs(lib.get_state()["u"] - want).max() < 1e-13
    # the mean is conserved up to rounding
    assert abs(lib.get_state()["u"].sum() - u0.sum()) < 1e-10


def test_gs_fixed_point_and_translation_covariance(tmp_path):
    n = (10, 12, 8)
    lib, _ = build_oracle_variant(tmp_path, "gsc", "gray-scott-3d-64.fmr", n, 4)
    lib.set_state({"U": np.ones(n), "V": np.zeros(n)})
    lib.forward()
    s = lib.get_state()
    assert np.array_equal(s["U"], np.ones(n)) and np.array_equal(s["V"], np.zeros(n))
    u, v = gray_scott_ic(n, 23, cube_frac=0.3)
    lib.set_state({"U": u, "V": v})
    lib.forward()
    a = lib.get_state()
    shift = (3, -5, 2)
    lib.set_state({"U": np.roll(u, shift, (0, 1, 2)), "V": np.roll(v, shift, (0, 1, 2))})
    lib.forward()
    b = lib.get_state()
    assert np.array_equal(np.roll(a["U"], shift, (0, 1, 2)), b["U"])
    assert np.array_equal(np.roll(a["V"], shift, (0, 1, 2)), b["V"])


def test_time_step_advances_by_monitor_interval2(tmp_path):
    lib, c = build_oracle_variant(tmp_path, "heatc", "heat-3d-512.fmr", (4, 4, 4), 5, T=2)
    assert c.layout.steps_per_forward == 8      # 5 rounded up to a multiple of 2T = 4
    lib.set_state({"u": np.zeros((4, 4, 4))})
    lib.forward()
    lib.forward()
    assert lib.time_step == 16
    assert (lib.navi.lower_x, lib.navi.upper_x) == (c.layout.margin[0], c.layout.margin[0] + 4)


def test_reference_generated_c_equals_oracle_bitwise(tmp_path, reference_tree):
    """examples-generated/pearson.c (600^2, 20 steps per Formura_Forward, its
    own region plan and ridge exchange) against the oracle for the same .fmr."""
    from tests.golden.pearson_c_interp import PearsonC
    d = os.path.join(reference_tree, "examples-generated")
    text = open(os.path.join(d, "pearson.fmr"), encoding="utf-8").read()
    lib, c = build_oracle_variant(tmp_path, "pearson600", None, (600, 600), 20, T=1, text_override=text)
    p = PearsonC(open(os.path.join(d, "pearson.c"), encoding="utf-8").read())
    assert p.steps_per_forward == c.layout.steps_per_forward == 20
    u, v = gray_scott_ic((600, 600), 42)
    lib.set_state({"U": u, "V": v})
    # pearson.c keeps the interior at rows/columns [2, 602) at the start of a step
    p.arrays["U"][2:602, 2:602] = u
    p.arrays["V"][2:602, 2:602] = v
    for _ in range(2):
        p.forward()
        lib.forward()
        assert np.array_equal(p.arrays["U"][2:602, 2:602], lib.interior("U"))
        assert np.array_equal(p.arrays["V"][2:602, 2:602], lib.interior("V"))
    assert lib.time_step == 40


def test_parameterised_oracle_equals_baked_constants():
    """The oracle with F, k, dt, cu, cv read at run time (used to check the
    batched launches) equals the oracle with the literals baked in, when the
    run-time values are computed with the program's own arithmetic."""
    import ctypes
    from formura_amd.batched import InstanceSpec
    from oracle.build_oracle import build_oracle, compiled, compiled_extra, oracle_path
    for n in ("gs2d", "gs2d_p256"):
        if not os.path.exists(oracle_path(n)):
            build_oracle(n)
    baked = StepLibrary(oracle_path("gs2d"), compiled("gs2d").layout, device=False)
    param = StepLibrary(oracle_path("gs2d_p256"), compiled_extra("gs2d_p256").layout, device=False)
    baked.init()
    param.init()
    p = (ctypes.c_double * 5)(*InstanceSpec(256, 256, 0.04, 0.06).params())
    assert param.lib.Formura_Set_Params(p, 5) == 0
    u, v = gray_scott_ic((256, 256), 77)
    baked.set_state({"U": u, "V": v})
    param.set_state({"U": u, "V": v})
    for _ in range(5):       # 5 x 32 steps
        baked.forward()
    for _ in range(4):       # 4 x 40 steps
        param.forward()
    assert baked.time_step == param.time_step == 160
    a, b = param.get_state(), baked.get_state()
    assert np.array_equal(a["U"], b["U"]) and np.array_equal(a["V"], b["V"])


def test_heat_rk2_oracle_fourier_mode(tmp_path):
    """Two-stage program (the half step is a manifest node): a Fourier mode is
    scaled per step by the explicit midpoint method's factor 1 + z + z^2/2,
    z = c * sum_a (2 cos(2 pi k_a / n_a) - 2), c = dt D / dx^2 = 0.1."""
    n = (16, 8, 12)
    steps = 6
    lib, c = build_oracle_variant(tmp_path, "heatrk2", "heat-3d-rk2.fmr", n, steps)
    assert sum(1 for k in c.mm.step.values() if k.micro[k.root()].inst[0] != "Store") == 1
    k = (1, 2, 3)
    grids = np.meshgrid(*[np.arange(s) for s in n], indexing="ij")
    u0 = np.cos(sum(2 * np.pi * kk * g / s for kk, g, s in zip(k, grids, n)))
    lib.set_state({"u": u0})
    lib.forward()
    z = 0.1 * sum(2.0 * np.cos(2 * np.pi * kk / s) - 2.0 for kk, s in zip(k, n))
    want = u0 * (1.0 + z + 0.5 * z * z) ** steps
    assert np.abs(lib.get_state()["u"] - want).max() < 1e-13


def test_gs_rk4_oracle_equals_numpy_restatement(tmp_path):
    """Classical RK4 with six manifest stage arrays, restated in numpy in the
    program's association order: bit-equal."""
    n = (10, 12, 8)
    lib, c = build_oracle_variant(tmp_path, "gsrk4", "gray-scott-3d-rk4.fmr", n, 4)
    u, v = gray_scott_ic(n, 5)
    lib.set_state({"U": u, "V": v})
    lib.forward()
    got = lib.get_state()
    cu = 2e-5 / ((1.0 / 64.0) * (1.0 / 64.0))
    cv = 1e-5 / ((1.0 / 64.0) * (1.0 / 64.0))
    dt = 2.0

    def ddt(a, b):
        return (((-((a * b) * b)) + (0.04 * (1.0 - a))) + (cu * lap3(a)),
                (((a * b) * b) - ((0.04 + 0.06) * b)) + (cv * lap3(b)))

    for _ in range(4):
        k1 = ddt(u, v)
        u2, v2 = u + (dt / 2.0) * k1[0], v + (dt / 2.0) * k1[1]
        k2 = ddt(u2, v2)
        u3, v3 = u + (dt / 2.0) * k2[0], v + (dt / 2.0) * k2[1]
        k3 = ddt(u3, v3)
        u4, v4 = u + dt * k3[0], v + dt * k3[1]
        k4 = ddt(u4, v4)
        u = u + (dt / 6.0) * (((k1[0] + 2.0 * k2[0]) + 2.0 * k3[0]) + k4[0])
        v = v + (dt / 6.0) * (((k1[1] + 2.0 * k2[1]) + 2.0 * k3[1]) + k4[1])
    assert np.array_equal(got["U"], u) and np.array_equal(got["V"], v)


def fv_euler_numpy(q, steps, visc=True):
    """Isothermal gas dynamics in conservation form, MUSCL face states with
    generalized minmod slopes (theta 3/2), Rusanov fluxes, SSP-RK3 in
    Shu-Osher form, restated with numpy rolls in the program's association
    order.  q = [r, mx, my, mz] on grids [x, y, z]; a[i+1] is np.roll(a, -1,
    axis); a face flux of face i+1/2 is stored at i.  visc=False drops the
    Rusanov dissipation term (a planted mistake for the pins below)."""
    lam = 0.1

    def minmod3(a, b, c):
        return np.fmin(0.0, np.fmax(a, np.fmax(b, c))) + np.fmax(0.0, np.fmin(a, np.fmin(b, c)))

    def flux(q, ax):
        nxt = [np.roll(u, -1, ax) for u in q]
        prv = [np.roll(u, 1, ax) for u in q]
        g = [minmod3(1.5 * (u - p), (n - p) / 2.0, 1.5 * (n - u)) for u, p, n in zip(q, prv, nxt)]
        qa = [u + gg / 2.0 for u, gg in zip(q, g)]
        qb = [np.roll(u - gg / 2.0, -1, ax) for u, gg in zip(q, g)]

        def phys(s):
            r, m = s[0], s[1:]
            w = m[ax] / r
            f = [m[ax]] + [mm * w for mm in m]
            f[1 + ax] = m[ax] * w + 1.0 * r
            return f, w

        fa, wa = phys(qa)
        fb, wb = phys(qb)
        sp = np.fmax(wa, np.fmax(-wa, np.fmax(wb, -wb))) + 1.0
        k = 1.0 if visc else 0.0
        return [((x + y) - k * sp * (b - a)) / 2.0 for x, y, a, b in zip(fa, fb, qa, qb)]

    def outflow(q):
        fx, fy, fz = (flux(q, ax) for ax in range(3))
        return [((x - np.roll(x, 1, 0)) + (y - np.roll(y, 1, 1))) + (z - np.roll(z, 1, 2))
                for x, y, z in zip(fx, fy, fz)]

    for _ in range(steps):
        q1 = [a - lam * d for a, d in zip(q, outflow(q))]
        q2 = [0.75 * a + 0.25 * (b - lam * d) for a, b, d in zip(q, q1, outflow(q1))]
        q = [(1.0 / 3.0) * a + (2.0 / 3.0) * (b - lam * d) for a, b, d in zip(q, q2, outflow(q2))]
    return q


def fv_state(n, seed):
    r = 1.0 + 0.2 * normal(seed, n)
    return {"r": r, "mx": r * 0.1 * normal(seed + 1, n), "my": r * 0.1 * normal(seed + 2, n),
            "mz": r * 0.1 * normal(seed + 3, n)}


FV = ("r", "mx", "my", "mz")


def test_fv_euler_oracle_matches_numpy_restatement(tmp_path):
    """48 manifest nodes (three SSP-RK3 stages, each with the face fluxes of
    four fields along three axes, and two intermediate states): the oracle's
    loop nests against the numpy restatement, bit for bit.  Without the
    Rusanov dissipation the restatement is far off, so that term is pinned."""
    n = (12, 10, 14)
    lib, c = build_oracle_variant(tmp_path, "fve", "fv-euler-3d.fmr", n, 2)
    st = fv_state(n, 3)
    lib.set_state(st)
    lib.forward()
    got = lib.get_state()
    want = fv_euler_numpy([st[f] for f in FV], 2)
    for f, w in zip(FV, want):
        err = np.abs(got[f] - w).max() / np.abs(w).max()
        assert np.array_equal(got[f], w), (f, err)
    plain = fv_euler_numpy([st[f] for f in FV], 2, visc=False)
    assert np.abs(got["mx"] - plain[1]).max() > 1e-3 * np.abs(plain[1]).max()


def test_fv_euler_oracle_conserves_mass_and_momentum(tmp_path):
    """Conservation form: every face flux leaves one cell and enters its
    neighbour, so the totals of r, mx, my and mz are conserved to rounding.
    A shift of the initial state by one cell along each axis shifts the
    result (periodic grid), bit for bit."""
    n = (16, 12, 10)
    lib, c = build_oracle_variant(tmp_path, "fve2", "fv-euler-3d.fmr", n, 2)
    st = fv_state(n, 9)
    lib.set_state(st)
    for _ in range(3):
        lib.forward()
    got = lib.get_state()
    for f in FV:
        assert np.isfinite(got[f]).all()
        scale = np.abs(st[f]).sum()
        assert abs(got[f].sum() - st[f].sum()) <= 1e-13 * scale, f
        assert not np.array_equal(got[f], st[f])
    moved = {f: np.roll(v, (1, 2, 3), axis=(0, 1, 2)) for f, v in st.items()}
    lib.set_state(moved)
    for _ in range(3):
        lib.forward()
    for f, v in lib.get_state().items():
        assert np.array_equal(v, np.roll(got[f], (1, 2, 3), axis=(0, 1, 2))), f


MHD = ("r", "mx", "my", "mz", "bx", "by", "bz", "rs", "psi")


def mhd_numpy(q, steps, plant=None):
    """Isothermal ideal MHD with GLM cleaning and a passive scalar
    (examples/mhd-glm-3d.fmr), restated with numpy rolls in the program's
    association order: MUSCL face states with generalized minmod slopes
    (theta 3/2), Rusanov fluxes with speed |v_n| + 3, SSP-RK3 in Shu-Osher
    form, lam = 1/20.  q = [r, mx, my, mz, bx, by, bz, rs, psi]; the face
    flux of face i+1/2 is stored at i.  plant="no_stress" drops the
    Maxwell stress -bn*bt of the tangential momenta, plant="no_glm" the
    cleaning potential in the normal field's flux (planted mistakes for the
    pins below)."""
    lam = 1.0 / 20.0

    def minmod3(a, b, c):
        return np.fmin(0.0, np.fmax(a, np.fmax(b, c))) + np.fmax(0.0, np.fmin(a, np.fmin(b, c)))

    def phys(s, ax):
        p = (ax, (ax + 1) % 3, (ax + 2) % 3)
        r, m, b, rs, psi = s[0], s[1:4], s[4:7], s[7], s[8]
        un, ut1, ut2 = (m[k] for k in p)
        bn, bt1, bt2 = (b[k] for k in p)
        vn = un / r
        pt = 1.0 * r + ((bn * bn + bt1 * bt1) + bt2 * bt2) / 2.0
        k = 0.0 if plant == "no_stress" else 1.0
        out = [un, (un * vn + pt) - bn * bn, ut1 * vn - k * (bn * bt1), ut2 * vn - k * (bn * bt2),
               0.0 * psi if plant == "no_glm" else psi,
               vn * bt1 - bn * (ut1 / r), vn * bt2 - bn * (ut2 / r), rs * vn, 1.0 * bn]
        f = [None] * 9
        f[0], f[7], f[8] = out[0], out[7], out[8]
        for j, kk in enumerate(p):
            f[1 + kk] = out[1 + j]
            f[4 + kk] = out[4 + j]
        return f, vn

    def flux(q, ax):
        nxt = [np.roll(u, -1, ax) for u in q]
        prv = [np.roll(u, 1, ax) for u in q]
        g = [minmod3(1.5 * (u - pv), (n - pv) / 2.0, 1.5 * (n - u)) for u, pv, n in zip(q, prv, nxt)]
        qa = [u + gg / 2.0 for u, gg in zip(q, g)]
        qb = [np.roll(u - gg / 2.0, -1, ax) for u, gg in zip(q, g)]
        fa, wa = phys(qa, ax)
        fb, wb = phys(qb, ax)
        sp = np.fmax(wa, np.fmax(-wa, np.fmax(wb, -wb))) + 3.0
        return [((x + y) - sp * (b - a)) / 2.0 for x, y, a, b in zip(fa, fb, qa, qb)]

    def outflow(q):
        fx, fy, fz = (flux(q, ax) for ax in range(3))
        return [((x - np.roll(x, 1, 0)) + (y - np.roll(y, 1, 1))) + (z - np.roll(z, 1, 2))
                for x, y, z in zip(fx, fy, fz)]

    for _ in range(steps):
        q1 = [a - lam * d for a, d in zip(q, outflow(q))]
        q2 = [0.75 * a + 0.25 * (b - lam * d) for a, b, d in zip(q, q1, outflow(q1))]
        q = [(1.0 / 3.0) * a + (2.0 / 3.0) * (b - lam * d) for a, b, d in zip(q, q2, outflow(q2))]
    return q


@pytest.fixture(scope="module")
def mhd_oracle(tmp_path_factory):
    n = (12, 10, 14)
    lib, c = build_oracle_variant(tmp_path_factory.mktemp("mhd"), "mhdo", "mhd-glm-3d.fmr", n, 2)
    return lib, n


def test_mhd_oracle_matches_numpy_restatement(mhd_oracle):
    """108 manifest nodes (three SSP-RK3 stages, each with the face fluxes of
    nine fields along three axes, and two intermediate states) in two
    forwards of two steps: the oracle's loop nests against the numpy
    restatement, bit for bit.  The restatement without the Maxwell stress,
    or without the cleaning potential's flux, is far off, so those terms
    are pinned."""
    from formura_amd.rng import mhd_state
    lib, n = mhd_oracle
    st = mhd_state(n, 5)
    lib.set_state(st)
    lib.forward()
    got = lib.get_state()
    want = mhd_numpy([st[f] for f in MHD], 2)
    for f, w in zip(MHD, want):
        err = np.abs(got[f] - w).max() / np.abs(w).max()
        assert np.array_equal(got[f], w), (f, err)
    for plant, f in (("no_stress", "my"), ("no_glm", "bx")):
        bad = mhd_numpy([st[g] for g in MHD], 2, plant=plant)[MHD.index(f)]
        assert np.abs(got[f] - bad).max() > 1e-4 * np.abs(bad).max(), plant


def test_mhd_oracle_conserves_and_shifts(mhd_oracle):
    """Conservation form: the totals of all nine fields are conserved to
    rounding over three forwards; a shift of the initial state by (1, 2, 3)
    cells shifts the result (periodic grid), bit for bit."""
    from formura_amd.rng import mhd_state
    lib, n = mhd_oracle
    st = mhd_state(n, 11)
    st["psi"] = 0.05 * normal(99, n)
    lib.set_state(st)
    for _ in range(3):
        lib.forward()
    got = lib.get_state()
    for f in MHD:
        assert np.isfinite(got[f]).all()
        scale = np.abs(st[f]).sum()
        assert abs(got[f].sum() - st[f].sum()) <= 1e-13 * scale, f
        assert not np.array_equal(got[f], st[f]), f
    moved = {f: np.roll(v, (1, 2, 3), axis=(0, 1, 2)) for f, v in st.items()}
    lib.set_state(moved)
    for _ in range(3):
        lib.forward()
    for f, v in lib.get_state().items():
        assert np.array_equal(v, np.roll(got[f], (1, 2, 3), axis=(0, 1, 2))), f
