"""Surrogate CNN kernels against the CPU oracle (oracle/cnn_oracle.py): the
forward pass, the loss and every gradient, on the same BF16-rounded weights
and inputs.  Tolerances: forward state 1e-5 relative L2 (the residual output
is dominated by the FP32 input state), loss 1e-3 relative, each layer's weight
and bias gradient 2e-2 relative L2 (FP32 accumulation orders differ and
BF16 roundings of intermediates may land one unit apart)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _case(B=2, seed=3):
    from formura_amd import surrogate as S
    specs = S.trajectory_specs(B, seed)
    state = S.initial_states(specs, seed)
    rng = np.random.default_rng(seed)
    state = state + rng.normal(0, 0.05, state.shape).astype(np.float32)
    target = state + rng.normal(0, 0.01, state.shape).astype(np.float32)
    return specs, state, target, S.fk_array(specs)


def test_forward_and_gradients_match_oracle():
    from formura_amd import surrogate as S
    from oracle import cnn_oracle as O
    sur = S.Surrogate.get(16)
    specs, state, target, fk = _case()
    B = len(specs)
    theta = S.init_params(11)
    # make the last layer non-trivial so every gradient path is exercised
    theta = theta * np.float32(1.0)
    sur.set_params(theta)
    scale = (0.02, 0.01)
    sur.set_scale(*scale)
    bs, bt, bf, bo = (sur.buffer(state.nbytes), sur.buffer(target.nbytes), sur.buffer(fk.nbytes),
                      sur.buffer(state.nbytes))
    bs.upload(state)
    bt.upload(target)
    bf.upload(fk)
    _say("[surrogate] forward")
    sur.forward(bs, bf, bo, B)
    out = bo.download(state.shape)
    ref = O.forward(theta.astype(np.float64), state.astype(np.float64), fk.astype(np.float64), scale)
    inc_gpu = out.astype(np.float64) - state
    inc_ref = ref["out"] - state
    e_inc = np.linalg.norm(inc_gpu - inc_ref) / np.linalg.norm(inc_ref)
    _say(f"[surrogate] forward increment rel L2 error {e_inc:.3e}")
    assert e_inc < 2e-3           # measured 3.8e-4: BF16 rounding ties and FP32 sum order
    assert np.linalg.norm(out - ref["out"]) / np.linalg.norm(ref["out"]) < 1e-5
    _say("[surrogate] gradients")
    loss = sur.gradients(bs, bf, bt, B)
    g = sur.get_grads().astype(np.float64)
    lref, gref = O.loss_and_grad(theta.astype(np.float64), state.astype(np.float64), fk.astype(np.float64),
                                 target.astype(np.float64), scale)
    _say(f"[surrogate] loss GPU {loss:.6e} oracle {lref:.6e}")
    assert abs(loss - lref) <= 1e-3 * lref
    off = 0
    for l, (ci, co) in enumerate(zip(O.CIN, O.COUT)):
        nw, nb = co * 9 * ci, co
        for name, sl in (("W", slice(off, off + nw)), ("b", slice(off + nw, off + nw + nb))):
            e = np.linalg.norm(g[sl] - gref[sl]) / max(np.linalg.norm(gref[sl]), 1e-30)
            _say(f"[surrogate] layer {l + 1} d{name}: rel L2 error {e:.3e}")
            assert e < 5e-4, f"layer {l + 1} d{name}"     # measured 8e-6 .. 4.3e-5
        off += nw + nb
    for b in (bs, bt, bf, bo):
        b.free()


def test_convolution_kernels_give_the_same_bits():
    """Every convolution kernel (Sur_Set_Conv) forms each output's sum over
    the same k-steps in the same order, so the inference forward pass, the
    training forward pass (pre-activations kept), the loss and every gradient
    (data gradients through the same kernels) are bit-identical across them."""
    from formura_amd import surrogate as S
    sur = S.Surrogate.get(16)
    specs, state, target, fk = _case(B=3, seed=5)
    B = len(specs)
    sur.set_params(S.init_params(13))
    sur.set_scale(0.02, 0.01)
    bs, bt, bf, bo = (sur.buffer(state.nbytes), sur.buffer(target.nbytes), sur.buffer(fk.nbytes),
                      sur.buffer(state.nbytes))
    bs.upload(state)
    bt.upload(target)
    bf.upload(fk)
    got = {}
    try:
        for kind in range(S.Surrogate.CONV_KINDS):
            _say(f"[surrogate] convolution kernel {kind}")
            sur.set_conv(kind)
            sur.forward(bs, bf, bo, B)
            out = bo.download(state.shape)
            loss = sur.gradients(bs, bf, bt, B)
            got[kind] = (out, loss, sur.get_grads())
    finally:
        sur.set_conv(S.Surrogate.DEFAULT_CONV)
    out0, loss0, g0 = got[0]
    assert np.isfinite(out0).all() and np.isfinite(g0).all()
    for kind, (out, loss, g) in got.items():
        assert np.array_equal(out.view(np.uint32), out0.view(np.uint32)), f"kernel {kind}: forward"
        assert loss == loss0, f"kernel {kind}: loss"
        assert np.array_equal(g.view(np.uint32), g0.view(np.uint32)), f"kernel {kind}: gradients"
    with pytest.raises(ValueError):
        sur.set_conv(S.Surrogate.CONV_KINDS)
    for b in (bs, bt, bf, bo):
        b.free()


def test_training_reduces_loss():
    """A short training run on the GPU (its own forward, backward and Adam
    kernels) lowers the loss on held-out stencil pairs -- trajectories with
    other F, k and seeds than the training ones -- to under half of the
    untrained network's."""
    from formura_amd import surrogate as S
    _say("[surrogate] short training run")
    res = S.train(n_traj=128, traj_steps=64 + 32 * 12, batch=16, lr=2e-3, max_seconds=60.0,
                  log=lambda m: _say("[surrogate] " + m))
    trained = S.Surrogate.get().get_params()
    xs, fk, ys = S.stencil_pairs(16, seed=4242)
    before = S.loss_on(S.init_params(7), res.scale, xs, fk, ys)
    after = S.loss_on(trained, res.scale, xs, fk, ys)
    _say(f"[surrogate] held-out loss {before:.3e} -> {after:.3e} after {res.steps} Adam steps")
    assert after < 0.5 * before


def test_trained_network_passes_the_checks_the_untrained_one_fails():
    """What the trained network gains over the method without it: with the
    committed weights (formura_amd/data/surrogate_gs2d_256.npz) every check of
    a checked rollout on held-out trajectories passes, so the network advances
    all 16 super-steps (512 stencil steps) and the stencil runs only the 2
    checked super-steps; an untrained network fails the first check and the
    rollout falls back to the stencil for everything."""
    from formura_amd import surrogate as S
    theta, scale, _ = S.load_weights()
    sur = S.Surrogate.get()
    specs = S.trajectory_specs(8, seed=99)
    x0 = S.initial_states(specs, 99000)
    st = S.Stencil(specs)
    for i in range(len(specs)):
        st.lib.set_state(i, [x0[i, 0].astype(np.float64), x0[i, 1].astype(np.float64)])
    st.advance(256)
    buf = sur.buffer(x0.nbytes)
    st.store(sur, buf)
    xs = buf.download(x0.shape)
    buf.free()
    fk = sur.buffer(8 * len(specs))
    fk.upload(S.fk_array(specs))
    sur.set_params(theta)
    sur.set_scale(*scale)
    ro = S.checked_rollout(sur, st, fk, xs, 16)
    _say(f"[surrogate] trained: checks {ro.checks}, fallback {ro.fallback_at}, network super-steps {ro.cnn_steps}")
    assert ro.fallback_at is None and ro.cnn_steps == 16 and max(ro.checks) <= S.BOUND
    sur.set_params(S.init_params(7))
    ro0 = S.checked_rollout(sur, st, fk, xs, 16)
    _say(f"[surrogate] untrained: checks {ro0.checks}, fallback {ro0.fallback_at}")
    assert ro0.fallback_at == 0 and ro0.checks[0] > S.BOUND
    fk.free()


def test_device_relative_l2_matches_numpy():
    """The rollout check's per-trajectory relative L2 (FP64 sums on the GPU)
    against numpy in float64."""
    from formura_amd import surrogate as S
    sur = S.Surrogate.get(16)
    rng = np.random.default_rng(4)
    B, n = 5, 2 * 256 * 256
    a = rng.standard_normal((B, n)).astype(np.float32)
    b = (a + 1e-3 * rng.standard_normal((B, n))).astype(np.float32)
    ba, bb = sur.buffer(a.nbytes), sur.buffer(b.nbytes)
    ba.upload(a)
    bb.upload(b)
    got = sur.rel_l2(ba, bb, B, n)
    want = [np.linalg.norm(a[i].astype(np.float64) - b[i]) / np.linalg.norm(b[i].astype(np.float64)) for i in range(B)]
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-12 * w
    ba.free()
    bb.free()


def test_scatter_and_gather_round_trip():
    """Sur_Scatter writes each slice of a buffer to its own device array and
    Sur_Gather reads them back in the same order, bit for bit."""
    from formura_amd import surrogate as S
    sur = S.Surrogate.get(16)
    rng = np.random.default_rng(8)
    nptr, n = 6, 1000
    src = rng.standard_normal((nptr, n)).astype(np.float32)
    bsrc, bback = sur.buffer(src.nbytes), sur.buffer(src.nbytes)
    bsrc.upload(src)
    parts = [sur.buffer(4 * n) for _ in range(nptr)]
    order = [3, 0, 5, 1, 4, 2]
    sur.scatter([parts[i].ptr for i in order], bsrc, n)
    for k, i in enumerate(order):
        assert np.array_equal(parts[i].download((n,)), src[k])
    sur.gather([parts[i].ptr for i in order], bback, n)
    assert np.array_equal(bback.download((nptr, n)), src)
    for b in parts + [bsrc, bback]:
        b.free()


def test_accumulated_check_measures_the_drift():
    """The accumulated-error check of a checked rollout: shadow instances,
    advanced by the FP32 stencil alone from the rollout's start, against the
    rollout's state of the same trajectories after each window.  Its last
    value must equal (to 1e-6 relative) the host-computed relative L2
    between the rollout's final state of those trajectories and an FP32
    stencil run of them for the same steps; and the rollout either keeps
    every accumulated error within the check's threshold (ACC_MARGIN of the
    bound) or falls back at the window where one exceeds it."""
    from formura_amd import surrogate as S
    theta, scale, _ = S.load_weights()
    sur = S.Surrogate.get()
    specs = S.trajectory_specs(8, seed=99)
    x0 = S.initial_states(specs, 99000)
    st = S.Stencil(specs)
    for i in range(len(specs)):
        st.lib.set_state(i, [x0[i, 0].astype(np.float64), x0[i, 1].astype(np.float64)])
    st.advance(256)
    buf = sur.buffer(x0.nbytes)
    st.store(sur, buf)
    xs = buf.download(x0.shape)
    fk = sur.buffer(8 * len(specs))
    fk.upload(S.fk_array(specs))
    sur.set_params(theta)
    sur.set_scale(*scale)
    shadows = S.shadow_sample(len(specs))
    n_super = 16
    sts = S.Stencil(specs, shadows=shadows)
    ro = S.checked_rollout(sur, sts, fk, xs, n_super)
    _say(f"[surrogate] accumulated errors {ro.accumulated}, local {ro.checks}, fallback {ro.fallback_at}")
    assert len(ro.accumulated) == len(ro.checks)
    acc_at = S.ACC_MARGIN * S.BOUND     # the accumulated check's threshold (the shadows are a sample)
    if ro.fallback_at is None:
        assert max(ro.accumulated) <= acc_at
        # the stencil alone, from the same start, for the shadowed trajectories
        st2 = S.Stencil(specs)
        buf.upload(xs)
        st2.load(sur, buf)
        st2.advance(S.SUPER * n_super)
        st2.store(sur, buf)
        alone = buf.download(x0.shape)
        host = max(S.rel_l2(ro.state[i].astype(np.float64), alone[i].astype(np.float64)) for i in shadows)
        _say(f"[surrogate] last accumulated error {ro.accumulated[-1]:.6e}, host {host:.6e}")
        assert abs(ro.accumulated[-1] - host) <= 1e-6 * host + 1e-12
    else:
        w = ro.fallback_at // S.CHECK_EVERY
        assert ro.accumulated[w] > acc_at or ro.checks[w] > S.BOUND
        assert all(a <= acc_at for a in ro.accumulated[:w])
    for b in (buf, fk):
        b.free()
