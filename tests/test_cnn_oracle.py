"""Pins for the CNN oracle: its float64 gradients against central finite
differences of its own loss, on a small image; BF16 rounding is
round-to-nearest-even; the circular convolution and its adjoint satisfy
<conv(a), z> = <a, conv_t(z)>."""
from __future__ import annotations

import numpy as np

from formura_amd.surrogate import init_params
from oracle import cnn_oracle as O


def small_case(seed=0, Hs=6, Ws=5):
    rng = np.random.default_rng(seed)
    state = rng.uniform(0, 1, (1, 2, Hs, Ws))
    target = state + rng.normal(0, 0.05, state.shape)
    fk = np.array([[0.03, 0.06]])
    theta = init_params(seed).astype(np.float64)
    theta[-(2 * 9 * 48 + 2):] *= 10.0        # a last layer large enough to matter
    return theta, state, target, fk


def test_gradients_match_finite_differences():
    theta, state, target, fk = small_case()
    scale = (0.3, 0.2)
    _, g = O.loss_and_grad(theta, state, fk, target, scale, quantize=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(theta.size, 12, replace=False),
                          [0, 4 * 9 * 4 + 3, theta.size - 1, theta.size - 3]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp, _ = O.loss_and_grad(tp, state, fk, target, scale, quantize=False)
        lm, _ = O.loss_and_grad(tm, state, fk, target, scale, quantize=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-8, abs(fd)) + 1e-12, (i, fd, g[i])


def test_conv_adjoint_identity():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(2, 5, 7, 3))
    z = rng.normal(size=(2, 5, 7, 4))
    w = rng.normal(size=(4, 3, 3, 3))
    lhs = float((O.conv(a, w) * z).sum())
    rhs = float((a * O.conv_t(z, w)).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # the weight gradient is the derivative of <conv(a, w), z> in w
    gw = O.conv_w(a, z)
    dw = rng.normal(size=w.shape)
    assert abs(float((O.conv(a, dw) * z).sum()) - float((gw * dw).sum())) <= 1e-10 * abs(lhs)


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-30, 65504.0], dtype=np.float32)
    r = O.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even -> 1; 1 + 3*2^-8 rounds up to 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5
    assert r[5] == 65536.0
    assert abs(r[4] - 3.0e-30) / 3.0e-30 <= 2 ** -8


def test_split_input_keeps_sixteen_bits():
    """hi = bf16(x), lo = bf16(x - hi): hi + lo is x to 2^-16 (relative), both
    parts are BF16 values, and the first layer applied to the split input with
    tied weights is the layer applied to x itself (to the same 2^-16)."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.5, 1.5, (1, 4, 5, 4)).astype(np.float32).astype(np.float64)
    s = O.split_input(x)
    hi, lo = s[..., :4], s[..., 4:]
    assert np.array_equal(O.bf16(hi), hi) and np.array_equal(O.bf16(lo), lo)
    assert np.all(np.abs(hi + lo - x) <= 2.0 ** -16 * np.abs(x))
    assert np.max(np.abs(hi - x)) > 2.0 ** -12      # hi alone is coarser than the pair
    w = O.bf16(rng.normal(size=(6, 3, 3, 4)))
    full = O.conv(x, w)
    split = O.conv(s, O.tie(w))
    bound = 2.0 ** -16 * O.conv(np.abs(x), np.abs(w))
    assert np.all(np.abs(split - full) <= bound + 1e-15)


def test_forward_matches_torch_float64():
    """The oracle's unquantised forward pass against an independent statement
    of the same network in torch (float64): circular padding by
    torch.nn.functional.pad(mode="circular"), torch's conv2d (cross-
    correlation: out[p] = sum_d W[d] a[p + d], the oracle's convention) and
    the exact (erf) GELU.  A wrong tap offset, a transposed kernel or a
    missing wrap would change the result far beyond the tolerance."""
    torch = __import__("pytest").importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(5)
    B, H, W = 2, 7, 9          # odd and unequal sides: a wrong axis or shift shows
    state = rng.uniform(0.0, 1.0, (B, 2, H, W))
    fk = np.array([[0.03, 0.06], [0.05, 0.052]])
    theta = rng.normal(0.0, 0.2, O.num_params())
    scale = (0.3, 0.2)
    got = O.forward(theta, state, fk, scale, quantize=False)

    x = torch.from_numpy(O.pack_input(state, fk)).permute(0, 3, 1, 2)   # [B][4][H][W]
    a = x
    layers = O.split(theta)
    for l, (w, b) in enumerate(layers):
        wt = torch.from_numpy(np.ascontiguousarray(w.transpose(0, 3, 1, 2)))   # [co][ci][3][3]
        z = F.conv2d(F.pad(a, (1, 1, 1, 1), mode="circular"), wt, torch.from_numpy(b))
        a = F.gelu(z, approximate="none") if l < len(layers) - 1 else z
    out = torch.from_numpy(state).clone()
    for c in range(2):
        out[:, c] += scale[c] * a[:, c]
    ref = out.numpy()
    inc_o, inc_t = got["out"] - state, ref - state
    assert np.linalg.norm(inc_o - inc_t) <= 1e-12 * np.linalg.norm(inc_t)
    # the last layer's pre-activation as well, channel by channel
    z5 = a.permute(0, 2, 3, 1).numpy()
    assert np.max(np.abs(got["z5"] - z5)) <= 1e-12 * np.max(np.abs(z5))
