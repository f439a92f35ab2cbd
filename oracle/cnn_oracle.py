"""ORACLE (test infrastructure only): CPU restatement of the surrogate CNN of
formura_amd/csrc/surrogate.hip -- forward pass, loss and gradients -- in
float64 numpy, for checking the GPU kernels.

With `quantize=True` the values are rounded to BF16 (round to nearest even)
exactly where the GPU rounds them: the packed input planes, the weights, the
stored pre-activations and activations, and the back-propagated
pre-activation gradients; everything else is float64.

The four input features x (u, v, F, k, affinely normalised in FP32) enter the
first layer as two BF16 parts each, hi = bf16(x) and lo = bf16(x - hi), in
channels 0..3 and 4..7, and the first layer applies the same weights to both
parts: W1 x is computed as W1 hi + W1 lo, so the input keeps about 16
significant bits instead of BF16's 8 (the 32-step increment of u is ~1% of u).  With `quantize=False`
it is the plain float64 network, whose gradients are pinned by central finite
differences in the CPU tests.

Parameter vector layout (FP32 on the GPU): for each layer l, W_l[co][3][3][ci]
then b_l[co]; layers 4->48->48->48->48->2.  Convolutions are circular 3x3:
z[p][co] = sum_{dy,dx,ci} W[co][dy+1][dx+1][ci] * a[p + (dy, dx)][ci] + b[co].
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np
from scipy.special import erf

CIN = (4, 48, 48, 48, 48)
COUT = (48, 48, 48, 48, 2)


def num_params() -> int:
    return sum(co * 9 * ci + co for ci, co in zip(CIN, COUT))


def split(theta: np.ndarray) -> List[Tuple[np.ndarray, np.ndarray]]:
    out, off = [], 0
    for ci, co in zip(CIN, COUT):
        w = theta[off:off + co * 9 * ci].reshape(co, 3, 3, ci)
        off += co * 9 * ci
        b = theta[off:off + co]
        off += co
        out.append((w, b))
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def pack_input(state: np.ndarray, fk: np.ndarray) -> np.ndarray:
    """state [B][2][H][W], fk [B][2] -> features [B][H][W][4], computed in
    FP32 as the GPU's pack_input does (returned as float64)."""
    B, _, H, W = state.shape
    s = state.astype(np.float32)
    f = fk.astype(np.float32)
    x = np.empty((B, H, W, 4), dtype=np.float32)
    x[..., 0] = (s[:, 0] - np.float32(0.5)) * np.float32(2.0)
    x[..., 1] = s[:, 1] * np.float32(4.0)
    x[..., 2] = ((f[:, 0] - np.float32(0.0375)) / np.float32(0.0175))[:, None, None]
    x[..., 3] = ((f[:, 1] - np.float32(0.0575)) / np.float32(0.0075))[:, None, None]
    return x.astype(np.float64)


def split_input(x: np.ndarray, quantize: bool = True) -> np.ndarray:
    """[B][H][W][4] features -> [B][H][W][8]: hi parts then lo parts."""
    if not quantize:
        return np.concatenate([x, np.zeros_like(x)], axis=-1)
    hi = bf16(x)
    return np.concatenate([hi, bf16(x - hi)], axis=-1)


def tie(w: np.ndarray) -> np.ndarray:
    """First-layer weights [co][3][3][4] -> [co][3][3][8] (hi and lo parts share them)."""
    return np.concatenate([w, w], axis=-1)


def conv(a: np.ndarray, w: np.ndarray) -> np.ndarray:
    """a [B][H][W][ci], w [co][3][3][ci] -> [B][H][W][co], circular."""
    out = 0.0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            sh = np.roll(a, shift=(-dy, -dx), axis=(1, 2))   # sh[p] = a[p + (dy, dx)]
            out = out + sh @ w[:, dy + 1, dx + 1, :].T
    return out


def conv_t(dz: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Adjoint of `conv`: da[q][ci] = sum W[co][d][ci] dz[q - d][co]."""
    out = 0.0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            sh = np.roll(dz, shift=(dy, dx), axis=(1, 2))    # sh[q] = dz[q - (dy, dx)]
            out = out + sh @ w[:, dy + 1, dx + 1, :]
    return out


def conv_w(a: np.ndarray, dz: np.ndarray) -> np.ndarray:
    """dW[co][d][ci] = sum_p dz[p][co] a[p + d][ci]."""
    co, ci = dz.shape[-1], a.shape[-1]
    g = np.empty((co, 3, 3, ci))
    dzf = dz.reshape(-1, co)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            sh = np.roll(a, shift=(-dy, -dx), axis=(1, 2)).reshape(-1, ci)
            g[:, dy + 1, dx + 1, :] = dzf.T @ sh
    return g


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(theta, state, fk, scale, quantize=True) -> Dict:
    q = bf16 if quantize else (lambda x: x)
    layers = split(np.asarray(theta, dtype=np.float64))
    x0 = split_input(pack_input(state, fk), quantize)
    acts, zs = [x0], []
    a = x0
    for l, (w, b) in enumerate(layers):
        z = conv(a, q(tie(w) if l == 0 else w)) + b
        if l < len(layers) - 1:
            zs.append(q(z))
            a = q(gelu(z))
            acts.append(a)
        else:
            z5 = z
    out = state.astype(np.float64).copy()
    for c in range(2):
        out[:, c] = out[:, c] + scale[c] * z5[..., c]
    return {"out": out, "acts": acts, "zs": zs, "z5": z5, "layers": layers}


def loss_and_grad(theta, state, fk, target, scale, quantize=True):
    """Mean over pixels and the two channels of (out - target)^2, and its
    gradient with respect to theta (same layout)."""
    q = bf16 if quantize else (lambda x: x)
    f = forward(theta, state, fk, scale, quantize)
    out, acts, zs, layers = f["out"], f["acts"], f["zs"], f["layers"]
    n = out.shape[0] * out.shape[2] * out.shape[3] * 2
    d = out - target
    loss = float((d * d).sum() / n)
    B, _, H, W = out.shape
    dz = np.zeros((B, H, W, 2))
    for c in range(2):
        dz[..., c] = 2.0 / n * d[:, c] * scale[c]
    dz = q(dz)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        w, b = layers[l]
        gw = conv_w(acts[l], dz)
        if l == 0:                       # tied hi / lo weights: the gradients add
            gw = gw[..., :4] + gw[..., 4:]
        gb = dz.reshape(-1, dz.shape[-1]).sum(axis=0)
        grads.append((gw, gb))
        if l > 0:
            da = conv_t(dz, q(w))
            dz = q(da * gelu_grad(zs[l - 1]))
    grads.reverse()
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, g
