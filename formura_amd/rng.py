"""Seeded synthetic inputs (initial conditions, sweep parameters).

All random numbers come from a project-defined SplitMix64 stream, so the same
seed gives the same state on any machine."""
from __future__ import annotations

import numpy as np

MASK = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n successive SplitMix64 outputs as uint64 (vectorised)."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, shape) -> np.ndarray:
    """U[0,1) doubles from the top 53 bits."""
    n = int(np.prod(shape))
    return ((splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)).reshape(shape)


def normal(seed: int, shape) -> np.ndarray:
    """N(0,1) by Box-Muller over two uniform streams."""
    n = int(np.prod(shape))
    u1 = uniform(seed, (n,))
    u2 = uniform(seed ^ 0x5DEECE66D, (n,))
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (r * np.cos(2.0 * np.pi * u2)).reshape(shape)


def gray_scott_ic(shape, seed: int, cube_frac: float = 0.125, noise: float = 0.01):
    """u = 1, v = 0 except a centred cube (square in 2D) of side n/8 cells with
    u = 0.5, v = 0.25; uniform noise in [-noise, noise) is added to u and v
    everywhere."""
    u = np.ones(shape)
    v = np.zeros(shape)
    sl = tuple(slice(n // 2 - max(1, int(n * cube_frac)) // 2,
                     n // 2 - max(1, int(n * cube_frac)) // 2 + max(1, int(n * cube_frac))) for n in shape)
    u[sl] = 0.5
    v[sl] = 0.25
    u += noise * (2.0 * uniform(seed, shape) - 1.0)
    v += noise * (2.0 * uniform(seed + 1, shape) - 1.0)
    return u, v


def mhd_state(shape, seed: int):
    """Fields of the MHD program (examples/mhd-glm-3d.fmr): density
    1 + 0.1 N(0,1) (positive), momentum r * 0.1 N(0,1) per axis, field
    0.3 N(0,1) per axis, scalar density r * U[0,1), cleaning potential 0."""
    r = 1.0 + 0.1 * normal(seed, shape)
    out = {"r": r}
    for k, f in enumerate(("mx", "my", "mz")):
        out[f] = r * 0.1 * normal(seed + 1 + k, shape)
    for k, f in enumerate(("bx", "by", "bz")):
        out[f] = 0.3 * normal(seed + 4 + k, shape)
    out["rs"] = r * uniform(seed + 7, shape)
    out["psi"] = 0.0 * r
    return out
