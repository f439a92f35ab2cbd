"""Seeded inputs for the tests (re-exported from formura_amd.rng)."""
from formura_amd.rng import gray_scott_ic, normal, splitmix64, uniform  # noqa: F401
