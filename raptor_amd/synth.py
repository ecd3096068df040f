"""Seeded synthetic inputs for the benchmark configurations.

The reference seeds its random vectors with time(NULL)
(core/vector.cpp:31-38), so its right-hand sides are not reproducible; this
build defines its own generator: SplitMix64 (state += 0x9E3779B97F4A7C15,
two xor-shift-multiply rounds), uniform u = (z >> 11) * 2^-53, and
Box-Muller N(0,1) samples sqrt(-2 ln u1) cos(2 pi u2) from consecutive pairs.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN
        z = (s ^ (s >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, n):
    """n samples of U[0, 1)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_range(seed, n, lo, hi):
    return lo + (hi - lo) * uniform(seed, n)


def normal(seed, n):
    """n samples of N(0, 1)."""
    u = uniform(seed, 2 * n)
    u1 = np.maximum(u[0::2], 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])
