"""Seeded synthetic fields.

SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): the state advances by the golden-ratio increment
0x9E3779B97F4A7C15 and each output is the variant-13 finaliser of the state;
the top 53 bits give u in [0, 1) and the value is lo + (hi - lo) * u.  The
reference seeds std::mt19937 from std::random_device
(sol/inc/initialconditions.h:2820-2831), which cannot be reproduced, so
seeded fields are generated here and injected.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(shape, seed: int, lo: float = -0.1, hi: float = 0.1,
            dtype=np.float64) -> np.ndarray:
    """Field of the given numpy shape (slowest axis first), values in [lo, hi)."""
    count = int(np.prod(shape))
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(shape).astype(dtype)
