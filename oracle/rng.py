"""std::mt19937 and libstdc++'s uniform_real_distribution<double>.

TeNeS draws its initial tensors with std::mt19937(seed) and
std::uniform_real_distribution<double>(-1, 1) (src/iTPS/tensors.cpp:83-128).
numpy's MT19937 with the legacy seeding is the same generator
(init_genrand(seed)), so its raw 32-bit outputs are the std::mt19937 stream.
libstdc++ builds a double from two 32-bit draws (generate_canonical with
53 bits: (x0 + x1 * 2^32) / 2^64, pulled below 1 if it rounds to 1.0).
"""
from __future__ import annotations

import numpy as np


class MT19937:
    """Raw std::mt19937 stream (32-bit outputs)."""

    def __init__(self, seed: int):
        self._bg = np.random.MT19937(0)
        self._bg._legacy_seeding(int(seed) & 0xFFFFFFFF)

    def raw(self, n: int) -> np.ndarray:
        return self._bg.random_raw(n).astype(np.uint64)


class UniformReal:
    """libstdc++ uniform_real_distribution<double>(a, b) on a MT19937."""

    def __init__(self, a: float = -1.0, b: float = 1.0):
        self.a = a
        self.b = b

    def __call__(self, gen: MT19937, n: int = 1) -> np.ndarray:
        r = gen.raw(2 * n).reshape(n, 2)
        s = r[:, 0].astype(np.float64) + r[:, 1].astype(np.float64) * 4294967296.0
        c = s / 18446744073709551616.0
        c = np.where(c >= 1.0, np.nextafter(1.0, 0.0), c)
        return self.a + (self.b - self.a) * c
