"""Kernel workload for counter collection (rocprofv3 --pmc): ten launches of
the dominant CTMRG GEMM shape (R = LT*LB at D=5, chi=100: 2500 x 2500 x 2500
FP64, row-major operands) and five launches of a plain copy kernel of known
size (4096 x 4096 FP64 transpose: 128 MiB read, 128 MiB written) whose
counter readings calibrate FETCH_SIZE / WRITE_SIZE.

    python3 tenes_amd/roofline_probe.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    import tenes_amd as ta

    ctx = ta.Context(0)
    rng = np.random.default_rng(0)
    n = 2500
    A = ctx.tensor(rng.standard_normal((n, n)))
    B = ctx.tensor(rng.standard_normal((n, n)))
    for _ in range(10):
        ta.tensordot(A, B, [1], [0])
    ctx.synchronize()
    X = ctx.tensor(rng.standard_normal((4096, 4096)))
    for _ in range(5):
        ta.transpose(X, [1, 0])
    ctx.synchronize()
    print("roofline probe: 10 GEMM 2500^3 + 5 copies of 128 MiB done", flush=True)


if __name__ == "__main__":
    main()
