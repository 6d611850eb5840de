"""Integer-valued chain inputs on which FP32 and BF16 arithmetic is exact.

Every operand entry is a small integer, so every product is an integer, and every
partial sum of every GEMM (in any order) is bounded by the row abs-sum of A_i times
the largest |X_i|. When that bound stays below 2^24 every partial sum is an exactly
representable fp32 integer: each fp32 addition is exact, and the result does not
depend on summation order, k blocking, split-K or the MFMA's internal order. The
only rounding left is the one the algorithm specifies, the bf16 round-to-nearest-even
of the BF16 chain's intermediates (multigemm/src/kernel_bf16.h:356-390): values of
several thousand and more are not bf16-representable, so that rounding shows in
nearly every output, and truncating instead of rounding changes it.

The bound uses max |X_i| over the whole intermediate, so it also holds when the
intermediate is permuted within panels (the reference's INT8/BF16 layout is a
bijection within every 32-wide panel's row pairs/quads).
"""
from __future__ import annotations

import numpy as np

EXACT_LIMIT = 1 << 24

# GEMM widths and entry bounds: K1 = 64, 96 -> 64 -> 80 output rows per GEMM.
K1, MS, BMAX, AMAX = 64, (96, 64, 80), 15, (15, 3, 1)


def integer_chain(rng, N: int, K1_: int = K1, Ms=MS, bmax: int = BMAX, amax=AMAX):
    """B (K1 x N) and the A_i as int64 arrays with entries in [-bmax, bmax] / [-amax_i, amax_i]."""
    Ks = [K1_] + list(Ms[:-1])
    B = rng.integers(-bmax, bmax + 1, (K1_, N)).astype(np.int64)
    As = [rng.integers(-a, a + 1, (m, k)).astype(np.int64) for a, m, k in zip(amax, Ms, Ks)]
    return B, As


def bf16_round(y: np.ndarray, mode: str) -> np.ndarray:
    """fp32 values (exact integers below 2^24) rounded to bf16 and back, as int64.
    mode "rne": round to nearest, ties to even; "trunc": drop the low 16 bits."""
    f = np.asarray(y, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    if mode == "rne":
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    elif mode == "trunc":
        u = u & 0xFFFF0000
    else:
        raise ValueError(mode)
    return u.astype(np.uint32).view(np.float32).astype(np.int64)


def exact_chain(B: np.ndarray, As, rounding: str | None):
    """The chain in exact integer arithmetic: C = A_n ... A_1 B with every intermediate
    rounded by bf16_round(., rounding) (None: FP32, no rounding). Returns (C as int64,
    the per-GEMM bound max_m sum_k |A_i[m][k]| * max |X_i|)."""
    X = np.asarray(B, np.int64)
    bounds = []
    for i, A in enumerate(As):
        A = np.asarray(A, np.int64)
        bounds.append(int(np.abs(A).sum(1).max() * np.abs(X).max()))
        Y = A @ X
        if i == len(As) - 1:
            return Y, bounds
        X = Y if rounding is None else bf16_round(Y, rounding)
    raise ValueError("empty chain")


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Small integers (|x| <= 256) as raw bf16 bits (exact)."""
    return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)
