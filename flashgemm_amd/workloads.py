"""Seeded synthetic inputs of the benchmark workloads, shared by bench.py and the tests.

Every generator builds device tensors with torch from a fixed seed, in the
reference's orientation (B is K1 x N, A_i is M_i x K_i, row-major; the chain is
C = A_n ... A_1 B, multigemm/src/flashgemm.h:21-25). The workloads:

* ``sgemm_operands``: FP32 A (M x K), B (K x N), uniform [-1, 1);
* ``chain2_f32``: the FP32 two-GEMM chain, N = 8192, 256 -> 256 -> 256, uniform [-1, 1);
* ``chain3``: X(16384 x 512).W1(512 x 512).W2(512 x 1024).W3(1024 x 512), INT8 entries
  uniform integers in [-127, 127] or BF16 entries N(0, 1/sqrt(K)) rounded to bf16;
* ``precision_chain``: FP32 four-GEMM chain, N = 16384, 1024 -> 4096 -> 1024 -> 4096 ->
  1024, entries N(0, 1) times a per-column scale 10^U(-3, 3) in every operand;
* ``serving_sets`` / ``serving_trace``: 64 resident weight sets (chain length 2..4,
  widths from {64 ... 2048}, 32 BF16 / 19 INT8 / 13 FP32 sets, N(0, 1/sqrt(K)) or INT8
  uniform in [-127, 127]) and a request stream (set uniform, request width N
  log-uniform in 1..4096, unit-rate Poisson gaps).
"""
from __future__ import annotations

import numpy as np

DTYPE_F32, DTYPE_BF16, DTYPE_INT8 = 0, 1, 2
SERVING_DIMS = (64, 128, 256, 512, 1024, 2048)
SERVING_MAX_N = 4096


def uniform(torch, shape, gen, dtype=None):
    x = torch.rand(shape, generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
    return x if dtype is None else x.to(dtype)


def sgemm_operands(torch, M: int, N: int, K: int, seed: int | None = None):
    gen = torch.Generator(device="cuda").manual_seed(M + N + K if seed is None else seed)
    return uniform(torch, (M, K), gen), uniform(torch, (K, N), gen)


def chain2_f32(torch, seed: int = 1):
    """B (256 x 8192), A1 (256 x 256), A2 (256 x 256)."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    N, K1, M1, M2 = 8192, 256, 256, 256
    B = uniform(torch, (K1, N), gen)
    A1 = uniform(torch, (M1, K1), gen)
    A2 = uniform(torch, (M2, M1), gen)
    return B, [A1, A2]


def chain3(torch, dtype: int, seed: int = 3, N: int = 16384):
    """B (512 x N) and the three weights; INT8 uniform in [-127, 127], BF16 N(0, 1/sqrt(K))."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    K1, Ms = 512, [512, 1024, 512]
    Ks = [K1] + Ms[:-1]
    if dtype == DTYPE_INT8:
        B = torch.randint(-127, 128, (K1, N), generator=gen, device="cuda", dtype=torch.int8)
        As = [torch.randint(-127, 128, (m, k), generator=gen, device="cuda", dtype=torch.int8) for m, k in zip(Ms, Ks)]
    else:
        B = (torch.randn((K1, N), generator=gen, device="cuda") / K1 ** 0.5).to(torch.bfloat16)
        As = [(torch.randn((m, k), generator=gen, device="cuda") / k ** 0.5).to(torch.bfloat16) for m, k in zip(Ms, Ks)]
    return B, As


PRECISION_WIDTHS = (1024, 4096, 1024, 4096, 1024)


def precision_chain(torch, seed: int = 4, N: int = 16384):
    """B (1024 x N) and four weights; every operand column scaled by 10^U(-3, 3)."""
    gen = torch.Generator(device="cuda").manual_seed(seed)
    w = PRECISION_WIDTHS

    def scaled(rows, cols):
        s = 10.0 ** (torch.rand((1, cols), generator=gen, device="cuda") * 6 - 3)
        return (torch.randn((rows, cols), generator=gen, device="cuda") * s).contiguous()

    B = scaled(w[0], N)
    As = [scaled(w[i + 1], w[i]) for i in range(4)]
    return B, As


def serving_sets(torch, seed: int = 5, n_sets: int = 64):
    """Weight sets: dict(dtype, w (widths, K1 first), As (device weights), B (a device
    buffer of w[0] x SERVING_MAX_N elements that every request of the set reads its
    first w[0] x N elements from, row-major with ldb = N))."""
    rng = np.random.default_rng(seed)
    mix = [DTYPE_BF16] * 32 + [DTYPE_INT8] * 19 + [DTYPE_F32] * 13  # 50 / 30 / 20 % of 64
    dts = rng.permutation(mix)[:n_sets]
    gen = torch.Generator(device="cuda").manual_seed(seed)
    sets = []
    for dt in dts:
        w = [int(rng.choice(SERVING_DIMS)) for _ in range(int(rng.integers(2, 5)) + 1)]
        As = []
        for m, k in zip(w[1:], w[:-1]):
            if dt == DTYPE_INT8:
                As.append(torch.randint(-127, 128, (m, k), generator=gen, device="cuda", dtype=torch.int8))
            else:
                a = torch.randn((m, k), generator=gen, device="cuda") / k ** 0.5
                As.append(a.to(torch.bfloat16) if dt == DTYPE_BF16 else a)
        if dt == DTYPE_INT8:
            B = torch.randint(-127, 128, (w[0] * SERVING_MAX_N,), generator=gen, device="cuda", dtype=torch.int8)
        else:
            B = torch.randn((w[0] * SERVING_MAX_N,), generator=gen, device="cuda")
            B = B.to(torch.bfloat16) if dt == DTYPE_BF16 else B
        sets.append(dict(dtype=int(dt), w=w, As=As, B=B))
    return sets


def serving_trace(n_req: int, n_sets: int = 64, seed: int = 6):
    """(set index per request, N per request, unit-rate Poisson gaps): divide the
    cumulative gaps by the offered rate for arrival times in seconds."""
    rng = np.random.default_rng(seed)
    set_of = rng.integers(0, n_sets, n_req)
    Ns = np.minimum(np.exp(rng.uniform(0.0, np.log(SERVING_MAX_N), n_req)).astype(np.int64) + 1, SERVING_MAX_N)
    gaps = rng.exponential(1.0, n_req)
    return set_of, Ns, gaps


def elem_bytes(dtype: int) -> int:
    return 1 if dtype == DTYPE_INT8 else (2 if dtype == DTYPE_BF16 else 4)
