"""GPU bit-exactness of the FP32 and BF16 paths on exact-arithmetic inputs.

On the integer-valued inputs of tests/exact_inputs.py every fp32 partial sum is an
exact integer below 2^24 (checked on the CPU by tests/test_exact_inputs.py), so the
GPU's summation order, MFMA blocking, split-K and fused / unfused plans cannot change
a bit: the results must equal the oracle exactly. The BF16 chains' intermediates are
large integers that are not bf16-representable, so this pins the bf16
round-to-nearest-even of the epilogues (v_cvt_pk_bf16_f32 in the fused kernel,
bf16_dev in the unfused one) against the reference's rule (kernel_bf16.h:356-390):
truncation instead would change almost every output (test_exact_inputs).
"""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from exact_inputs import EXACT_LIMIT, exact_chain, integer_chain, to_bf16_bits
from test_gpu_parity import run_chain

pytestmark = pytest.mark.gpu

N = 256


def bf16_case(seed):
    B, As = integer_chain(np.random.default_rng(seed), N)
    want, bounds = exact_chain(B, As, "rne")
    assert max(bounds) < EXACT_LIMIT
    return to_bf16_bits(B), [to_bf16_bits(a) for a in As], want.astype(np.float32)


def f32_case(seed):
    B, As = integer_chain(np.random.default_rng(seed), N)
    want, bounds = exact_chain(B, As, None)
    assert max(bounds) < EXACT_LIMIT
    return B.astype(np.float32), [a.astype(np.float32) for a in As], want.astype(np.float32)


def same_bits(got, want):
    bad = int((got.view(np.uint32) != want.view(np.uint32)).sum())
    assert bad == 0, f"{bad} of {got.size} differ; max |diff| {np.abs(got - want).max()}"


@pytest.mark.parametrize("plan", [dict(fuse=fg.FUSE_ON, tile_n=32), dict(fuse=fg.FUSE_ON, tile_n=64),
                                  dict(fuse=fg.FUSE_ON, tile_n=128), dict(fuse=fg.FUSE_ON, tile_n=256),
                                  dict(fuse=fg.FUSE_OFF), dict(fuse=fg.FUSE_OFF, splits=3)])
def test_bf16_chain_exact_rne(gpu, plan):
    print(f"bf16 exact chain {plan}", flush=True)
    B, As, want = bf16_case(0)
    got = run_chain(fg.DTYPE_BF16, B, As, **plan)
    same_bits(got, want)
    same_bits(got, O.chain_bf16(B, As, O.LAYOUT_MATH))


@pytest.mark.parametrize("threads", [1, 2, 4])
def test_bf16_chain_exact_reference_layout(gpu, threads):
    """The reference's intermediate layout (32-wide panels for N = 256 and 1, 2, 4
    threads) through the device API, bit for bit against the oracle's layout."""
    print(f"bf16 exact chain, reference layout, {threads} threads", flush=True)
    B, As, _ = bf16_case(1)
    got = run_chain(fg.DTYPE_BF16, B, As, layout=fg.LAYOUT_REFERENCE, ref_threads=threads)
    same_bits(got, O.chain_bf16(B, As, O.LAYOUT_REFERENCE, threads))


def test_bf16_chain_exact_drop_in(gpu):
    """flashgemm_multi_bf16bf16f32 (host pointers, reference layout at
    flashgemm_thread_num = 2), bit for bit."""
    B, As, _ = bf16_case(2)
    C = np.zeros((As[-1].shape[0], N), np.float32)
    triples = []
    for a in As:
        triples += [a, a.shape[0], a.shape[1]]
    fg.flashgemm_set_thread_num(2)
    try:
        fg.flashgemm_multi_bf16bf16f32(C, B, N, len(As), *triples)
    finally:
        fg.flashgemm_set_thread_num(1)
    same_bits(C, O.chain_bf16(B, As, O.LAYOUT_REFERENCE, 2))


@pytest.mark.parametrize("plan", [dict(fuse=fg.FUSE_ON, tile_n=32), dict(fuse=fg.FUSE_ON, tile_n=96),
                                  dict(fuse=fg.FUSE_ON, tile_n=128), dict(fuse=fg.FUSE_OFF),
                                  dict(fuse=fg.FUSE_OFF, splits=4)])
def test_f32_chain_exact(gpu, plan):
    print(f"f32 exact chain {plan}", flush=True)
    B, As, want = f32_case(3)
    got = run_chain(fg.DTYPE_F32, B, As, **plan)
    same_bits(got, want)
    same_bits(got, O.chain_f32(B, As))


@pytest.mark.parametrize("M,Nn,K,splits", [(2048, 256, 512, 0), (256, 2048, 4096, 0), (64, 64, 65536, 0),
                                           (1000, 64, 2000, 0), (300, 200, 3000, 7)])
def test_sgemm_exact(gpu, M, Nn, K, splits):
    """Single FP32 GEMMs on small integers (|A| <= 3, |B| <= 7, K * 21 < 2^24): the tiled
    kernel, the skinny kernel and split-K with its fixed-order reduction are exact."""
    import torch
    from test_gpu_parity import dev, empty
    print(f"sgemm exact {M}x{Nn}x{K} splits={splits}", flush=True)
    assert 21 * K < EXACT_LIMIT
    rng = np.random.default_rng(M + K)
    A = rng.integers(-3, 4, (M, K)).astype(np.float32)
    B = rng.integers(-7, 8, (K, Nn)).astype(np.float32)
    C = empty((M, Nn), torch.float32)
    fg.sgemm(C, dev(A), dev(B), M, Nn, K, splits=splits)
    torch.cuda.synchronize()
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    same_bits(C.cpu().numpy(), want)
