"""Fused chains on wide tiles whose waves split over the columns (256 columns, and
128 for INT8/BF16: each wave one 32-column block, all rows).

Every output element sees the same fragments in the same k order whichever wave
computes it, so the bar is bit-identity with the 32-column tile (every wave on all
columns) for all kinds, plus the usual bars against the oracle (INT8 bit-exact,
FP32 relative Frobenius <= 1e-5, BF16 <= 5e-4).
"""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O

from test_gpu_parity import assert_rel, TOL_BF16, TOL_F32, rand_chain, run_chain

pytestmark = pytest.mark.gpu

SHAPES = {
    "conv_like": (288, 1024, [32, 24, 144]),      # the reference's conv test row shape, narrower N
    "passes": (96, 512, [600, 72, 130]),          # many 64-row passes per wave, M tails
    "deep_k": (480, 768, [40, 64]),               # several weight chunks, the held-pass arena
    "n_tail": (64, 700, [64, 96]),                # a partial last tile
}


def oracle_of(dtype, B, As):
    if dtype == fg.DTYPE_F32:
        return O.chain_f32(B, As)
    if dtype == fg.DTYPE_BF16:
        return O.chain_bf16(B, As, O.LAYOUT_MATH)
    return O.chain_int8(B, As, O.LAYOUT_MATH)


@pytest.mark.parametrize("tile_n", [256, 128])
@pytest.mark.parametrize("dtype", [fg.DTYPE_F32, fg.DTYPE_BF16, fg.DTYPE_INT8])
@pytest.mark.parametrize("name", sorted(SHAPES))
def test_wide_tile(gpu, dtype, tile_n, name):
    K1, N, Ms = SHAPES[name]
    rng = np.random.default_rng(K1 + N + tile_n + dtype)
    B, As = rand_chain(rng, dtype, K1, N, Ms)
    Ks = [K1] + Ms[:-1]
    fused, tn = fg.chain_plan(dtype, N, Ms, Ks, fg.opts(tile_n=tile_n))
    if not fused:
        pytest.skip(f"{tile_n}-column tile does not fit LDS for this chain")
    got = run_chain(dtype, B, As, fuse=fg.FUSE_ON, tile_n=tile_n)
    narrow = run_chain(dtype, B, As, fuse=fg.FUSE_ON, tile_n=32)
    assert np.array_equal(got, narrow)
    want = oracle_of(dtype, B, As)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, want)
    else:
        assert_rel(got, want, (TOL_F32 if dtype == fg.DTYPE_F32 else TOL_BF16))


def test_wide_tile_x3(gpu):
    rng = np.random.default_rng(31)
    K1, N, Ms = 96, 512, [32, 64]
    B, As = rand_chain(rng, fg.DTYPE_F32, K1, N, Ms)
    got = run_chain(fg.DTYPE_F32, B, As, fuse=fg.FUSE_ON, tile_n=256, f32_mode=fg.F32_BF16X3)
    narrow = run_chain(fg.DTYPE_F32, B, As, fuse=fg.FUSE_ON, tile_n=32, f32_mode=fg.F32_BF16X3)
    assert np.array_equal(got, narrow)


def test_wide_tile_conv_and_epilogue(gpu):
    # the image gather and the epilogue extensions on a 256-column tile
    import torch
    from test_gpu_epilogue import int8_case, to_fg
    from test_gpu_parity import dev, empty
    g = fg.conv_geom(1, 16, 24, 24, 3, 3, pad=1)
    K1, N = fg.conv_dims(g)
    rng = np.random.default_rng(32)
    X = rng.integers(-127, 128, (1, 16, 24, 24)).astype(np.int8)
    _, As, epi = int8_case(rng, K1, N, [32, 48])
    e, keep = to_fg(epi)
    C = empty((48, N), torch.int32)
    fg.chain_conv(fg.DTYPE_INT8, C, dev(X), g, [dev(a) for a in As], o=fg.opts(fuse=fg.FUSE_ON, tile_n=256), epi=e)
    torch.cuda.synchronize()
    want = O.chain_int8_epi(O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1)), As, epi)
    assert np.array_equal(C.cpu().numpy(), want)


def test_planner_picks_wide_tiles_for_small_m():
    # the reference's conv test row 1 (multigemm/test/test_int8_3.c:13-48): M of 24..144
    fused, tn = fg.chain_plan(fg.DTYPE_INT8, 65536, [32, 24, 144], [288, 32, 24])
    assert fused and tn == 256
    # the large 3-chain keeps 64 columns with every wave on all of them
    fused, tn = fg.chain_plan(fg.DTYPE_INT8, 16384, [512, 1024, 512], [512, 512, 1024])
    assert fused and tn == 64
