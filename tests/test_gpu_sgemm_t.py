"""GPU parity of FP32 GEMMs with transposed operands (fg_sgemm_t): op(A) . op(B)
with A stored K x M (transa) and/or B stored N x K (transb).

Bars: relative Frobenius <= 1e-5 against the oracle's sgemm on the transposed
copies (FP32 MFMA; <= 3e-5 against FP64 for the split-bf16 form), and bit-identical
to fg_sgemm on the transposed copies with the same split-K: only the staging reads
differ, the LDS layout and the MFMA order are the same.
"""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O

from test_gpu_parity import assert_rel, TOL_F32, TOL_X3, dev, empty

pytestmark = pytest.mark.gpu

SHAPES = [(256, 192, 320), (130, 70, 45), (2048, 2048, 512), (128, 128, 8192), (200, 1000, 33)]


def stored(X, trans):
    return np.ascontiguousarray(X.T) if trans else X


@pytest.mark.parametrize("ta,tb", [(0, 1), (1, 0), (1, 1), (0, 0)])
@pytest.mark.parametrize("shape", SHAPES)
def test_sgemm_transposed(gpu, ta, tb, shape):
    import torch
    M, N, K = shape
    rng = np.random.default_rng(M + N + K + 2 * ta + tb)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C = empty((M, N), torch.float32)
    splits = 4 if K >= 4096 else 1
    fg.sgemm_t(ta, tb, C, dev(stored(A, ta)), dev(stored(B, tb)), M, N, K, splits=splits)
    Cn = empty((M, N), torch.float32)
    fg.sgemm(Cn, dev(A), dev(B), M, N, K, splits=splits)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert_rel(got, O.sgemm(A, B), TOL_F32)
    assert np.array_equal(got, Cn.cpu().numpy())


@pytest.mark.parametrize("ta,tb", [(0, 1), (1, 0), (1, 1)])
def test_sgemm_transposed_bf16x3(gpu, ta, tb):
    import torch
    M, N, K = 384, 320, 256
    rng = np.random.default_rng(21 + ta + 2 * tb)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C = empty((M, N), torch.float32)
    fg.sgemm_t(ta, tb, C, dev(stored(A, ta)), dev(stored(B, tb)), M, N, K, f32_mode=fg.F32_BF16X3, splits=1)
    Cn = empty((M, N), torch.float32)
    fg.sgemm(Cn, dev(A), dev(B), M, N, K, f32_mode=fg.F32_BF16X3, splits=1)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert_rel(got, A.astype(np.float64) @ B.astype(np.float64), TOL_X3)
    assert np.array_equal(got, Cn.cpu().numpy())


def test_sgemm_t_rejects_bad_flags(gpu):
    import torch
    C = empty((8, 8), torch.float32)
    X = dev(np.zeros((8, 8), np.float32))
    with pytest.raises(fg.FlashgemmError) as e:
        fg.sgemm_t(2, 0, C, X, X, 8, 8, 8)
    assert e.value.status == fg.ERR_INVALID
