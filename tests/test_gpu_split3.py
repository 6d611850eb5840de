"""GPU tests of the three-piece split FP32 mode (FG_F32_BF16X6): x = hi + mid + lo in
bf16, the six products of order <= 2 accumulated in fp32 as one BF16 GEMM over 6K.
Its error against FP64 must stay within twice plain FP32 MFMA's on the same inputs (the
target the two-piece split cannot reach), small-integer inputs must give the exact
result, and the paths that cannot run it refuse it."""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from test_gpu_parity import dev, empty

pytestmark = pytest.mark.gpu


def sgemm(A, B, mode):
    import torch
    M, K = A.shape
    N = B.shape[1]
    C = empty((M, N), torch.float32)
    fg.sgemm(C, dev(A), dev(B), M, N, K, f32_mode=mode)
    torch.cuda.synchronize()
    return C.cpu().numpy()


def rel(got, want):
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


@pytest.mark.parametrize("M,N,K", [(512, 384, 1000), (130, 77, 4099), (2048, 2048, 2048)])
def test_x6_sgemm_within_twice_fp32(gpu, M, N, K):
    rng = np.random.default_rng(M + N + K)
    # 24 significant bits per entry, signs and four decades of magnitude
    A = (rng.uniform(-1, 1, (M, K)) * 10.0 ** rng.uniform(-2, 2, (M, K))).astype(np.float32)
    B = (rng.uniform(-1, 1, (K, N)) * 10.0 ** rng.uniform(-2, 2, (K, N))).astype(np.float32)
    want = A.astype(np.float64) @ B.astype(np.float64)
    e32 = rel(sgemm(A, B, fg.F32_MFMA), want)
    e3 = rel(sgemm(A, B, fg.F32_BF16X3), want)
    e6 = rel(sgemm(A, B, fg.F32_BF16X6), want)
    print(f"  {M}x{N}x{K}: fp32 {e32:.2e}  x3 {e3:.2e} ({e3 / e32:.1f}x)  x6 {e6:.2e} ({e6 / e32:.2f}x)", flush=True)
    assert e6 <= 2 * e32
    assert e6 < e3


def test_x6_exact_on_small_integers(gpu):
    """Integers below 2^8 are whole bf16 values (mid = lo = 0) and every partial sum
    stays below 2^24: the result is exact, so the bits are the integer product."""
    rng = np.random.default_rng(41)
    A = rng.integers(-15, 16, (200, 640)).astype(np.float32)
    B = rng.integers(-15, 16, (640, 150)).astype(np.float32)
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    assert np.array_equal(sgemm(A, B, fg.F32_BF16X6), want)


def test_x6_chain_within_twice_fp32(gpu):
    import torch
    rng = np.random.default_rng(42)
    K1, N, Ms = 300, 257, [512, 96, 200]
    Ks = [K1] + Ms[:-1]
    B = rng.standard_normal((K1, N)).astype(np.float32)
    As = [(rng.standard_normal((m, k)) / np.sqrt(k)).astype(np.float32) for m, k in zip(Ms, Ks)]
    want = O.chain_f64(B.astype(np.float64), [a.astype(np.float64) for a in As])
    errs = {}
    for name, mode in (("fp32", fg.F32_MFMA), ("x6", fg.F32_BF16X6)):
        C = empty((Ms[-1], N), torch.float32)
        fg.chain(fg.DTYPE_F32, C, dev(B), [dev(a) for a in As], o=fg.opts(f32_mode=mode))
        torch.cuda.synchronize()
        errs[name] = rel(C.cpu().numpy(), want)
    print(f"  chain: {errs}", flush=True)
    assert errs["x6"] <= 2 * errs["fp32"]
    assert fg.chain_plan(fg.DTYPE_F32, N, Ms, Ks, fg.opts(f32_mode=fg.F32_BF16X6))[0] is False


def test_x6_refused_where_it_cannot_run(gpu):
    import torch
    A, B = dev(np.ones((64, 64), np.float32)), dev(np.ones((64, 64), np.float32))
    C = empty((64, 64), torch.float32)
    for call, status in (
            (lambda: fg.chain(fg.DTYPE_F32, C, B, [A], o=fg.opts(f32_mode=fg.F32_BF16X6, fuse=fg.FUSE_ON)),
             fg.ERR_UNSUPPORTED),
            (lambda: fg.sgemm_t(1, 0, C, A, B, 64, 64, 64, f32_mode=fg.F32_BF16X6), fg.ERR_UNSUPPORTED),
            (lambda: fg.grouped_chain([fg.chain_req(fg.DTYPE_F32, C, B, [A], f32_mode=fg.F32_BF16X6)]),
             fg.ERR_UNSUPPORTED),
            (lambda: fg.sgemm(C, A, B, 64, 64, 64, f32_mode=7), fg.ERR_INVALID),
            (lambda: fg.chain(fg.DTYPE_F32, C, B, [A], o=fg.opts(f32_mode=-1)), fg.ERR_INVALID)):
        with pytest.raises(fg.FlashgemmError) as e:
            call()
        assert e.value.status == status
    torch.cuda.synchronize()
