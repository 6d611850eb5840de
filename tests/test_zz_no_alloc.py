"""fg_sgemm_noalloc through the C ABI: the same results as fg_sgemm, bit for bit,
with no device allocation (its split-K slabs live in the library's static device
array), and a clear error for a plan whose slabs exceed that array."""
import numpy as np
import pytest

import flashgemm_amd as fg
from test_gpu_parity import dev, empty

pytestmark = pytest.mark.gpu

# largest slabs first, then smaller inputs on the same array
CASES = [(256, 256, 8192, fg.F32_MFMA),   # tiled kernel, 64 split-K slabs of 256 KB (16 MiB)
         (64, 64, 65536, fg.F32_MFMA),    # skinny quadrant kernel, 256 slabs of 16 KB
         (200, 40, 3000, fg.F32_MFMA),    # skinny, ragged, smaller again
         (128, 96, 5000, fg.F32_BF16X3)]  # split-bf16 x3 through the tiled kernel's slabs


def test_noalloc_matches_sgemm_calls_in_a_row(gpu):
    import torch
    rng = np.random.default_rng(17)
    inputs = []
    for M, N, K, mode in CASES:
        A = dev(rng.uniform(-1, 1, (M, K)).astype(np.float32))
        B = dev(rng.uniform(-1, 1, (K, N)).astype(np.float32))
        inputs.append((A, B, empty((M, N), torch.float32), empty((M, N), torch.float32)))
    # warm-up on the smallest case (loads the kernels), then nothing may be allocated
    A, B, _, Cn = inputs[2]
    M, N, K, mode = CASES[2]
    fg.sgemm_noalloc(Cn, A, B, M, N, K, f32_mode=mode)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for (M, N, K, mode), (A, B, _, Cn) in zip(CASES, inputs):
        print(f"noalloc sgemm {M}x{N}x{K} mode={mode}", flush=True)
        fg.sgemm_noalloc(Cn, A, B, M, N, K, f32_mode=mode)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, f"device memory changed by {free0 - free1} bytes"
    for (M, N, K, mode), (A, B, Cs, Cn) in zip(CASES, inputs):
        fg.sgemm(Cs, A, B, M, N, K, f32_mode=mode)
        torch.cuda.synchronize()
        got, want = Cn.cpu().numpy(), Cs.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"{M}x{N}x{K} differs from fg_sgemm"


def test_noalloc_rejects_slabs_beyond_its_array(gpu):
    import torch
    M = N = 1024
    K = 4096
    A = empty((M, K), torch.float32)
    B = empty((K, N), torch.float32)
    C = empty((M, N), torch.float32)
    assert 64 * M * N * 4 > fg.NOALLOC_SCRATCH_BYTES
    with pytest.raises(fg.FlashgemmError) as e:
        fg.sgemm_noalloc(C, A, B, M, N, K, splits=64)  # 64 slabs of 4 MiB
    assert e.value.status == fg.ERR_INVALID
    assert "workspace" in str(e.value)
