"""GPU tests of the grouped level launch (fg_grouped_levels): many chains run unfused,
GEMM i of all of them (any mix of dtypes) in one launch, GEMMs with few tiles split over
k by their shape alone. Every request's result must have the same bits in a batch as
launched alone (the split and the slab order do not depend on the batch); INT8 must
equal the oracle and the unsplit unfused chain (integer sums are exact in any order),
and FP32 / BF16 / split-bf16 stay within the parity bars."""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from test_gpu_parity import TOL_BF16, TOL_F32, TOL_X3, assert_rel, dev, empty, rand_chain

pytestmark = pytest.mark.gpu


def make(rng, specs):
    """specs: (dtype, f32_mode, dims, N); dims = [K1, M1, ..., Mn]."""
    import torch
    out = []
    for dtype, mode, dims, N in specs:
        B, As = rand_chain(rng, dtype, dims[0], N, dims[1:], full_int8=True)
        C = empty((dims[-1], N), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
        dB, dA = dev(B), [dev(a) for a in As]
        out.append(dict(dtype=dtype, mode=mode, B=B, As=As, C=C, dB=dB, dA=dA, N=N,
                        req=fg.chain_req(dtype, C, dB, dA, N=N, f32_mode=mode)))
    return out


def alone(q):
    """The same request through the level launch on its own."""
    import torch
    C = torch.empty_like(q["C"])
    fg.grouped_levels([fg.chain_req(q["dtype"], C, q["dB"], q["dA"], N=q["N"], f32_mode=q["mode"])])
    torch.cuda.synchronize()
    return C.cpu().numpy()


def unfused(q):
    """The same chain through fg_chain, unfused, without split-K."""
    import torch
    C = torch.empty_like(q["C"])
    fg.chain(q["dtype"], C, q["dB"], q["dA"], N=q["N"],
             o=fg.opts(f32_mode=q["mode"], fuse=fg.FUSE_OFF, splits=1))
    torch.cuda.synchronize()
    return C.cpu().numpy()


def check(q, got):
    if q["dtype"] == fg.DTYPE_INT8:
        assert np.array_equal(got, O.chain_int8(q["B"], q["As"], O.LAYOUT_MATH))
    elif q["dtype"] == fg.DTYPE_BF16:
        assert_rel(got, O.chain_bf16(q["B"], q["As"], O.LAYOUT_MATH), TOL_BF16, "bf16")
    else:
        assert_rel(got, O.chain_f32(q["B"], q["As"]), TOL_X3 if q["mode"] == fg.F32_BF16X3 else TOL_F32,
                   "x3" if q["mode"] == fg.F32_BF16X3 else "f32")


def run_levels(qs):
    import torch
    for q in qs:
        q["C"].fill_(-7)
    torch.cuda.synchronize()
    fg.grouped_levels([q["req"] for q in qs])
    torch.cuda.synchronize()
    return [q["C"].cpu().numpy() for q in qs]


I8, BF, F32 = fg.DTYPE_INT8, fg.DTYPE_BF16, fg.DTYPE_F32
MFMA, X3 = fg.F32_MFMA, fg.F32_BF16X3


def mixed_specs(rng, count):
    """Mixed dtypes and modes, chain lengths 1..4, widths with and without whole
    128-row tiles, N from 1 to 700 (ragged, odd)."""
    widths = [32, 96, 128, 200, 256, 520, 1024]
    kinds = [(BF, MFMA), (I8, MFMA), (F32, MFMA), (F32, X3)]
    specs = []
    for r in range(count):
        dtype, mode = kinds[r % 4]
        n = 1 + r % 4
        dims = [int(rng.choice(widths)) for _ in range(n + 1)]
        specs.append((dtype, mode, dims, int(rng.integers(1, 700))))
    return specs


def test_levels_mixed_batch_same_bits_as_alone(gpu):
    rng = np.random.default_rng(31)
    qs = make(rng, mixed_specs(rng, 24))
    got = run_levels(qs)
    for i, (q, g) in enumerate(zip(qs, got)):
        print(f"  request {i}: dtype={q['dtype']} mode={q['mode']} dims={[q['As'][0].shape[1]] + [a.shape[0] for a in q['As']]} N={q['N']}",
              flush=True)
        a = alone(q)
        assert np.array_equal(g.view(np.uint32), a.view(np.uint32)), f"request {i}: {(g != a).sum()} differ"
        if q["dtype"] == fg.DTYPE_INT8:
            assert np.array_equal(g, unfused(q))
        check(q, g)


def test_levels_serving_sized_chains(gpu):
    """Chains too large for the on-chip grouped tile (2048-wide intermediates), as the
    serving engine hands them over; two batches in a row on one stream, the second
    smaller (scratch and table slots reused)."""
    rng = np.random.default_rng(32)
    first = make(rng, [(BF, MFMA, [2048, 2048, 512, 2048], 300), (I8, MFMA, [1024, 2048, 1024], 77),
                       (F32, MFMA, [512, 2048, 64], 1), (BF, MFMA, [64, 2048, 2048, 128, 256], 512)])
    second = make(rng, [(I8, MFMA, [2048, 1024, 64], 33), (F32, MFMA, [256, 1024, 1024], 129)])
    for qs in (first, second):
        got = run_levels(qs)
        for q, g in zip(qs, got):
            assert np.array_equal(g.view(np.uint32), alone(q).view(np.uint32))
            if q["dtype"] == fg.DTYPE_INT8:
                assert np.array_equal(g, unfused(q))
            check(q, g)


def test_levels_argument_errors(gpu):
    import torch
    rng = np.random.default_rng(33)
    (q,) = make(rng, [(BF, MFMA, [64, 64, 64], 32)])
    fg.grouped_levels([])  # nothing to do
    packed = fg.chain_req(BF, q["C"], q["dB"], q["dA"], N=32, a_packed=1)
    with pytest.raises(fg.FlashgemmError) as e:
        fg.grouped_levels([q["req"], packed])
    assert e.value.status == fg.ERR_UNSUPPORTED
    bad = fg.chain_req(BF, q["C"], q["dB"], q["dA"], N=32, Ms=[64, 64], Ks=[64, 32])
    with pytest.raises(fg.FlashgemmError) as e:
        fg.grouped_levels([bad])
    assert e.value.status == fg.ERR_INVALID
    torch.cuda.synchronize()
