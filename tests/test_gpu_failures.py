"""Failure handling of device allocations: fg_inject_alloc_failures makes the library's
next device allocations fail as out of memory (forced from the host; no kernel runs
differently). Each caller must report FG_ERR_NOMEM, keep its state consistent, and
succeed on the same call once the injection is spent. The reference has no such
path: its calls take memory with posix_memalign and do not check the result
(multigemm/src/flashgemm.c:71-72)."""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from test_gpu_parity import assert_rel, TOL_BF16, TOL_F32, dev, empty, rand_chain

pytestmark = pytest.mark.gpu


@pytest.fixture
def injection():
    yield fg.inject_alloc_failures
    fg.inject_alloc_failures(0)


def test_scratch_allocation_failure_then_recovers(gpu, injection):
    """A split-K GEMM on a stream with no scratch yet: its slab allocation fails, then
    the same call allocates and computes."""
    import torch
    M, N, K = 64, 64, 8192
    print(f"sgemm {M}x{N}x{K} on a fresh stream, first allocation fails", flush=True)
    rng = np.random.default_rng(91)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dA, dB = dev(A), dev(B)
    C = empty((M, N), torch.float32)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    injection(1)
    with pytest.raises(fg.FlashgemmError) as e:
        fg.sgemm(C, dA, dB, M, N, K, stream=s)
    assert e.value.status == fg.ERR_NOMEM, e.value
    fg.sgemm(C, dA, dB, M, N, K, stream=s)
    s.synchronize()
    assert_rel(C.cpu().numpy(), O.sgemm(A, B), TOL_F32)


def test_grouped_table_allocation_failure_then_recovers(gpu, injection):
    """A grouped launch on a fresh stream: the request table's device copy cannot be
    allocated; the retry gives the single launch's bits."""
    import torch
    print("grouped chain, table allocation fails once", flush=True)
    rng = np.random.default_rng(92)
    reqs = []
    for dtype, dims, N in ((fg.DTYPE_BF16, [64, 96, 32], 40), (fg.DTYPE_INT8, [64, 128, 64], 36)):
        B, As = rand_chain(rng, dtype, dims[0], N, dims[1:])
        C = empty((dims[-1], N), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
        dB, dA = dev(B), [dev(a) for a in As]
        reqs.append(dict(dtype=dtype, B=B, As=As, C=C, dB=dB, dA=dA, N=N,
                         req=fg.chain_req(dtype, C, dB, dA, N=N, Ms=dims[1:], Ks=dims[:-1])))
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    injection(2)  # whichever device allocations the first call needs
    with pytest.raises(fg.FlashgemmError) as e:
        fg.grouped_chain([q["req"] for q in reqs], stream=s)
    assert e.value.status == fg.ERR_NOMEM, e.value
    injection(0)
    fg.grouped_chain([q["req"] for q in reqs], stream=s)
    s.synchronize()
    for q in reqs:
        got = q["C"].cpu().numpy()
        if q["dtype"] == fg.DTYPE_INT8:
            assert np.array_equal(got, O.chain_int8(q["B"], q["As"], O.LAYOUT_MATH))
        else:
            assert_rel(got, O.chain_bf16(q["B"], q["As"], O.LAYOUT_MATH), TOL_BF16)


def test_server_weight_registration_failure_releases_budget(gpu, injection):
    """A server whose budget holds one copy of a weight set: the registration whose
    packed-weight allocation fails gives its reservation back, so the retry fits, and
    a request on the set then matches the oracle."""
    import torch
    print("server weight registration, allocation fails once", flush=True)
    rng = np.random.default_rng(93)
    dims = [256, 512, 256]
    dtype = fg.DTYPE_BF16
    _, As = rand_chain(rng, dtype, dims[0], 1, dims[1:])
    dA = [dev(a) for a in As]
    # what registration charges: the caller's weights and the server's packed copy
    wbytes = sum(a.size * 2 + fg.packed_weight_bytes(dtype, a.shape[0], a.shape[1]) for a in As)
    srv = fg.Server(budget_bytes=wbytes + (64 << 10), num_streams=1, use_graph=False)
    try:
        injection(1)
        with pytest.raises(fg.FlashgemmError) as e:
            srv.add_weights(dtype, dA)
        assert e.value.status == fg.ERR_NOMEM, e.value
        sid = srv.add_weights(dtype, dA)  # a leaked reservation would exceed the budget here
        N = 32
        B, _ = rand_chain(rng, dtype, dims[0], N, [1])
        dB = dev(B)
        C = torch.zeros((dims[-1], N), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        srv.submit(sid, dB, C)
        srv.drain()
        assert_rel(C.cpu().numpy(), O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)
    finally:
        srv.close()


def test_server_cancel_waiting_requests(gpu):
    """One request slot held by a long chain (about 4 ms of GPU work) while nine short
    requests wait for admission: the last five are cancelled. Cancelled requests leave
    C untouched, report latency -2 and do not hold up drain; the others complete and
    match the oracle. Ids that are unknown or already done are refused."""
    import torch
    print("server cancellation of waiting requests", flush=True)
    rng = np.random.default_rng(94)
    dtype = fg.DTYPE_BF16
    big_dims, dims = [2048, 2048, 2048, 2048], [256, 512, 256]
    _, big_As = rand_chain(rng, dtype, big_dims[0], 1, big_dims[1:])
    _, As = rand_chain(rng, dtype, dims[0], 1, dims[1:])
    srv = fg.Server(max_inflight=1, num_streams=1, use_graph=False)
    try:
        big_sid = srv.add_weights(dtype, [dev(a) for a in big_As])
        sid = srv.add_weights(dtype, [dev(a) for a in As])
        N0 = 131072
        big_B = torch.randn((big_dims[0], N0), device="cuda").to(torch.bfloat16)
        big_C = torch.zeros((big_dims[-1], N0), dtype=torch.float32, device="cuda")
        small = []
        for _ in range(9):
            B, _ = rand_chain(rng, dtype, dims[0], 64, [1])
            small.append((B, dev(B), torch.zeros((dims[-1], 64), dtype=torch.float32, device="cuda")))
        torch.cuda.synchronize()
        ids = [srv.submit(big_sid, big_B, big_C)] + [srv.submit(sid, dB, C) for _, dB, C in small]
        cancelled = [srv.cancel(r) for r in ids[5:]]
        srv.drain()
        assert all(cancelled), cancelled  # the slot is held for milliseconds
        lat = srv.latencies()
        for i, (B, _, C) in enumerate(small, start=1):
            got = C.cpu().numpy()
            if i >= 5:
                assert lat[i] == -2.0 and not got.any()
            else:
                assert lat[i] > 0
                assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)
        assert lat[0] > 0
        cols = big_B[:, :16].contiguous().view(torch.int16).cpu().numpy().view(np.uint16)  # bf16 bits
        assert_rel(big_C[:, :16].cpu().numpy(), O.chain_bf16(cols, big_As, O.LAYOUT_MATH), TOL_BF16)
        assert srv.cancel(ids[0]) is False  # done
        with pytest.raises(fg.FlashgemmError):
            srv.cancel(10 ** 6)
    finally:
        srv.close()
