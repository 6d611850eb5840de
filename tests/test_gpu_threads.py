"""Host-side reentrancy: the library is used from two threads at once, each on its own
stream (ctypes releases the GIL, so the C calls overlap). Per-stream scratch, the plan
and pack paths, the grouped tables and the error state must not mix: every result is
bit-identical to the same call made alone."""
import threading

import numpy as np
import pytest

import flashgemm_amd as fg
from test_gpu_parity import dev, empty, rand_chain

pytestmark = pytest.mark.gpu


def test_two_threads_two_streams_same_bits(gpu):
    import torch
    print("two host threads, two streams, 40 calls each", flush=True)
    rng = np.random.default_rng(95)
    # thread 0: a split-K FP32 GEMM (slabs in its stream's scratch) and an unfused BF16 chain
    A = dev(rng.uniform(-1, 1, (64, 8192)).astype(np.float32))
    Bg = dev(rng.uniform(-1, 1, (8192, 96)).astype(np.float32))
    Bb, Asb = rand_chain(rng, fg.DTYPE_BF16, 192, 300, [256, 64])
    dBb, dAb = dev(Bb), [dev(a) for a in Asb]
    # thread 1: a fused INT8 chain (weights packed per call) and a grouped launch
    Bi, Asi = rand_chain(rng, fg.DTYPE_INT8, 128, 512, [96, 160, 64], full_int8=True)
    dBi, dAi = dev(Bi), [dev(a) for a in Asi]
    Bq, Asq = rand_chain(rng, fg.DTYPE_F32, 64, 40, [96, 32])
    dBq, dAq = dev(Bq), [dev(a) for a in Asq]

    def run0(s, out):
        fg.sgemm(out[0], A, Bg, 64, 96, 8192, stream=s)
        fg.chain(fg.DTYPE_BF16, out[1], dBb, dAb, o=fg.opts(fuse=fg.FUSE_OFF), stream=s)

    def run1(s, out):
        fg.chain(fg.DTYPE_INT8, out[0], dBi, dAi, stream=s)
        fg.grouped_chain([fg.chain_req(fg.DTYPE_F32, out[1], dBq, dAq, N=40, Ms=[96, 32], Ks=[64, 96])], stream=s)

    def outs0():
        return [empty((64, 96), torch.float32), empty((64, 300), torch.float32)]

    def outs1():
        return [empty((64, 512), torch.int32), empty((32, 40), torch.float32)]

    # each workload alone, on a stream of its own
    ref = []
    for run, mk in ((run0, outs0), (run1, outs1)):
        s = torch.cuda.Stream()
        o = mk()
        run(s, o)
        s.synchronize()
        ref.append([x.clone() for x in o])

    errors = []
    results = [None, None]

    def worker(k):
        try:
            s = torch.cuda.Stream()
            run, mk = ((run0, outs0), (run1, outs1))[k]
            o = mk()
            for _ in range(40):
                run(s, o)
            s.synchronize()
            results[k] = o
        except Exception as e:  # reported below, from the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a worker thread did not finish"
    assert not errors, errors
    for k in range(2):
        for got, want in zip(results[k], ref[k]):
            assert torch.equal(got, want), f"thread {k}: result differs from the call made alone"
