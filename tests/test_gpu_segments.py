"""Fusion depth on the GPU: chains whose whole arena does not fit LDS run as segments
(fused runs where they fit, the intermediates between segments spilled to HBM
requantized exactly as the unfused epilogue does), and the autotune table's plan for a
shape bucket steers both fg_chain and the grouped launch. INT8 is bit-exact against the
oracle and against the all-unfused plan; FP32 / BF16 within the usual bars."""
import math

import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O
from test_gpu_parity import TOL_BF16, TOL_F32, assert_rel, dev, empty, rand_chain, run_chain

pytestmark = pytest.mark.gpu

# (dtype, K1, Ms, N, expected segments): a fused head whose last GEMM writes the next
# operand to HBM, then unfused GEMMs (a 4096 / 8192-row intermediate fits no arena);
# and an unfused head (K1 = 8192 rows of B fit no arena) before a fused tail.
CASES = [
    (fg.DTYPE_INT8, 128, [128, 128, 8192, 64], 256, "head"),
    (fg.DTYPE_BF16, 128, [128, 128, 4096, 64], 256, "head"),
    (fg.DTYPE_F32, 128, [128, 128, 4096, 64], 256, "head"),
    (fg.DTYPE_INT8, 8192, [128, 128, 64], 256, "tail"),
    (fg.DTYPE_BF16, 4096, [128, 96, 64], 200, "tail"),
]


@pytest.mark.parametrize("dtype,K1,Ms,N,where", CASES)
def test_segmented_chain(gpu, dtype, K1, Ms, N, where):
    print(f"segments dtype={dtype} K1={K1} Ms={Ms} N={N}", flush=True)
    Ks = [K1] + Ms[:-1]
    segs = fg.chain_plan_segments(dtype, N, Ms, Ks)
    print("  plan", segs, flush=True)
    fused = [(i, e, t) for i, (e, t) in enumerate(segs) if t > 0]
    assert len(fused) == 1, segs
    i, end, tile = fused[0]
    start = segs[i - 1][0] + 1 if i > 0 else 0
    assert end - start >= 1  # a fused run of two GEMMs or more
    assert (start == 0) == (where == "head") and (end == len(Ms) - 1) == (where == "tail")
    rng = np.random.default_rng(K1 + sum(Ms) + dtype)
    B, As = rand_chain(rng, dtype, K1, N, Ms)
    got = run_chain(dtype, B, As)
    if dtype == fg.DTYPE_INT8:
        want = O.chain_int8(B, As, O.LAYOUT_MATH)
        assert np.array_equal(got, want), f"{(got != want).sum()} of {got.size} differ"
        assert np.array_equal(got, run_chain(dtype, B, As, fuse=fg.FUSE_OFF))
    elif dtype == fg.DTYPE_BF16:
        assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)
    else:
        assert_rel(got, O.chain_f32(B, As), TOL_F32)


def bucket(x):
    return 0 if x <= 1 else math.ceil(math.log2(x))


def table_line(kind, N, Ms, Ks, segs, splits=0):
    head = [kind, len(Ms), bucket(N)] + [bucket(m) for m in Ms] + [bucket(k) for k in Ks]
    body = [len(segs)] + [v for s in segs for v in s] + [splits]
    return " ".join(str(v) for v in head + body) + " 0.5"


def test_table_steers_grouped_launch_and_planner(gpu, tmp_path):
    """An entry for a request's shape bucket changes the grouped plan (64-column tiles,
    or out of the grouped launch onto its own segmented plan) and fg_chain's plan; the
    INT8 results stay bit-identical to the untuned run and equal the oracle."""
    import torch
    fg.autotune_clear()
    try:
        rng = np.random.default_rng(12)
        K1, Ms, N = 96, [64, 160, 64], 512
        Ks = [K1] + Ms[:-1]
        B, As = rand_chain(rng, fg.DTYPE_INT8, K1, N, Ms, full_int8=True)
        dB, dA = dev(B), [dev(a) for a in As]
        want = O.chain_int8(B, As, O.LAYOUT_MATH)

        def grouped():
            C = empty((Ms[-1], N), torch.int32)
            req = fg.chain_req(fg.DTYPE_INT8, C, dB, dA, N=N)
            tile = fg.grouped_tile(req)
            fg.grouped_chain([req])
            torch.cuda.synchronize()
            return tile, C.cpu().numpy()

        tile0, got0 = grouped()
        assert tile0 == 32 and np.array_equal(got0, want)
        for segs, tile_want, plan_want in (
                ([(2, 64)], 64, [(2, 64)]),                       # whole chain at 64 columns
                ([(1, 32), (2, 0)], 0, [(1, 32), (2, 0)]),        # fused head, unfused last GEMM
                ([(0, 0), (1, 0), (2, 0)], 0, [(0, 0), (1, 0), (2, 0)])):  # all unfused
            path = tmp_path / "table.txt"
            path.write_text("flashgemm-autotune v2 gfx950\n" + table_line(2, N, Ms, Ks, segs) + "\n")
            fg.autotune_clear()
            fg.autotune_load(str(path))
            assert fg.chain_plan_segments(fg.DTYPE_INT8, N, Ms, Ks) == plan_want
            tile, got = grouped()
            print(f"  table {segs}: grouped tile {tile}", flush=True)
            assert tile == tile_want
            assert np.array_equal(got, got0)
            assert np.array_equal(run_chain(fg.DTYPE_INT8, B, As), want)
    finally:
        fg.autotune_clear()


def test_autotune_measures_segments(gpu):
    """The autotuner times the candidates of a chain that does not fit whole (random
    data) and its pick is what the planner then uses; results equal the oracle."""
    fg.autotune_clear()
    try:
        K1, Ms, N = 128, [128, 128, 4096, 64], 1024
        Ks = [K1] + Ms[:-1]
        best = fg.autotune_chain(fg.DTYPE_BF16, N, Ms, Ks)
        segs = fg.chain_plan_segments(fg.DTYPE_BF16, N, Ms, Ks)
        print(f"  autotuned {best} -> {segs}", flush=True)
        assert fg.autotune_size() == 1 and not best["fused"]
        assert segs[-1][0] == len(Ms) - 1
        rng = np.random.default_rng(3)
        B, As = rand_chain(rng, fg.DTYPE_BF16, K1, N, Ms)
        assert_rel(run_chain(fg.DTYPE_BF16, B, As), O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)
    finally:
        fg.autotune_clear()
