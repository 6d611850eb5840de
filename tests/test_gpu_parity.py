"""GPU parity: the HIP path (through the C ABI) against the oracle and against the
reference's own outputs (golden fixtures).

Bars, stated per test:
  * INT8: bit-exact (int32 accumulation, saturating requantization, the reference's
    u8 . s8 reading and, in REFERENCE layout, its intermediate layout);
  * FP32 (FP32 MFMA): relative Frobenius error <= 1e-5 against the reference /
    oracle and <= 1e-5 against FP64; plus the reference's own singlegemm check
    |C - C_ref| <= 1e-3 (singlegemm/FP32_ID23.c:125-128);
  * BF16 chains: relative Frobenius error <= 5e-4 against the reference / oracle
    (bf16 rounding of intermediates under a different fp32 summation order; on
    exact-arithmetic inputs the BF16 chains are bit-exact, test_gpu_exact.py);
  * FP32 split into bf16 hi/lo (3 products): relative Frobenius <= 3e-5 vs FP64.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden

import flashgemm_amd as fg
import oracle as O

pytestmark = pytest.mark.gpu

# Bars: about 3x (FP32 5x) the largest error measured over the GPU suite (BF16 1.6e-4,
# split-bf16 x3 8.7e-6 on the precision study, FP32 2.0e-6); the measured value is
# printed by every check (assert_rel).
TOL_F32 = 1e-5
TOL_BF16 = 5e-4
TOL_X3 = 3e-5


def assert_rel(got, want, tol, what=""):
    """Relative Frobenius error of got against want, printed, then held to tol."""
    err = O.rel_fro(got, want)
    print(f"  rel err {what}: {err:.3e} (bar {tol:.0e})", flush=True)
    assert err <= tol, f"rel err {what} {err:.3e} > {tol:.0e}"


def dev(x):
    import torch
    x = np.ascontiguousarray(x)
    if x.dtype == np.uint16:
        x = x.view(np.int16)
    if x.dtype == np.uint8:
        x = x.view(np.int8)
    return torch.from_numpy(x).cuda()


def empty(shape, dtype):
    import torch
    return torch.empty(shape, dtype=dtype, device="cuda")


def run_chain(dtype, B, As, **kw):
    import torch
    N = B.shape[1]
    M = As[-1].shape[0]
    C = empty((M, N), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
    dA = [dev(a) for a in As]
    fg.chain(dtype, C, dev(B), dA, N=N, Ms=[a.shape[0] for a in As], Ks=[a.shape[1] for a in As],
             o=fg.opts(**kw))
    torch.cuda.synchronize()
    return C.cpu().numpy()


def rand_chain(rng, dtype, K1, N, Ms, full_int8=False):
    Ks = [K1] + Ms[:-1]
    if dtype == fg.DTYPE_F32:
        return rng.uniform(-1, 1, (K1, N)).astype(np.float32), \
            [rng.uniform(-1, 1, (m, k)).astype(np.float32) for m, k in zip(Ms, Ks)]
    if dtype == fg.DTYPE_BF16:
        B = O.bf16_rne(rng.standard_normal((K1, N)).astype(np.float32))
        return B, [O.bf16_rne((rng.standard_normal((m, k)) / np.sqrt(k)).astype(np.float32)) for m, k in zip(Ms, Ks)]
    B = rng.integers(-127, 128, (K1, N)).astype(np.int8)
    lo, hi = (-127, 128) if full_int8 else (-3, 4)
    return B, [rng.integers(lo, hi, (m, k)).astype(np.int8) for m, k in zip(Ms, Ks)]


# ------------------------------------------------------------ single GEMM --

@pytest.mark.parametrize("M,N,K", [(64, 64, 64), (128, 128, 256), (130, 70, 50), (1, 33, 17), (257, 129, 96),
                                   # small M or N: the 64 x 64 split-K path, aligned and ragged
                                   (1000, 64, 2000), (64, 300, 1024), (37, 45, 5001), (200, 8, 777)])
def test_sgemm_fp32_mfma_vs_oracle(gpu, M, N, K):
    import torch
    print(f"sgemm {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C = empty((M, N), torch.float32)
    fg.sgemm(C, dev(A), dev(B), M, N, K)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert_rel(got, O.sgemm(A, B), TOL_F32)
    assert_rel(got, A.astype(np.float64) @ B.astype(np.float64), TOL_F32)


@pytest.mark.parametrize("f32_mode", [fg.F32_MFMA, fg.F32_BF16X3])
def test_sgemm_large_tile_vs_f64(gpu, f32_mode):
    """Enough 256 x 256 tiles for the large-tile kernel, ragged in M and N."""
    import torch
    M, N, K = 4000, 3500, 200
    print(f"sgemm large tile {M}x{N}x{K} mode={f32_mode}", flush=True)
    rng = np.random.default_rng(41)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    C = empty((M, N), torch.float32)
    fg.sgemm(C, dev(A), dev(B), M, N, K, f32_mode=f32_mode)
    torch.cuda.synchronize()
    tol = TOL_F32 if f32_mode == fg.F32_MFMA else TOL_X3
    assert_rel(C.cpu().numpy(), A.astype(np.float64) @ B.astype(np.float64), tol)


@pytest.mark.parametrize("dtype", [fg.DTYPE_INT8, fg.DTYPE_BF16])
def test_unfused_chain_large_tile(gpu, dtype):
    """A 2000 x 6000 first GEMM runs on the large-tile kernel (unfused chain)."""
    print(f"unfused large-tile chain dtype={dtype}", flush=True)
    rng = np.random.default_rng(43)
    B, As = rand_chain(rng, dtype, 300, 6000, [2000, 96])
    got = run_chain(dtype, B, As, fuse=fg.FUSE_OFF)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, O.chain_int8(B, As, O.LAYOUT_MATH))
    else:
        assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)


@pytest.mark.parametrize("M,N,K,splits", [(64, 64, 4096, 8), (200, 40, 3000, 5), (64, 64, 65536, 0)])
def test_sgemm_split_k_deterministic(gpu, M, N, K, splits):
    import torch
    print(f"sgemm split-K {M}x{N}x{K} splits={splits}", flush=True)
    rng = np.random.default_rng(K)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    dA, dB = dev(A), dev(B)
    outs = []
    for _ in range(2):
        C = empty((M, N), torch.float32)
        fg.sgemm(C, dA, dB, M, N, K, splits=splits)
        torch.cuda.synchronize()
        outs.append(C.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))  # fixed-order reduction
    assert_rel(outs[0], A.astype(np.float64) @ B.astype(np.float64), TOL_F32)


def test_sgemm_split_bf16x3_vs_f64(gpu):
    import torch
    rng = np.random.default_rng(11)
    M, N, K = 192, 160, 512
    A = (rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-3, 3, (1, K))).astype(np.float32)
    B = (rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-3, 3, (1, N))).astype(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    C = empty((M, N), torch.float32)
    fg.sgemm(C, dev(A), dev(B), M, N, K, f32_mode=fg.F32_BF16X3)
    torch.cuda.synchronize()
    err_x3 = O.rel_fro(C.cpu().numpy(), ref)
    C2 = empty((M, N), torch.float32)
    fg.sgemm(C2, dev(A), dev(B), M, N, K)
    torch.cuda.synchronize()
    err_f32 = O.rel_fro(C2.cpu().numpy(), ref)
    print(f"bf16x3 err {err_x3:.3e}  fp32 mfma err {err_f32:.3e}")
    assert err_x3 <= TOL_X3
    assert err_f32 <= TOL_F32


@pytest.mark.parametrize("name", golden_names("sgemm"))
def test_sgemm_golden_through_reference_entry_point(gpu, name):
    d = load_golden(name)
    A, B = d["A"], d["B"]
    M, K = A.shape
    N = B.shape[1]
    C = np.zeros((M, N), np.float32)
    fg.LibShalom_set_thread_nums(d["threads"])
    fg.LibShalom_sgemm_mp(0, 0, C, A, B, M, N, K)
    assert np.max(np.abs(C - d["C"])) <= 1e-3  # the reference's own check
    assert_rel(C, d["C"], TOL_F32)


# ------------------------------------------------------------------ chains --

@pytest.mark.parametrize("name", golden_names("int8") + golden_names("bf16") + golden_names("f32"))
def test_chain_golden_through_reference_entry_point(gpu, name):
    """The reference's outputs, reproduced through the reference's own C entry
    points (host pointers, flashgemm_thread_num = the fixture's thread count)."""
    d = load_golden(name)
    B, As = d["B"], d["As"]
    N = B.shape[1]
    triples = []
    for a in As:
        triples += [a, a.shape[0], a.shape[1]]
    C = np.zeros_like(d["C"])
    fg.flashgemm_set_thread_num(d["threads"])
    try:
        fn = {"f32": fg.flashgemm_multi_f32f32f32, "bf16": fg.flashgemm_multi_bf16bf16f32,
              "int8": fg.flashgemm_multi_int8uint8int32}[d["dtype"]]
        fn(C, B, N, len(As), *triples)
    finally:
        fg.flashgemm_set_thread_num(1)
    if d["dtype"] == "int8":
        assert np.array_equal(C, d["C"]), f"{(C != d['C']).sum()} of {C.size} differ"
    elif d["dtype"] == "bf16":
        assert_rel(C, d["C"], TOL_BF16)
    else:
        assert_rel(C, d["C"], TOL_F32)


def expect_unfusable(dtype, K1, N, Ms, fuse):
    """When the fused path cannot take a shape (arena of (K_i + M_i) rows beyond LDS,
    N % 4 != 0, or a K_i row not a whole number of 16-byte groups), asking for it
    must fail with FG_ERR_UNSUPPORTED rather than run something else. Returns True
    when that was checked (nothing more to test for this case)."""
    if fuse != fg.FUSE_ON or fg.chain_plan(dtype, N, Ms, [K1] + Ms[:-1])[0]:
        return False
    with pytest.raises(fg.FlashgemmError) as e:
        run_chain(dtype, *rand_chain(np.random.default_rng(0), dtype, K1, N, Ms), fuse=fuse)
    assert e.value.status == -7
    return True


CHAIN_SHAPES = [
    (64, 64, [32]),
    (64, 200, [96, 64]),                 # ragged N
    (96, 333, [48, 130, 20]),            # M, K not multiples of 32
    (512, 1024, [512, 1024, 512]),       # the 3-chain widths, narrow N
    (40, 17, [24, 36, 12, 16]),          # 4 GEMMs, tiny N
]


@pytest.mark.parametrize("K1,N,Ms", CHAIN_SHAPES)
@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
def test_int8_chain_bit_exact_math_layout(gpu, K1, N, Ms, fuse):
    print(f"int8 chain K1={K1} N={N} Ms={Ms} fuse={fuse}", flush=True)
    if expect_unfusable(fg.DTYPE_INT8, K1, N, Ms, fuse):
        return
    rng = np.random.default_rng(K1 + N)
    B, As = rand_chain(rng, fg.DTYPE_INT8, K1, N, Ms)
    got = run_chain(fg.DTYPE_INT8, B, As, fuse=fuse)
    want = O.chain_int8(B, As, O.LAYOUT_MATH)
    assert np.array_equal(got, want), f"{(got != want).sum()} of {got.size} differ"


@pytest.mark.parametrize("tile_n", [32, 64, 96, 128, 256])
def test_int8_chain_every_tile_width(gpu, tile_n):
    rng = np.random.default_rng(tile_n)
    B, As = rand_chain(rng, fg.DTYPE_INT8, 128, 300, [64, 96, 48], full_int8=True)
    if tile_n == 96:  # INT8 tiles: 32 and 64 columns on every wave, 128 and 256 split over the waves
        with pytest.raises(fg.FlashgemmError) as e:
            run_chain(fg.DTYPE_INT8, B, As, fuse=fg.FUSE_ON, tile_n=tile_n)
        assert e.value.status == fg.ERR_UNSUPPORTED
        return
    got = run_chain(fg.DTYPE_INT8, B, As, fuse=fg.FUSE_ON, tile_n=tile_n)
    assert np.array_equal(got, O.chain_int8(B, As, O.LAYOUT_MATH))


@pytest.mark.parametrize("f32_mode", [fg.F32_MFMA, fg.F32_BF16X3])
@pytest.mark.parametrize("tile_n", [32, 64, 96, 128, 256])
def test_f32_chain_every_tile_width(gpu, tile_n, f32_mode):
    rng = np.random.default_rng(tile_n + 7)
    B, As = rand_chain(rng, fg.DTYPE_F32, 100, 300, [72, 40, 136])
    if not fg.chain_plan(fg.DTYPE_F32, 300, [72, 40, 136], [100, 72, 40],
                         fg.opts(tile_n=tile_n, f32_mode=f32_mode))[0]:
        # 256 fp32 columns take 4 KB of LDS per row: this chain's arena does not fit
        with pytest.raises(fg.FlashgemmError) as e:
            run_chain(fg.DTYPE_F32, B, As, fuse=fg.FUSE_ON, tile_n=tile_n, f32_mode=f32_mode)
        assert e.value.status == fg.ERR_UNSUPPORTED
        return
    got = run_chain(fg.DTYPE_F32, B, As, fuse=fg.FUSE_ON, tile_n=tile_n, f32_mode=f32_mode)
    tol = TOL_F32 if f32_mode == fg.F32_MFMA else TOL_X3
    assert_rel(got, O.chain_f32(B, As), tol)


@pytest.mark.parametrize("K1,N,Ms", CHAIN_SHAPES)
@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
def test_bf16_chain_math_layout(gpu, K1, N, Ms, fuse):
    print(f"bf16 chain K1={K1} N={N} Ms={Ms} fuse={fuse}", flush=True)
    if expect_unfusable(fg.DTYPE_BF16, K1, N, Ms, fuse):
        return
    rng = np.random.default_rng(K1 * 3 + N)
    B, As = rand_chain(rng, fg.DTYPE_BF16, K1, N, Ms)
    got = run_chain(fg.DTYPE_BF16, B, As, fuse=fuse)
    assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)


@pytest.mark.parametrize("K1,N,Ms", CHAIN_SHAPES)
@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
def test_f32_chain(gpu, K1, N, Ms, fuse):
    if expect_unfusable(fg.DTYPE_F32, K1, N, Ms, fuse):
        return
    rng = np.random.default_rng(K1 * 5 + N)
    B, As = rand_chain(rng, fg.DTYPE_F32, K1, N, Ms)
    got = run_chain(fg.DTYPE_F32, B, As, fuse=fuse)
    assert_rel(got, O.chain_f32(B, As), TOL_F32)


@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
def test_f32_chain_split_bf16x3(gpu, fuse):
    rng = np.random.default_rng(5)
    B, As = rand_chain(rng, fg.DTYPE_F32, 256, 512, [256, 256])
    got = run_chain(fg.DTYPE_F32, B, As, fuse=fuse, f32_mode=fg.F32_BF16X3)
    ref = O.chain_f64(B.astype(np.float64), [a.astype(np.float64) for a in As])
    assert_rel(got, ref, TOL_X3)


@pytest.mark.parametrize("dtype,f32_mode", [(fg.DTYPE_INT8, fg.F32_MFMA), (fg.DTYPE_BF16, fg.F32_MFMA),
                                            (fg.DTYPE_F32, fg.F32_MFMA), (fg.DTYPE_F32, fg.F32_BF16X3)])
def test_prepacked_weights_same_bits(gpu, dtype, f32_mode):
    """Weights packed once (fg_pack_weight) and passed with a_packed=1 give the same bits
    as the call that packs them itself; a shape the fused path cannot take is refused."""
    import torch
    print(f"prepacked dtype={dtype} f32_mode={f32_mode}", flush=True)
    rng = np.random.default_rng(10 * dtype + f32_mode)
    K1, N, Ms = 96, 300, [130, 48, 72]
    Ks = [K1] + Ms[:-1]
    B, As = rand_chain(rng, dtype, K1, N, Ms, full_int8=True)
    want = run_chain(dtype, B, As, fuse=fg.FUSE_ON, f32_mode=f32_mode)
    dA = [dev(a) for a in As]
    packed = [empty((fg.packed_weight_bytes(dtype, m, k, f32_mode),), torch.uint8) for m, k in zip(Ms, Ks)]
    for a, m, k, p in zip(dA, Ms, Ks, packed):
        fg.pack_weight(dtype, a, m, k, p, f32_mode=f32_mode)
    C = empty((Ms[-1], N), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
    dB = dev(B)
    fg.chain(dtype, C, dB, packed, N=N, Ms=Ms, Ks=Ks, o=fg.opts(f32_mode=f32_mode, a_packed=1))
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    assert np.array_equal(got, want), f"{(got != want).sum()} of {got.size} differ"
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, O.chain_int8(B, As, O.LAYOUT_MATH))
    # N % 4 != 0 has no fused form, and packed weights have no other consumer
    with pytest.raises(fg.FlashgemmError) as e:
        fg.chain(dtype, C, dB, packed, N=N - 1, Ms=Ms, Ks=Ks, o=fg.opts(f32_mode=f32_mode, a_packed=1))
    assert e.value.status == fg.ERR_UNSUPPORTED
    torch.cuda.synchronize()


@pytest.mark.parametrize("tile_n", [32, 64, 96, 128, 256])
@pytest.mark.parametrize("K1,N,Ms", [(64, 1024, [128, 64, 96]), (256, 768, [256, 256]), (128, 512, [64]),
                                     (64, 2048, [64, 64, 32])])
def test_f32_row_major_weights_same_bits(gpu, K1, N, Ms, tile_n):
    """FP32 chains whose weights need no padding (every M_i % 32 == 0, every K_i % 64 ==
    0) skip the per-call pack: the fused kernel reads the row-major weights in fragment
    order, across row passes and GEMM boundaries (K differs from GEMM to GEMM here).
    Same bits as the same chain on weights packed beforehand, at every tile width the
    shape fits (the last shape fits all of them); where it does not, both calls refuse."""
    import torch
    print(f"f32 row-major weights K1={K1} N={N} Ms={Ms} tile_n={tile_n}", flush=True)
    rng = np.random.default_rng(K1 + N + tile_n)
    Ks = [K1] + Ms[:-1]
    B, As = rand_chain(rng, fg.DTYPE_F32, K1, N, Ms)
    packed = [empty((fg.packed_weight_bytes(fg.DTYPE_F32, m, k),), torch.uint8) for m, k in zip(Ms, Ks)]
    for a, m, k, p in zip(As, Ms, Ks, packed):
        fg.pack_weight(fg.DTYPE_F32, dev(a), m, k, p)
    C = empty((Ms[-1], N), torch.float32)
    try:
        got = run_chain(fg.DTYPE_F32, B, As, fuse=fg.FUSE_ON, tile_n=tile_n)
    except fg.FlashgemmError as e:
        assert e.status == fg.ERR_UNSUPPORTED and Ms != [64, 64, 32]
        with pytest.raises(fg.FlashgemmError):
            fg.chain(fg.DTYPE_F32, C, dev(B), packed, N=N, Ms=Ms, Ks=Ks, o=fg.opts(tile_n=tile_n, a_packed=1))
        return
    fg.chain(fg.DTYPE_F32, C, dev(B), packed, N=N, Ms=Ms, Ks=Ks, o=fg.opts(tile_n=tile_n, a_packed=1))
    torch.cuda.synchronize()
    want = C.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"{(got != want).sum()} of {got.size} differ"
    assert_rel(got, O.chain_f32(B, As), TOL_F32)


@pytest.mark.parametrize("dtype,K1,Ms", [(fg.DTYPE_F32, 64, [128, 64, 96]), (fg.DTYPE_BF16, 128, [256, 64])])
def test_wide_tile_arena_whole_chunks(gpu, dtype, K1, Ms):
    """256-column tiles walk each GEMM in passes shorter than a chunk; when their final
    passes are held in registers the arena bound Kc + (passes - 1) * pass is not a whole
    number of chunks, and an arena of that size let a chunk of X run past its end. The
    arena is a whole number of chunks: these chains then need more than the 160 KB of
    LDS at 256 columns, so the forced width is refused, and the planner's own choice
    matches the oracle."""
    print(f"wide-tile arena dtype={dtype} K1={K1} Ms={Ms}", flush=True)
    rng = np.random.default_rng(K1)
    B, As = rand_chain(rng, dtype, K1, 1024, Ms)
    with pytest.raises(fg.FlashgemmError) as e:
        run_chain(dtype, B, As, fuse=fg.FUSE_ON, tile_n=256)
    assert e.value.status == fg.ERR_UNSUPPORTED
    got = run_chain(dtype, B, As)
    if dtype == fg.DTYPE_F32:
        assert_rel(got, O.chain_f32(B, As), TOL_F32)
    else:
        assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_MATH), TOL_BF16)


@pytest.mark.parametrize("threads,N,Ms", [(1, 96, [24, 36, 12]), (4, 160, [12, 24, 36]), (2, 64, [36, 48, 24, 12]),
                                          (3, 96, [24, 12])])
def test_int8_reference_layout_device_api(gpu, threads, N, Ms):
    rng = np.random.default_rng(threads * 100 + N)
    B, As = rand_chain(rng, fg.DTYPE_INT8, 64, N, Ms)
    got = run_chain(fg.DTYPE_INT8, B, As, layout=fg.LAYOUT_REFERENCE, ref_threads=threads)
    assert np.array_equal(got, O.chain_int8(B, As, O.LAYOUT_REFERENCE, threads))


@pytest.mark.parametrize("threads,N,Ms", [(1, 96, [24, 36, 12]), (4, 160, [20, 24, 12]), (2, 128, [28, 20, 16])])
def test_bf16_reference_layout_device_api(gpu, threads, N, Ms):
    rng = np.random.default_rng(threads * 10 + N)
    B, As = rand_chain(rng, fg.DTYPE_BF16, 64, N, Ms)
    got = run_chain(fg.DTYPE_BF16, B, As, layout=fg.LAYOUT_REFERENCE, ref_threads=threads)
    assert_rel(got, O.chain_bf16(B, As, O.LAYOUT_REFERENCE, threads), TOL_BF16)


@pytest.mark.parametrize("dtype", [fg.DTYPE_INT8, fg.DTYPE_BF16])
@pytest.mark.parametrize("tile_n", [32, 64, 128, 256])
@pytest.mark.parametrize("threads", [1, 2, 4, 3])
def test_reference_layout_fused(gpu, dtype, tile_n, threads):
    """The reference's intermediate layout on the fused kernel: with 32-aligned panels
    (1, 2, 4 threads over N = 512) and intermediates of whole 32-row blocks, the fused
    chain writes the permuted intermediates itself; INT8 equals the unfused gather path
    and the oracle bit for bit, BF16 the oracle within TOL_BF16. With 3 threads the
    panels are 176 wide (a 16-wide edge), so only the unfused path applies."""
    print(f"reference layout fused dtype={dtype} tile={tile_n} threads={threads}", flush=True)
    rng = np.random.default_rng(tile_n * 10 + threads + dtype)
    K1, N, Ms = 64, 512, [96, 64, 128]
    Ks = [K1] + Ms[:-1]
    B, As = rand_chain(rng, dtype, K1, N, Ms, full_int8=True)
    kw = dict(layout=fg.LAYOUT_REFERENCE, ref_threads=threads)
    fused, _ = fg.chain_plan(dtype, N, Ms, Ks, fg.opts(tile_n=tile_n, **kw))
    want = (O.chain_int8 if dtype == fg.DTYPE_INT8 else O.chain_bf16)(B, As, O.LAYOUT_REFERENCE, threads)
    if threads == 3 or not fused:
        assert threads == 3 or (dtype == fg.DTYPE_BF16 and tile_n == 256)
        with pytest.raises(fg.FlashgemmError) as e:
            run_chain(dtype, B, As, fuse=fg.FUSE_ON, tile_n=tile_n, **kw)
        assert e.value.status == fg.ERR_UNSUPPORTED
        got = run_chain(dtype, B, As, **kw)
    else:
        got = run_chain(dtype, B, As, fuse=fg.FUSE_ON, tile_n=tile_n, **kw)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, want), f"{(got != want).sum()} of {got.size} differ"
        assert np.array_equal(got, run_chain(dtype, B, As, fuse=fg.FUSE_OFF, **kw))
    else:
        assert_rel(got, want, TOL_BF16)


def test_reference_layout_rejects_undefined_panels(gpu):
    rng = np.random.default_rng(0)
    B, As = rand_chain(rng, fg.DTYPE_INT8, 64, 40, [24])  # N=40 -> a panel of width 8
    with pytest.raises(fg.FlashgemmError) as e:
        run_chain(fg.DTYPE_INT8, B, As, layout=fg.LAYOUT_REFERENCE, ref_threads=1)
    assert e.value.status == -2


def test_invalid_chain_shapes_rejected(gpu):
    import torch
    C = empty((8, 8), torch.float32)
    B = empty((8, 8), torch.float32)
    A = empty((8, 4), torch.float32)  # K=4 but B has 8 rows: mismatch with M=8 of A1
    with pytest.raises(fg.FlashgemmError):
        fg.chain(fg.DTYPE_F32, C, B, [B, A], N=8, Ms=[8, 8], Ks=[8, 4])


def test_legacy_odd_k_prints_and_leaves_c(gpu, capfd):
    A = np.ones((12, 3), np.float32)
    B = np.ones((3, 32), np.float32)
    C = np.full((12, 32), 7.0, np.float32)
    fg.flashgemm_multi_f32f32f32(C, B, 32, 1, A, 12, 3)
    out = capfd.readouterr().out
    assert "K must be a multiple of 2" in out and "Dimension error for gemm 1" in out
    assert np.all(C == 7.0)


# ------------------------------------------------------- full-size checks --

def test_int8_three_chain_full_size_properties(gpu):
    """The 16384-wide 3-chain (512 -> 512 -> 1024 -> 512): fused and unfused kernels
    agree bit for bit over the whole output, and sampled columns equal the oracle
    (in the mathematical layout every column of C depends only on its column of B)."""
    import torch
    rng = np.random.default_rng(2024)
    N, K1, Ms = 16384, 512, [512, 1024, 512]
    B, As = rand_chain(rng, fg.DTYPE_INT8, K1, N, Ms, full_int8=True)
    As = [rng.integers(-2, 3, a.shape).astype(np.int8) for a in As]  # keep intermediates unsaturated
    fused = run_chain(fg.DTYPE_INT8, B, As, fuse=fg.FUSE_ON)
    spill = run_chain(fg.DTYPE_INT8, B, As, fuse=fg.FUSE_OFF)
    assert np.array_equal(fused, spill)
    cols = np.sort(rng.choice(N, 96, replace=False))
    want = O.chain_int8(np.ascontiguousarray(B[:, cols]), As, O.LAYOUT_MATH)
    assert np.array_equal(fused[:, cols], want)


def test_measure_stream_entry_point(gpu):
    """The stream-rate measurement runs (a positive time) and rejects bad sizes."""
    assert fg.measure_stream(32, 3, 8) > 0
    with pytest.raises(fg.FlashgemmError) as e:
        fg.measure_stream(0, 3, 8)
    assert e.value.status == fg.ERR_INVALID
