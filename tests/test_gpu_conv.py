"""GPU parity of the convolution front end (B formed from an NCHW image).

Bars, stated per test:
  * fg_im2col: bit-exact against the oracle's gather for every dtype;
  * fused conv chains, INT8: bit-exact against the oracle chain on im2col(X);
  * fused conv chains, FP32 / BF16: bit-identical to fg_chain on the materialized
    patch matrix (the same kernel after staging, so the same arithmetic), and
    within the chain tolerances of the oracle (FP32 relative Frobenius <= 1e-5,
    BF16 <= 5e-4);
  * unfused and reference-layout conv chains: the same bars through the
    materialized path.
"""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O

from test_gpu_parity import assert_rel, TOL_BF16, TOL_F32, dev, empty

pytestmark = pytest.mark.gpu

GEOMS = {  # batch, channels, H, W, kh, kw, stride, pad, dilation
    "3x3_pad1": (1, 8, 12, 12, 3, 3, (1, 1), (1, 1), (1, 1)),
    "5x5_s2_b2": (2, 6, 15, 13, 5, 5, (2, 2), (2, 2), (1, 1)),
    "3x2_dil_odd_n": (1, 5, 9, 7, 3, 2, (1, 2), (1, 0), (2, 1)),
    "1x1": (2, 32, 8, 8, 1, 1, (1, 1), (0, 0), (1, 1)),
}


def geom_of(t):
    nb, C, H, W, kh, kw, s, p, d = t
    return fg.conv_geom(nb, C, H, W, kh, kw, stride=s[0], stride_w=s[1], pad=p[0], pad_w=p[1], dilation=d[0],
                        dilation_w=d[1])


def image(rng, dtype, shape):
    if dtype == fg.DTYPE_F32:
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if dtype == fg.DTYPE_BF16:
        return O.bf16_rne(rng.standard_normal(shape).astype(np.float32))
    return rng.integers(-127, 128, shape).astype(np.int8)


def weights(rng, dtype, Ms, K1):
    Ks = [K1] + Ms[:-1]
    if dtype == fg.DTYPE_F32:
        return [rng.uniform(-1, 1, (m, k)).astype(np.float32) for m, k in zip(Ms, Ks)]
    if dtype == fg.DTYPE_BF16:
        return [O.bf16_rne((rng.standard_normal((m, k)) / np.sqrt(k)).astype(np.float32)) for m, k in zip(Ms, Ks)]
    return [rng.integers(-3, 4, (m, k)).astype(np.int8) for m, k in zip(Ms, Ks)]


def oracle_chain(dtype, B, As, layout=O.LAYOUT_MATH, threads=1):
    if dtype == fg.DTYPE_F32:
        return O.chain_f32(B, As)
    if dtype == fg.DTYPE_BF16:
        return O.chain_bf16(B, As, layout, threads)
    return O.chain_int8(B, As, layout, threads)


def out_dtype(dtype):
    import torch
    return torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32


def run_conv(dtype, X, g, As, **kw):
    import torch
    K1, N = fg.conv_dims(g)
    C = empty((As[-1].shape[0], N), out_dtype(dtype))
    fg.chain_conv(dtype, C, dev(X), g, [dev(a) for a in As], Ms=[a.shape[0] for a in As],
                  Ks=[a.shape[1] for a in As], o=fg.opts(**kw))
    torch.cuda.synchronize()
    return C.cpu().numpy()


def run_matrix(dtype, B, As, **kw):
    import torch
    C = empty((As[-1].shape[0], B.shape[1]), out_dtype(dtype))
    fg.chain(dtype, C, dev(B), [dev(a) for a in As], N=B.shape[1], Ms=[a.shape[0] for a in As],
             Ks=[a.shape[1] for a in As], o=fg.opts(**kw))
    torch.cuda.synchronize()
    return C.cpu().numpy()


@pytest.mark.parametrize("dtype", [fg.DTYPE_F32, fg.DTYPE_BF16, fg.DTYPE_INT8])
@pytest.mark.parametrize("name", sorted(GEOMS))
def test_im2col_kernel(gpu, dtype, name):
    import torch
    t = GEOMS[name]
    g = geom_of(t)
    X = image(np.random.default_rng(3), dtype, t[:4])
    want = O.im2col(X, t[4], t[5], t[6], t[7], t[8])
    K1, N = fg.conv_dims(g)
    assert (K1, N) == want.shape
    tdt = {fg.DTYPE_F32: torch.float32, fg.DTYPE_BF16: torch.int16, fg.DTYPE_INT8: torch.int8}[dtype]
    B = empty((K1, N), tdt)
    fg.im2col(dtype, B, dev(X), g)
    torch.cuda.synchronize()
    got = B.cpu().numpy()
    assert np.array_equal(got.view(want.dtype), want)


@pytest.mark.parametrize("dtype", [fg.DTYPE_F32, fg.DTYPE_BF16, fg.DTYPE_INT8])
@pytest.mark.parametrize("name", sorted(GEOMS))
def test_fused_conv_chain(gpu, dtype, name):
    t = GEOMS[name]
    g = geom_of(t)
    rng = np.random.default_rng(4)
    X = image(rng, dtype, t[:4])
    K1, N = fg.conv_dims(g)
    Ms = [48, 64, 24]
    As = weights(rng, dtype, Ms, K1)
    B = O.im2col(X, t[4], t[5], t[6], t[7], t[8])
    got = run_conv(dtype, X, g, As, fuse=fg.FUSE_ON)
    want = oracle_chain(dtype, B, As)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, want)
        return
    tol = TOL_F32 if dtype == fg.DTYPE_F32 else TOL_BF16
    assert_rel(got, want, tol)
    if N % 4 == 0:  # the matrix chain fuses too: identical arithmetic after staging
        assert np.array_equal(got, run_matrix(dtype, B, As, fuse=fg.FUSE_ON))


@pytest.mark.parametrize("tile_n", [32, 64])
def test_fused_conv_chain_int8_tiles_tails(gpu, tile_n):
    # N = 2 * 11 * 9 = 198: a partial last tile and a partial 4-column group
    g = fg.conv_geom(2, 7, 11, 9, 3, 3, pad=1)
    rng = np.random.default_rng(5)
    X = image(rng, fg.DTYPE_INT8, (2, 7, 11, 9))
    K1, N = fg.conv_dims(g)
    assert N == 198
    As = [rng.integers(-127, 128, (m, k)).astype(np.int8) for m, k in ((40, K1), (72, 40))]
    want = O.chain_int8(O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1)), As, O.LAYOUT_MATH)
    assert np.array_equal(run_conv(fg.DTYPE_INT8, X, g, As, fuse=fg.FUSE_ON, tile_n=tile_n), want)


@pytest.mark.parametrize("dtype", [fg.DTYPE_INT8, fg.DTYPE_BF16])
def test_fused_conv_chain_unaligned_image(gpu, dtype):
    # an image that is not dword aligned takes the element-wise gather instead of the
    # dword one: same bits as the aligned image
    import torch
    g = fg.conv_geom(1, 4, 16, 16, 3, 3, pad=1)
    rng = np.random.default_rng(9)
    X = image(rng, dtype, (1, 4, 16, 16))
    K1, N = fg.conv_dims(g)
    As = weights(rng, dtype, [32, 16], K1)
    tdt = torch.int8 if dtype == fg.DTYPE_INT8 else torch.int16
    flat = dev(X.reshape(-1))
    raw = torch.zeros(flat.numel() + 1, dtype=tdt, device="cuda")
    raw[1:] = flat  # one element in: 1 byte (INT8) or 2 bytes (BF16) past a dword
    C = empty((16, N), out_dtype(dtype))
    fg.chain_conv(dtype, C, raw.data_ptr() + raw.element_size(), g, [dev(a) for a in As],
                  o=fg.opts(fuse=fg.FUSE_ON))
    torch.cuda.synchronize()
    aligned = run_conv(dtype, X, g, As, fuse=fg.FUSE_ON)
    assert np.array_equal(C.cpu().numpy(), aligned)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(aligned, O.chain_int8(O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1)), As, O.LAYOUT_MATH))


@pytest.mark.parametrize("dtype", [fg.DTYPE_F32, fg.DTYPE_INT8])
def test_unfused_conv_chain(gpu, dtype):
    t = GEOMS["5x5_s2_b2"]
    g = geom_of(t)
    rng = np.random.default_rng(6)
    X = image(rng, dtype, t[:4])
    K1, N = fg.conv_dims(g)
    As = weights(rng, dtype, [32, 16], K1)
    B = O.im2col(X, t[4], t[5], t[6], t[7], t[8])
    got = run_conv(dtype, X, g, As, fuse=fg.FUSE_OFF)
    want = oracle_chain(dtype, B, As)
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(got, want)
    else:
        assert_rel(got, want, TOL_F32)


def test_reference_layout_conv_chain(gpu):
    # the reference layout runs unfused on the formed patch matrix: same bits as fg_chain on it
    g = fg.conv_geom(1, 4, 16, 16, 3, 3, pad=1)  # K1 = 36, N = 256
    rng = np.random.default_rng(7)
    X = image(rng, fg.DTYPE_INT8, (1, 4, 16, 16))
    As = [rng.integers(-127, 128, (m, k)).astype(np.int8) for m, k in ((32, 36), (16, 32))]
    B = O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1))
    got = run_conv(fg.DTYPE_INT8, X, g, As, layout=fg.LAYOUT_REFERENCE, ref_threads=4)
    assert np.array_equal(got, O.chain_int8(B, As, O.LAYOUT_REFERENCE, 4))


@pytest.mark.parametrize("dtype", [fg.DTYPE_INT8, fg.DTYPE_BF16])
def test_conv_chain_host_form(gpu, dtype):
    # fg_chain_conv_host (host arrays, blocking) gives the device call's bits
    t = GEOMS["5x5_s2_b2"]
    g = geom_of(t)
    rng = np.random.default_rng(10)
    X = image(rng, dtype, t[:4])
    K1, N = fg.conv_dims(g)
    As = weights(rng, dtype, [40, 24], K1)
    C = np.empty((24, N), np.int32 if dtype == fg.DTYPE_INT8 else np.float32)
    fg.chain_conv_host(dtype, C, np.ascontiguousarray(X), g, [np.ascontiguousarray(a) for a in As])
    assert np.array_equal(C, run_conv(dtype, X, g, As))
    if dtype == fg.DTYPE_INT8:
        assert np.array_equal(C, O.chain_int8(O.im2col(X, t[4], t[5], t[6], t[7], t[8]), As, O.LAYOUT_MATH))


def test_conv_chain_rejects_wrong_k1(gpu):
    g = fg.conv_geom(1, 4, 8, 8, 3, 3, pad=1)  # K1 = 36
    rng = np.random.default_rng(8)
    X = image(rng, fg.DTYPE_INT8, (1, 4, 8, 8))
    As = [rng.integers(-3, 4, (16, 32)).astype(np.int8)]
    with pytest.raises(fg.FlashgemmError) as e:
        run_conv(fg.DTYPE_INT8, X, g, As)
    assert e.value.status == fg.ERR_INVALID
