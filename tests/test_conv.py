"""Convolution front end on the CPU: the oracle's patch matrix (im2col) against two
independent statements of the same thing, and the C ABI's geometry checks.

* torch.nn.functional.unfold gives the same matrix (row order (c, ky, kx), column
  order (b, oy, ox)) for every geometry below, bit for bit;
* A . im2col(X) equals torch's direct float64 convolution of X with A viewed as
  M x C x kh x kw filters, which pins what the front end is for.
"""
import numpy as np
import pytest

import oracle as O

GEOMS = [  # batch, channels, H, W, kh, kw, stride, pad, dilation
    (1, 3, 7, 9, 3, 3, (1, 1), (1, 1), (1, 1)),
    (2, 4, 10, 8, 5, 3, (2, 1), (2, 0), (1, 2)),
    (1, 1, 5, 5, 1, 1, (1, 1), (0, 0), (1, 1)),
    (3, 2, 6, 11, 3, 2, (3, 2), (1, 1), (2, 1)),
    (1, 5, 4, 4, 4, 4, (1, 1), (3, 3), (1, 1)),  # padding wider than the image
]


@pytest.mark.parametrize("geom", GEOMS)
def test_im2col_matches_unfold(geom):
    import torch
    nb, C, H, W, kh, kw, s, p, d = geom
    X = np.random.default_rng(1).standard_normal((nb, C, H, W))
    B = O.im2col(X, kh, kw, s, p, d)
    U = torch.nn.functional.unfold(torch.from_numpy(X), (kh, kw), dilation=d, padding=p, stride=s)
    U = U.permute(1, 0, 2).reshape(C * kh * kw, -1).numpy()
    assert B.shape == U.shape
    assert np.array_equal(B, U)


@pytest.mark.parametrize("geom", GEOMS[:4])
def test_im2col_gemm_is_convolution(geom):
    import torch
    nb, C, H, W, kh, kw, s, p, d = geom
    rng = np.random.default_rng(2)
    X = rng.standard_normal((nb, C, H, W))
    A = rng.standard_normal((6, C * kh * kw))
    Y = A @ O.im2col(X, kh, kw, s, p, d)  # 6 x (nb * OH * OW)
    ref = torch.nn.functional.conv2d(torch.from_numpy(X), torch.from_numpy(A.reshape(6, C, kh, kw)), stride=s,
                                     padding=p, dilation=d).numpy()  # nb, 6, OH, OW
    assert np.allclose(Y, ref.transpose(1, 0, 2, 3).reshape(6, -1), rtol=1e-12, atol=1e-12)


def test_im2col_integer_bytes():
    # a pure gather: int8 images come out unchanged, padding as zero bytes
    X = np.arange(-60, 60, dtype=np.int8).reshape(1, 2, 6, 10)
    B = O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1))
    assert B.dtype == np.int8 and B.shape == (18, 60)
    assert B[4, 0] == X[0, 0, 0, 0]          # centre tap of the first pixel
    assert B[0, 0] == 0                      # top-left tap of the first pixel is padding
    assert B[13, 59] == X[0, 1, 5, 9]        # centre tap of channel 1, last pixel


@pytest.mark.parametrize("geom", GEOMS)
def test_conv_dims_abi(geom):
    import flashgemm_amd as fg
    nb, C, H, W, kh, kw, s, p, d = geom
    g = fg.conv_geom(nb, C, H, W, kh, kw, stride=s[0], stride_w=s[1], pad=p[0], pad_w=p[1], dilation=d[0],
                     dilation_w=d[1])
    K1, N = fg.conv_dims(g)
    B = O.im2col(np.zeros((nb, C, H, W), np.float32), kh, kw, s, p, d)
    assert (K1, N) == B.shape


@pytest.mark.parametrize("bad", [
    (0, 1, 4, 4, 3, 3, 1, 1, 1),      # no images
    (1, 1, 2, 2, 5, 5, 1, 0, 1),      # kernel larger than the padded image
    (1, 1, 4, 4, 3, 3, 0, 0, 1),      # zero stride
    (1, 1, 4, 4, 3, 3, 1, -1, 1),     # negative padding
    (1, 1, 4, 4, 3, 3, 1, 0, 0),      # zero dilation
    (4096, 512, 1024, 1024, 1, 1, 1, 0, 1),  # 2^41 elements
])
def test_conv_dims_rejects(bad):
    import flashgemm_amd as fg
    g = fg.conv_geom(*bad)
    with pytest.raises(fg.FlashgemmError) as e:
        fg.conv_dims(g)
    assert e.value.status == fg.ERR_INVALID
