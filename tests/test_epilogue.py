"""Epilogue extensions in the oracle (CPU), pinned without retyping their formulas:

* every extension unset gives the reference's chain (the C restatement, itself
  pinned by the reference's golden vectors);
* a per-row int32 bias equals one extra k-row of ones in B with the bias as an
  extra u8 column of A, computed by the C restatement;
* scaled requantization rounds half to even: checked value by value against
  Python's round() on exactly representable quotients;
* ReLU with a zero point: outputs never fall below it and a second ReLU is a no-op.
"""
import numpy as np
import pytest

import oracle as O


def rand_int8(rng, K1, N, Ms, lo=-127, hi=128):
    Ks = [K1] + Ms[:-1]
    return rng.integers(-127, 128, (K1, N)).astype(np.int8), \
        [rng.integers(lo, hi, (m, k)).astype(np.int8) for m, k in zip(Ms, Ks)]


def test_defaults_are_the_reference_chain():
    rng = np.random.default_rng(1)
    B, As = rand_int8(rng, 96, 72, [64, 80, 40])
    want = O.chain_int8(B, As, O.LAYOUT_MATH)
    assert np.array_equal(O.chain_int8_epi(B, As, {}), want)
    assert np.array_equal(O.chain_int8_epi(B, As, {"bias": [None] * 3, "zero_point": [0, 0, 0], "relu": [0, 0, 0]}),
                          want)


def test_bias_is_an_extra_k_row():
    rng = np.random.default_rng(2)
    B, As = rand_int8(rng, 64, 50, [32, 24], lo=-3, hi=4)
    bias = [rng.integers(0, 256, 32), rng.integers(0, 256, 24)]
    got = O.chain_int8_epi(B, As, {"bias": bias})
    # the same chain with bias folded into the GEMMs: B and each X gain a row of ones
    # (s8 1) and A_i a u8 column holding the bias; K grows by one per GEMM
    Bx = np.vstack([B, np.ones((1, B.shape[1]), np.int8)])
    A1 = np.hstack([As[0], bias[0].astype(np.uint8).view(np.int8)[:, None]])
    Y1 = O.chain_int8(Bx, [A1], O.LAYOUT_MATH)  # int32, single GEMM: no requantization
    X1 = np.clip(Y1, -128, 127).astype(np.int8)
    X1x = np.vstack([X1, np.ones((1, X1.shape[1]), np.int8)])
    A2 = np.hstack([As[1], bias[1].astype(np.uint8).view(np.int8)[:, None]])
    want = O.chain_int8(X1x, [A2], O.LAYOUT_MATH)
    assert np.array_equal(got, want)


def test_scale_rounds_half_to_even():
    acc = np.array([[-7, -6, -5, -3, -1, 0, 1, 3, 5, 6, 7, 255, 257, -255]], np.int32)
    for sc in (0.5, 0.25, 0.125, 1.0, 2.0):
        got = O.epi_int8(acc, scale=np.array([sc], np.float32))
        want = [max(-128, min(127, round(int(a) * sc))) for a in acc[0]]  # Python round: half to even
        assert got[0].tolist() == want, sc


def test_scale_zero_point_saturation():
    acc = np.array([[100000, -100000, 2 ** 31 - 1, -(2 ** 31), 37]], np.int32)
    got = O.epi_int8(acc, scale=np.array([0.5], np.float32), zero_point=-20)
    assert got[0].tolist() == [127, -128, 127, -128, round(18.5) - 20]


@pytest.mark.parametrize("zp", [-128, -5, 0, 17, 127])
def test_relu_floor_is_the_zero_point(zp):
    rng = np.random.default_rng(3)
    acc = rng.integers(-2000, 2000, (8, 50)).astype(np.int32)
    sc = rng.uniform(0.01, 0.2, 8).astype(np.float32)
    q = O.epi_int8(acc, scale=sc, zero_point=zp, relu=True)
    assert q.min() >= zp
    assert np.array_equal(np.maximum(q.astype(np.int64), zp), q)
    plain = O.epi_int8(acc, scale=sc, zero_point=zp)
    assert np.array_equal(q, np.maximum(plain, zp).astype(np.int8))


def test_float_epilogue_order():
    # acc * scale first, then + bias: with scale 3 and bias -1 the ReLU boundary sits at 1/3
    acc = np.array([[0.0, 0.25, 1.0 / 3.0, 0.5]], np.float32)
    y = O.epi_float(acc, bias=np.array([-1.0], np.float32), scale=np.array([3.0], np.float32), relu=True)
    assert y[0, 0] == 0 and y[0, 1] == 0 and y[0, 3] == np.float32(0.5)
    assert np.isnan(O.epi_float(np.array([[np.nan]], np.float32))[0, 0])
    assert O.epi_float(np.array([[np.nan]], np.float32), relu=True)[0, 0] == 0
