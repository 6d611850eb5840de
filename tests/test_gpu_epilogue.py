"""GPU parity of the epilogue extensions (per-row bias and scale, INT8 zero point,
ReLU) on the fused and unfused chain paths.

Bars: INT8 bit-exact against the oracle (chain_int8_epi); FP32 relative Frobenius
<= 1e-5 and BF16 <= 5e-4 against chain_float_epi; all extensions unset gives the
plain chain's bits.
"""
import numpy as np
import pytest

import flashgemm_amd as fg
import oracle as O

from test_gpu_parity import assert_rel, TOL_BF16, TOL_F32, dev, empty

pytestmark = pytest.mark.gpu


def int8_case(rng, K1, N, Ms):
    Ks = [K1] + Ms[:-1]
    B = rng.integers(-127, 128, (K1, N)).astype(np.int8)
    As = [rng.integers(-127, 128, (m, k)).astype(np.int8) for m, k in zip(Ms, Ks)]
    n = len(Ms)
    # scales that keep the intermediates mostly unsaturated: ~1/(K * 64 * 128) per channel
    epi = {"bias": [rng.integers(-20000, 20000, m).astype(np.int32) for m in Ms],
           "scale": [rng.uniform(0.5, 2.0, m).astype(np.float32) / (k * 40.0) for m, k in zip(Ms[:-1], Ks[:-1])]
           + [None],
           "zero_point": [int(z) for z in rng.integers(-30, 30, n - 1)] + [0],
           "relu": [1] * (n - 1) + [1]}
    return B, As, epi


def to_fg(epi):
    dv = lambda x: None if x is None else dev(np.ascontiguousarray(x))
    keep = {"bias": [dv(b) for b in epi.get("bias", [])], "scale": [dv(s) for s in epi.get("scale", [])]}
    e = fg.epilogue(bias=keep["bias"], scale=keep["scale"], zero_point=epi.get("zero_point"), relu=epi.get("relu"))
    return e, keep  # keep the device buffers alive


def run(dtype, B, As, epi, **kw):
    import torch
    C = empty((As[-1].shape[0], B.shape[1]), torch.int32 if dtype == fg.DTYPE_INT8 else torch.float32)
    e, keep = to_fg(epi)
    fg.chain_epi(dtype, C, dev(B), [dev(a) for a in As], e, N=B.shape[1], Ms=[a.shape[0] for a in As],
                 Ks=[a.shape[1] for a in As], o=fg.opts(**kw))
    torch.cuda.synchronize()
    del keep
    return C.cpu().numpy()


@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
@pytest.mark.parametrize("shape", [(512, 1024, [256, 512, 128]), (200, 300, [72, 40, 96])])
def test_int8_epilogue(gpu, fuse, shape):
    K1, N, Ms = shape
    rng = np.random.default_rng(10 + K1)
    B, As, epi = int8_case(rng, K1, N, Ms)
    got = run(fg.DTYPE_INT8, B, As, epi, fuse=fuse)
    want = O.chain_int8_epi(B, As, epi)
    assert np.array_equal(got, want)


def test_int8_scaled_intermediates_are_not_saturated(gpu):
    # the reference's sat-only rule clips almost every intermediate at these
    # magnitudes; with the per-channel scale most of them stay inside int8
    rng = np.random.default_rng(11)
    B, As, epi = int8_case(rng, 512, 512, [256, 128])
    acc = np.asarray(As[0]).view(np.uint8).astype(np.int64) @ B.astype(np.int64)
    sat_plain = np.mean(np.abs(acc) >= 127)
    q = O.epi_int8(acc.astype(np.int32), epi["bias"][0], epi["scale"][0], epi["zero_point"][0], False)
    sat_scaled = np.mean((q == 127) | (q == -128))
    assert sat_plain > 0.9 and sat_scaled < 0.1
    assert np.array_equal(run(fg.DTYPE_INT8, B, As, epi, fuse=fg.FUSE_ON), O.chain_int8_epi(B, As, epi))


def test_int8_unset_epilogue_is_the_plain_chain(gpu):
    rng = np.random.default_rng(12)
    B, As, _ = int8_case(rng, 256, 512, [128, 64])
    import torch
    plain = empty((64, 512), torch.int32)
    fg.chain(fg.DTYPE_INT8, plain, dev(B), [dev(a) for a in As])
    torch.cuda.synchronize()
    assert np.array_equal(run(fg.DTYPE_INT8, B, As, {}), plain.cpu().numpy())


def float_case(rng, dtype, K1, N, Ms):
    Ks = [K1] + Ms[:-1]
    if dtype == fg.DTYPE_BF16:
        B = O.bf16_rne(rng.standard_normal((K1, N)).astype(np.float32))
        As = [O.bf16_rne((rng.standard_normal((m, k)) / np.sqrt(k)).astype(np.float32)) for m, k in zip(Ms, Ks)]
    else:
        B = rng.uniform(-1, 1, (K1, N)).astype(np.float32)
        As = [(rng.uniform(-1, 1, (m, k)) / np.sqrt(k)).astype(np.float32) for m, k in zip(Ms, Ks)]
    epi = {"bias": [rng.uniform(-0.5, 0.5, m).astype(np.float32) for m in Ms],
           "scale": [rng.uniform(0.5, 2.0, m).astype(np.float32) for m in Ms],
           "relu": [1] * (len(Ms) - 1) + [0]}
    return B, As, epi


@pytest.mark.parametrize("fuse", [fg.FUSE_ON, fg.FUSE_OFF])
@pytest.mark.parametrize("dtype", [fg.DTYPE_F32, fg.DTYPE_BF16])
def test_float_epilogue(gpu, dtype, fuse):
    rng = np.random.default_rng(13 + dtype)
    B, As, epi = float_case(rng, dtype, 256, 768, [192, 256, 96])
    got = run(dtype, B, As, epi, fuse=fuse)
    want = O.chain_float_epi(B, As, epi, bf16=dtype == fg.DTYPE_BF16)
    assert_rel(got, want, (TOL_F32 if dtype == fg.DTYPE_F32 else TOL_BF16))


def test_conv_chain_with_epilogue(gpu):
    import torch
    g = fg.conv_geom(1, 8, 14, 14, 3, 3, pad=1)
    K1, N = fg.conv_dims(g)
    rng = np.random.default_rng(14)
    X = rng.integers(-127, 128, (1, 8, 14, 14)).astype(np.int8)
    _, As, epi = int8_case(rng, K1, N, [48, 32])
    e, keep = to_fg(epi)
    C = empty((32, N), torch.int32)
    fg.chain_conv(fg.DTYPE_INT8, C, dev(X), g, [dev(a) for a in As], o=fg.opts(fuse=fg.FUSE_ON), epi=e)
    torch.cuda.synchronize()
    want = O.chain_int8_epi(O.im2col(X, 3, 3, (1, 1), (1, 1), (1, 1)), As, epi)
    assert np.array_equal(C.cpu().numpy(), want)


def test_epilogue_rejects(gpu):
    rng = np.random.default_rng(15)
    B, As, _ = int8_case(rng, 64, 64, [32, 16])
    with pytest.raises(fg.FlashgemmError) as e:
        run(fg.DTYPE_INT8, B, As, {"zero_point": [200, 0]})
    assert e.value.status == fg.ERR_INVALID
    with pytest.raises(fg.FlashgemmError) as e:
        run(fg.DTYPE_INT8, B, As, {"scale": [None, np.ones(16, np.float32)]})
    assert e.value.status == fg.ERR_UNSUPPORTED
    Bf, Af, _ = float_case(rng, fg.DTYPE_F32, 64, 64, [32, 16])
    with pytest.raises(fg.FlashgemmError) as e:
        run(fg.DTYPE_F32, Bf, Af, {"zero_point": [3, 0]})
    assert e.value.status == fg.ERR_UNSUPPORTED
