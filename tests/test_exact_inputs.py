"""CPU checks of the exact-arithmetic chain inputs (tests/exact_inputs.py) used by the
GPU bit-exactness tests of the FP32 and BF16 chains, and of the oracle on them.

* The partial-sum bound stays below 2^24 for every GEMM, so the fp32 arithmetic of
  any summation order is exact on these inputs.
* The oracle's BF16 and FP32 chains equal the exact integer computation bit for bit
  (an independent restatement: numpy int64 products and sums, RNE by its definition).
* A truncating bf16 intermediate (the planted mistake) changes nearly every output,
  so a bit-exact comparison catches it; the oracle's truncating conversion
  (oracle_bf16_trunc, the reference's float_to_bf16, multigemm/src/utils.h:20-33) is
  pinned by its definition on textbook cases and random floats.
"""
import numpy as np
import pytest

import oracle as O
from exact_inputs import EXACT_LIMIT, bf16_round, exact_chain, integer_chain, to_bf16_bits


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_exact_bounds(seed):
    B, As = integer_chain(np.random.default_rng(seed), 256)
    for rounding in (None, "rne", "trunc"):
        _, bounds = exact_chain(B, As, rounding)
        assert max(bounds) < EXACT_LIMIT, (rounding, bounds)


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_equals_exact_integer_chain(seed):
    rng = np.random.default_rng(seed)
    B, As = integer_chain(rng, 256)
    want_bf16, _ = exact_chain(B, As, "rne")
    got = O.chain_bf16(to_bf16_bits(B), [to_bf16_bits(a) for a in As], O.LAYOUT_MATH)
    assert np.array_equal(got, want_bf16.astype(np.float32))
    want_f32, _ = exact_chain(B, As, None)
    got32 = O.chain_f32(B.astype(np.float32), [a.astype(np.float32) for a in As])
    assert np.array_equal(got32, want_f32.astype(np.float32))


def test_truncating_intermediate_is_caught():
    """The planted mistake: the same chain with truncated bf16 intermediates differs
    from the RNE chain in almost every output element."""
    B, As = integer_chain(np.random.default_rng(0), 256)
    rne, _ = exact_chain(B, As, "rne")
    trunc, _ = exact_chain(B, As, "trunc")
    assert (rne != trunc).mean() > 0.9
    # the first intermediate: about half its elements are not bf16-representable, and
    # truncation and RNE disagree on about a quarter of them
    y1 = As[0] @ B
    assert (bf16_round(y1, "rne") != y1).mean() > 0.3
    assert (bf16_round(y1, "rne") != bf16_round(y1, "trunc")).mean() > 0.15


def test_oracle_bf16_trunc_pinned():
    # textbook cases of "drop the low 16 bits" against round to nearest even
    one_and_3ulp_half = 1.0 + 3 * 2.0 ** -9       # 1.0058593750: halfway-plus, rne rounds up
    cases = [(1.0, 0x3F80, 0x3F80), (-2.5, 0xC020, 0xC020), (one_and_3ulp_half, 0x3F80, 0x3F81),
             (-one_and_3ulp_half, 0xBF80, 0xBF81), (1.0 + 2.0 ** -8, 0x3F80, 0x3F80),  # tie -> even
             (np.float32(3.4028235e38), 0x7F7F, 0x7F80)]                          # rne overflows to inf
    for x, trunc, rne in cases:
        assert O.bf16_trunc_scalar(x) == trunc, x
        assert O.bf16_rne_scalar(x) == rne, x
    rng = np.random.default_rng(3)
    xs = (rng.standard_normal(2000) * 10.0 ** rng.uniform(-30, 30, 2000)).astype(np.float32)
    want = (xs.view(np.uint32) >> 16).astype(np.uint16)
    assert all(O.bf16_trunc_scalar(x) == int(w) for x, w in zip(xs, want))
