"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

* golden fixtures written by the compiled reference (tests/golden/gen_golden.py):
  every oracle routine must reproduce them bit for bit, including the reference's
  INT8/BF16 intermediate layout at 1..4 threads with 32- and 16-wide panels;
* brute force in pure Python on tiny shapes (INT8 u8.s8, FP32 sequential fma
  order, the BF16 pair order);
* closed forms: ones-chains, saturation boundaries, textbook RNE cases;
* a planted mistake (a transposed intermediate) must fail the golden check.
"""
import math

import numpy as np
import pytest

from conftest import golden_names, load_golden

import oracle as O


def run_oracle(d, layout=O.LAYOUT_REFERENCE):
    if d["dtype"] == "f32":
        return O.chain_f32(d["B"], d["As"])
    if d["dtype"] == "bf16":
        return O.chain_bf16(d["B"], d["As"], layout, d["threads"])
    if d["dtype"] == "int8":
        return O.chain_int8(d["B"], d["As"], layout, d["threads"])
    return O.sgemm(d["A"], d["B"])


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(oracle_lib, name):
    d = load_golden(name)
    got = run_oracle(d)
    assert got.dtype == d["C"].dtype
    assert np.array_equal(got.view(np.uint32), d["C"].view(np.uint32)), name


def test_planted_transpose_fails_golden(oracle_lib):
    """Reading the intermediate transposed within each panel must break parity."""
    d = load_golden("int8_chain3_T1")
    B, (A1, A2, A3) = d["B"], d["As"]
    Y1 = O.chain_int8(B, [A1], O.LAYOUT_MATH)
    Z1 = np.clip(Y1, -128, 127).astype(np.int8)
    wrong = Z1.reshape(Z1.shape[0], -1, 32)[:, :, ::-1].reshape(Z1.shape)  # mirrored panels
    C = O.chain_int8(wrong, [A2, A3], O.LAYOUT_REFERENCE, 1)
    assert not np.array_equal(C, d["C"])


def test_math_layout_differs_from_reference_layout_for_int8(oracle_lib):
    d = load_golden("int8_chain3_T1")
    math_c = O.chain_int8(d["B"], d["As"], O.LAYOUT_MATH)
    assert not np.array_equal(math_c, d["C"])


def test_f32_chain_identical_in_both_orders(oracle_lib):
    """The reference's FP32 chain is the mathematical chain: it equals chaining
    single GEMMs of the oracle."""
    d = load_golden("f32_chain3_T3")
    X = d["B"]
    for A in d["As"]:
        X = O.chain_f32(X, [A])
    assert np.array_equal(X, d["C"])


def _py_seq_fma(a_row, b_col):
    acc = np.float32(0)
    for a, b in zip(a_row, b_col):
        acc = np.float32(math.fma(float(a), float(b), float(acc))) if hasattr(math, "fma") else \
            np.float32(np.float64(a) * np.float64(b) + np.float64(acc))
    return acc


def test_f32_sequential_fma_brute_force(oracle_lib):
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, (3, 40)).astype(np.float32)
    B = rng.uniform(-1, 1, (40, 5)).astype(np.float32)
    C = O.chain_f32(B, [A])
    for m in range(3):
        for n in range(5):
            # fp32 fma = exact product + add, rounded once: float64 holds the exact
            # product of two floats and the sum of it with a float rounds correctly
            # to fp32 except in double-rounding corner cases, absent for these data.
            ref = _py_seq_fma(A[m], B[:, n])
            assert C[m, n] == ref


def test_int8_brute_force_u8_times_s8(oracle_lib):
    rng = np.random.default_rng(2)
    A = rng.integers(-128, 128, (4, 12)).astype(np.int8)
    B = rng.integers(-128, 128, (12, 3)).astype(np.int8)
    C = O.chain_int8(B, [A])
    for m in range(4):
        for n in range(3):
            s = sum(int(np.uint8(A[m, k].view(np.uint8))) * int(B[k, n]) for k in range(12))
            assert C[m, n] == s


def test_int8_saturation_closed_form(oracle_lib):
    """All-ones chain: Y1 = K1 (>127 saturates to 127), C = 127 * M1 per element."""
    K1, M1, M2, N = 200, 24, 12, 32
    B = np.ones((K1, N), np.int8)
    A1 = np.ones((M1, K1), np.int8)
    A2 = np.ones((M2, M1), np.int8)
    C = O.chain_int8(B, [A1, A2])
    assert np.all(C == 127 * M1)
    # negative side: A bytes 0xFF read as 255, B = -1 -> Y = -255*K -> -128
    A1n = np.full((M1, K1), -1, np.int8)
    Bn = np.full((K1, N), -1, np.int8)
    C = O.chain_int8(Bn, [A1n, A2])
    assert np.all(C == -128 * M1)
    # boundary: Y exactly 127 and 128, -128 and -129
    for y, z in ((127, 127), (128, 127), (-128, -128), (-129, -128)):
        Bv = np.zeros((4, 16), np.int8)
        Bv[0, :] = 1
        A1v = np.zeros((12, 4), np.int8)
        A1v[:, 0] = np.int8(np.uint8(abs(y)).view(np.int8)) if abs(y) < 256 else 0
        if y < 0:
            Bv[0, :] = -1
        A2v = np.eye(12, dtype=np.int8)
        C = O.chain_int8(Bv, [A1v, A2v])
        assert np.all(C == z), (y, z, C[0, 0])


def test_ones_chain_closed_form_f32_bf16(oracle_lib):
    K1, M1, M2, N = 64, 24, 36, 48
    ones = lambda *s: np.ones(s, np.float32)
    C = O.chain_f32(ones(K1, N), [ones(M1, K1), ones(M2, M1)])
    assert np.all(C == K1 * M1)
    one_bf = np.uint16(0x3F80)
    Cb = O.chain_bf16(np.full((K1, N), one_bf), [np.full((M1, K1), one_bf), np.full((M2, M1), one_bf)])
    assert np.all(Cb == K1 * M1)


def test_bf16_rne_textbook_cases(oracle_lib):
    L = O.lib()
    cases = {
        1.0: 0x3F80,
        1.0 + 2 ** -8: 0x3F80,            # tie -> even (down)
        1.0 + 3 * 2 ** -8: 0x3F82,        # tie -> even (up)
        1.0 + 2 ** -8 + 2 ** -20: 0x3F81, # above the tie -> up
        -2.5: 0xC020,
        float("inf"): 0x7F80,
        1e-40: 0x0000,                    # denormal -> +0
        -1e-40: 0x8000,                   # denormal -> -0
    }
    for x, want in cases.items():
        assert L.oracle_bf16_rne(x) == want, (x, hex(L.oracle_bf16_rne(x)))
    assert L.oracle_bf16_rne(float("nan")) & 0x7FC0 == 0x7FC0
    v = np.array(list(cases), np.float32)
    assert list(O.bf16_rne(v)) == [L.oracle_bf16_rne(float(x)) for x in v]


def test_bf16_pair_order(oracle_lib):
    """vdpbf16ps adds the odd k of each pair before the even one: a case where the
    order changes the rounded fp32 result."""
    big, small = np.float32(2 ** 24), np.float32(1.0)
    bf = lambda x: O.bf16_rne(np.array([x], np.float32))[0]
    A = np.array([[bf(small), bf(big), bf(small), bf(-big)]], np.uint16)  # k = 0..3
    B = np.full((4, 1), bf(1.0), np.uint16)
    # odd-first: ((0 + big) + 1) = big (1 lost), then (big - big) + 1 = 1
    C = O.chain_bf16(B, [A])
    assert C[0, 0] == 1.0


def test_partition_matches_reference_formula(oracle_lib):
    assert O.partition(16384, 24) == (688, 16384 % 688, 24)
    assert O.partition(16384, 1) == (16384, 0, 1)
    assert O.partition(160, 4) == (48, 16, 4)
    assert O.partition(32, 3) == (16, 0, 2)


def test_f64_oracle_matches_numpy(oracle_lib):
    rng = np.random.default_rng(4)
    B = rng.standard_normal((24, 40))
    As = [rng.standard_normal((16, 24)), rng.standard_normal((8, 16))]
    np.testing.assert_allclose(O.chain_f64(B, As), As[1] @ (As[0] @ B), rtol=1e-12, atol=1e-12)


def test_sgemm_kblocks_add_in_order(oracle_lib):
    """K > 2048: blocks are summed as C = block0 + block1 (not one long fma chain)."""
    d = load_golden("sgemm_12x16x4096_kblocks")
    A, B = d["A"], d["B"]
    one_chain = O.chain_f32(B, [A])  # single fma chain over all K
    blocked = O.sgemm(A, B)
    assert np.array_equal(blocked, d["C"])
    assert not np.array_equal(one_chain, d["C"])
