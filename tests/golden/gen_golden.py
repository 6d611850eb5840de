"""Write the golden fixtures in tests/golden/ from the compiled reference.

Run where the reference sources exist:

    make -C oracle ref && python tests/golden/gen_golden.py

Each fixture is an .npz holding the inputs (reference orientation: B is K1 x N,
A_i is M_i x K_i, row-major), the reference's output ``C`` and the metadata
(``dtype``, ``threads`` = the reference's flashgemm_thread_num). Every case lies in
the domain where the reference computes its own algorithm; outside it the
reference's kernels leave rows unwritten or read misaligned data, measured as:

  * INT8: every M_i % 12 == 0 (rows past the last full 12-row block are wrong);
  * BF16 with a 16-wide panel: first GEMM M_1 <= 20 and later M_i % 12 == 0;
  * singlegemm: K % 2048 == 8 crashes the reference.

The INT8 bound is also a CPU test (tests/test_reference_domain.py, run when oracle/_ref
is built).

The fixtures depend on the order of the calls below: the reference's singlegemm
keeps its block sizes in process globals that earlier calls set
(singlegemm/src/set_TmTn.h:25-37). ``sgemm_20x16x2336_ktail`` is what the reference
computes after the calls before it in this script; issued first in a fresh process,
the same call crashes the reference. The fixture stays valid as data: the oracle
restates the arithmetic the reference performs in this order (2048-K blocks added in
order, then the K tail), and tests compare the oracle with the saved C.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from oracle.reflib import Reference  # noqa: E402

CHAINS = [
    # name, dtype, Ms, K1, N, threads, value range tag
    ("f32_chain1_T1", "f32", [24], 64, 64, 1, "u"),
    ("f32_chain2_T2_tails", "f32", [20, 28], 32, 96, 2, "u"),
    ("f32_chain3_T3", "f32", [36, 24, 12], 64, 160, 3, "u"),
    ("f32_chain4_T4", "f32", [12, 48, 24, 36], 128, 128, 4, "u"),
    ("bf16_chain1_T1", "bf16", [24], 64, 64, 1, "n"),
    ("bf16_chain3_T1", "bf16", [24, 36, 12], 64, 96, 1, "n"),
    ("bf16_chain3_T2_tails", "bf16", [28, 20, 16], 64, 128, 2, "n"),
    ("bf16_chain3_T4_edge16", "bf16", [20, 24, 12], 64, 160, 4, "n"),
    ("int8_chain1_T1_full", "int8", [24], 64, 64, 1, "full"),
    ("int8_chain3_T1", "int8", [24, 36, 12], 64, 96, 1, "small"),
    ("int8_chain3_T4_edge16", "int8", [12, 24, 36], 64, 160, 4, "small"),
    ("int8_chain4_T2", "int8", [36, 48, 24, 12], 128, 64, 2, "small"),
    ("int8_chain2_T1_saturating", "int8", [36, 24], 64, 64, 1, "full"),
]

SGEMMS = [
    # name, M, N, K, threads
    ("sgemm_64x64x128", 64, 64, 128, 4),
    ("sgemm_12x32x64", 12, 32, 64, 1),
    ("sgemm_12x16x4096_kblocks", 12, 16, 4096, 4),
    ("sgemm_128x16x256_mgn", 128, 16, 256, 4),
    ("sgemm_20x16x2336_ktail", 20, 16, 2336, 2),
]


def make_chain_inputs(rng, dtype, Ms, K1, N, tag):
    Ks = [K1] + Ms[:-1]
    if dtype == "f32":
        B = rng.uniform(-1, 1, (K1, N)).astype(np.float32)
        As = [rng.uniform(-1, 1, (m, k)).astype(np.float32) for m, k in zip(Ms, Ks)]
    elif dtype == "bf16":
        B = O.bf16_rne(rng.standard_normal((K1, N)).astype(np.float32) / 4)
        As = [O.bf16_rne((rng.standard_normal((m, k)) / np.sqrt(k)).astype(np.float32))
              for m, k in zip(Ms, Ks)]
    else:
        B = rng.integers(-128, 128, (K1, N)).astype(np.int8)
        lo, hi = (-128, 128) if tag == "full" else (-3, 4)
        As = [rng.integers(lo, hi, (m, k)).astype(np.int8) for m, k in zip(Ms, Ks)]
    return B, As


def main() -> None:
    ref = Reference(1)
    rng = np.random.default_rng(20241203)
    for name, dtype, Ms, K1, N, threads, tag in CHAINS:
        B, As = make_chain_inputs(rng, dtype, Ms, K1, N, tag)
        ref.set_threads(threads)
        C = {"f32": ref.chain_f32, "bf16": ref.chain_bf16, "int8": ref.chain_int8}[dtype](B, As)
        arrays = {"B": B, "C": C, "dtype": np.array(dtype), "threads": np.array(threads)}
        for i, a in enumerate(As):
            arrays[f"A{i}"] = a
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(f"{name}: C {C.shape} {C.dtype}")
    for name, M, N, K, threads in SGEMMS:
        A = (rng.random((M, K)) - 0.5).astype(np.float32)  # drand48()-0.5 style data
        B = (rng.random((K, N)) - 0.5).astype(np.float32)
        ref.set_threads(threads)
        C = ref.sgemm(A, B)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), A=A, B=B, C=C,
                            dtype=np.array("sgemm"), threads=np.array(threads))
        print(f"{name}: C {C.shape}")


if __name__ == "__main__":
    main()
