"""The compiled reference's INT8 domain (CPU, needs oracle/_ref built from the reference
sources; skipped without it): its INT8 GEMM computes its own algorithm only on whole
12-row blocks of A. Rows past the last full block read misaligned A data, so parity
with the reference is defined for M_i % 12 == 0 and, elsewhere, against the oracle's
restatement of the algorithm. The reference runs in a child process: outside its
domain it reads memory it does not own."""
import json
import os
import subprocess
import sys

import pytest

from oracle import reflib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle as O
from oracle.reflib import Reference
ref = Reference(1)
rng = np.random.default_rng(12)
out = {}
for M in (12, 16, 24, 36, 40, 512):
    B = rng.integers(-128, 128, (64, 32)).astype(np.int8)
    A = rng.integers(-3, 4, (M, 64)).astype(np.int8)
    got, want = ref.chain_int8(B, [A]), O.chain_int8(B, [A], O.LAYOUT_REFERENCE, 1)
    out[M] = [int(r) for r in np.nonzero((got != want).any(axis=1))[0]]
print(json.dumps(out))
"""


@pytest.mark.skipif(not reflib.available(), reason="oracle/_ref (the compiled reference) is not built")
def test_int8_reference_domain_is_whole_12_row_blocks():
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    bad = {int(k): v for k, v in json.loads(r.stdout.strip().splitlines()[-1]).items()}
    print(bad)
    for M, rows in bad.items():
        full = M // 12 * 12
        # every row of a whole 12-row block agrees with the restated algorithm
        assert all(r >= full for r in rows), (M, rows)
        if M % 12 == 0:
            assert rows == [], (M, rows)
    # and the tail rows of a partial block do not (the measured defect), e.g. rows
    # 504-511 of M = 512
    assert bad[16] and bad[40] and bad[512]
    assert set(bad[512]) <= set(range(504, 512))
