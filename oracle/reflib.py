"""ctypes bindings over the compiled reference (oracle/_ref/, built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: used to write golden fixtures (tests/golden/) in a
container that holds the reference sources, and by bench.py's cpu_baseline leg.
Call signatures are the reference's own: multigemm/src/flashgemm.h:19-25 and the
singlegemm dispatcher restated in oracle/ref_singlegemm.cpp.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
MULTI = os.path.join(REF_DIR, "libref_multigemm.so")
SINGLE = os.path.join(REF_DIR, "libref_singlegemm.so")


def available() -> bool:
    return os.path.exists(MULTI) and os.path.exists(SINGLE)


class Reference:
    """One loaded copy of the reference libraries (the thread count is process-global)."""

    def __init__(self, threads: int = 1):
        self.multi = ctypes.CDLL(MULTI)
        self.single = ctypes.CDLL(SINGLE)
        self.multi.flashgemm_set_thread_num.argtypes = [ctypes.c_int]
        self.single.ref_singlegemm_set_threads.argtypes = [ctypes.c_int]
        self.single.ref_sgemm.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_long] * 3
        self.set_threads(threads)

    def set_threads(self, threads: int) -> None:
        self.threads = threads
        self.multi.flashgemm_set_thread_num(threads)
        self.single.ref_singlegemm_set_threads(threads)

    def _chain(self, fn, C, B, As):
        args = [ctypes.c_void_p(C.ctypes.data), ctypes.c_void_p(B.ctypes.data),
                ctypes.c_long(B.shape[1]), ctypes.c_int(len(As))]
        for a in As:
            args += [ctypes.c_void_p(a.ctypes.data), ctypes.c_long(a.shape[0]),
                     ctypes.c_long(a.shape[1])]
        fn(*args)
        return C

    def chain_f32(self, B, As):
        B = np.ascontiguousarray(B, np.float32)
        As = [np.ascontiguousarray(a, np.float32) for a in As]
        C = np.zeros((As[-1].shape[0], B.shape[1]), np.float32)
        return self._chain(self.multi.flashgemm_multi_f32f32f32, C, B, As)

    def chain_bf16(self, B, As):
        B = np.ascontiguousarray(B, np.uint16)
        As = [np.ascontiguousarray(a, np.uint16) for a in As]
        C = np.zeros((As[-1].shape[0], B.shape[1]), np.float32)
        return self._chain(self.multi.flashgemm_multi_bf16bf16f32, C, B, As)

    def chain_int8(self, B, As):
        B = np.ascontiguousarray(B).view(np.uint8)
        As = [np.ascontiguousarray(a).view(np.int8) for a in As]
        C = np.zeros((As[-1].shape[0], B.shape[1]), np.int32)
        return self._chain(self.multi.flashgemm_multi_int8uint8int32, C, B, As)

    def sgemm(self, A, B):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        C = np.zeros((A.shape[0], B.shape[1]), np.float32)
        self.single.ref_sgemm(C.ctypes.data, A.ctypes.data, B.ctypes.data,
                              A.shape[0], B.shape[1], A.shape[1])
        return C
