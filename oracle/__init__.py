"""CPU oracle for flashgemm: ctypes bindings over oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY. Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this package, as the checker; the product library
(``flashgemm_amd``) never does.

Each wrapper takes numpy arrays in the reference orientation (B is K1 x N, each
A_i is M_i x K_i, row-major) and returns C (M_n x N). The C routines restate the
reference arithmetic; see oracle/oracle.h for the citations and pins.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LAYOUT_MATH = 0
LAYOUT_REFERENCE = 1

_lib = None


def build(force: bool = False) -> None:
    """Compile liboracle.so with oracle/Makefile."""
    args = ["make", "-s", "-C", HERE, "all"]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, lp, i, l = ctypes.c_void_p, ctypes.POINTER(ctypes.c_long), ctypes.c_int, ctypes.c_long
        for name in ("oracle_chain_f32", "oracle_chain_f64"):
            getattr(L, name).argtypes = [vp, vp, l, i, ctypes.POINTER(vp), lp, lp]
            getattr(L, name).restype = i
        for name in ("oracle_chain_bf16", "oracle_chain_int8"):
            getattr(L, name).argtypes = [vp, vp, l, i, ctypes.POINTER(vp), lp, lp, i, i]
            getattr(L, name).restype = i
        L.oracle_sgemm.argtypes = [vp, vp, vp, l, l, l]
        L.oracle_sgemm.restype = i
        L.oracle_bf16_rne.argtypes = [ctypes.c_float]
        L.oracle_bf16_rne.restype = ctypes.c_uint16
        L.oracle_bf16_trunc.argtypes = [ctypes.c_float]
        L.oracle_bf16_trunc.restype = ctypes.c_uint16
        L.oracle_partition.argtypes = [l, i, lp, lp, lp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _chain_args(As):
    n = len(As)
    ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in As])
    Ms = (ctypes.c_long * n)(*[a.shape[0] for a in As])
    Ks = (ctypes.c_long * n)(*[a.shape[1] for a in As])
    return n, ptrs, Ms, Ks


def _check_rc(rc: int, what: str) -> None:
    if rc == -1:
        raise ValueError(f"{what}: invalid shapes")
    if rc == -2:
        raise ValueError(f"{what}: outside the reference layout domain (panel width or row group)")
    if rc != 0:
        raise RuntimeError(f"{what}: rc={rc}")


def chain_f32(B: np.ndarray, As) -> np.ndarray:
    B = np.ascontiguousarray(B, np.float32)
    As = [np.ascontiguousarray(a, np.float32) for a in As]
    N = B.shape[1]
    C = np.empty((As[-1].shape[0], N), np.float32)
    n, ptrs, Ms, Ks = _chain_args(As)
    _check_rc(lib().oracle_chain_f32(_ptr(C), _ptr(B), N, n, ptrs, Ms, Ks), "chain_f32")
    return C


def chain_f64(B: np.ndarray, As) -> np.ndarray:
    B = np.ascontiguousarray(B, np.float64)
    As = [np.ascontiguousarray(a, np.float64) for a in As]
    N = B.shape[1]
    C = np.empty((As[-1].shape[0], N), np.float64)
    n, ptrs, Ms, Ks = _chain_args(As)
    _check_rc(lib().oracle_chain_f64(_ptr(C), _ptr(B), N, n, ptrs, Ms, Ks), "chain_f64")
    return C


def chain_bf16(B: np.ndarray, As, layout: int = LAYOUT_MATH, nthreads: int = 1) -> np.ndarray:
    """B and A_i hold raw bf16 bits (uint16)."""
    B = np.ascontiguousarray(B, np.uint16)
    As = [np.ascontiguousarray(a, np.uint16) for a in As]
    N = B.shape[1]
    C = np.empty((As[-1].shape[0], N), np.float32)
    n, ptrs, Ms, Ks = _chain_args(As)
    _check_rc(lib().oracle_chain_bf16(_ptr(C), _ptr(B), N, n, ptrs, Ms, Ks, layout, nthreads),
              "chain_bf16")
    return C


def chain_int8(B: np.ndarray, As, layout: int = LAYOUT_MATH, nthreads: int = 1) -> np.ndarray:
    """B holds bytes read as signed, A_i bytes read as unsigned (the reference's u8.s8)."""
    B = np.ascontiguousarray(B).view(np.int8)
    As = [np.ascontiguousarray(a).view(np.int8) for a in As]
    N = B.shape[1]
    C = np.empty((As[-1].shape[0], N), np.int32)
    n, ptrs, Ms, Ks = _chain_args(As)
    _check_rc(lib().oracle_chain_int8(_ptr(C), _ptr(B), N, n, ptrs, Ms, Ks, layout, nthreads),
              "chain_int8")
    return C


def sgemm(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    C = np.empty((M, N), np.float32)
    _check_rc(lib().oracle_sgemm(_ptr(C), _ptr(A), _ptr(B), M, N, K), "sgemm")
    return C


def bf16_rne(x: np.ndarray) -> np.ndarray:
    """Vectorised restatement of oracle_bf16_rne (RNE, denormal -> signed zero, qNaN)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).copy()
    den = (u & 0x7F800000) == 0
    u[den] &= 0x80000000
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    out = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    out[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return out


def bf16_trunc_scalar(x: float) -> int:
    """oracle_bf16_trunc: the reference generator's float_to_bf16 (multigemm/src/utils.h:20-33),
    which keeps sign, exponent and the top 7 mantissa bits (truncation)."""
    return int(lib().oracle_bf16_trunc(float(x)))


def bf16_rne_scalar(x: float) -> int:
    return int(lib().oracle_bf16_rne(float(x)))


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def partition(N: int, nthreads: int):
    nb, ne, num = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
    lib().oracle_partition(N, nthreads, ctypes.byref(nb), ctypes.byref(ne), ctypes.byref(num))
    return nb.value, ne.value, num.value


def rel_fro(x: np.ndarray, ref: np.ndarray) -> float:
    """Relative Frobenius error ||x - ref||_F / ||ref||_F in float64."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))


def im2col(X: np.ndarray, kernel_h: int, kernel_w: int, stride=(1, 1), pad=(0, 0), dilation=(1, 1)) -> np.ndarray:
    """Patch matrix of an NCHW image: B[k][n] = X[b][c][iy][ix], zero outside the image.

    k = (c * kernel_h + ky) * kernel_w + kx, n = (b * OH + oy) * OW + ox,
    iy = oy * stride_h - pad_h + ky * dilation_h, ix = ox * stride_w - pad_w + kx * dilation_w.
    This front end has no reference counterpart: the reference's callers form B
    themselves (its conv-shaped tests, multigemm/test/test_f32_3.c:14-48, pass a
    random B). The row order is torch's unfold order; tests pin it against
    torch.nn.functional.unfold and a direct convolution. Any dtype (a pure gather).
    """
    X = np.asarray(X)
    nb, C, H, W = X.shape
    sh, sw = stride
    ph, pw = pad
    dh, dw = dilation
    OH = (H + 2 * ph - dh * (kernel_h - 1) - 1) // sh + 1
    OW = (W + 2 * pw - dw * (kernel_w - 1) - 1) // sw + 1
    Xp = np.zeros((nb, C, H + 2 * ph, W + 2 * pw), X.dtype)
    Xp[:, :, ph:ph + H, pw:pw + W] = X
    B = np.empty((C, kernel_h, kernel_w, nb, OH, OW), X.dtype)
    for ky in range(kernel_h):
        for kx in range(kernel_w):
            y0, x0 = ky * dh, kx * dw
            win = Xp[:, :, y0:y0 + sh * (OH - 1) + 1:sh, x0:x0 + sw * (OW - 1) + 1:sw]  # nb, C, OH, OW
            B[:, ky, kx] = win.transpose(1, 0, 2, 3)
    return B.reshape(C * kernel_h * kernel_w, nb * OH * OW)


def _wrap32(a: np.ndarray) -> np.ndarray:
    return ((np.asarray(a, np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)).astype(np.int64)


def epi_int8(acc: np.ndarray, bias=None, scale=None, zero_point: int = 0, relu: bool = False) -> np.ndarray:
    """INT8 intermediate with the epilogue extensions (no reference counterpart; the
    reference only saturates, multigemm/src/kernel_int8.h:329-385). Row m of acc is
    output channel m: a = acc + bias[m] (int32 wrap); with a scale, t = rne(fp32(a) *
    fp32 scale[m]) (half to even), else t = a; t is clamped to [-512, 512] (which
    cannot change the result); q = sat8(t + zero_point); ReLU: max(q, zero_point)."""
    a = _wrap32(np.asarray(acc, np.int64) + (0 if bias is None else np.asarray(bias, np.int64)[:, None]))
    if scale is not None:
        t = a.astype(np.float32) * np.asarray(scale, np.float32)[:, None]
        t = np.rint(np.clip(t, np.float32(-512), np.float32(512))).astype(np.int64)
    else:
        t = np.clip(a, -512, 512)
    q = np.clip(t + zero_point, -128, 127)
    if relu:
        q = np.maximum(q, zero_point)
    return q.astype(np.int8)


def epi_float(acc: np.ndarray, bias=None, scale=None, relu: bool = False) -> np.ndarray:
    """FP32 / BF16 epilogue extensions on fp32 values: y = acc * scale[m], then
    y + bias[m] (two fp32 roundings), then max(y, 0) under ReLU (NaN -> 0)."""
    y = np.asarray(acc, np.float32)
    if scale is not None:
        y = (y * np.asarray(scale, np.float32)[:, None]).astype(np.float32)
    if bias is not None:
        y = (y + np.asarray(bias, np.float32)[:, None]).astype(np.float32)
    if relu:
        y = np.where(y > 0, y, np.float32(0)).astype(np.float32)
    return y


def chain_int8_epi(B: np.ndarray, As, epi: dict) -> np.ndarray:
    """INT8 chain (mathematical layout) with epilogue extensions. epi holds per-GEMM
    lists "bias", "scale", "zero_point", "relu" (None / 0 entries: not set). The GEMMs
    are exact integer products u8(A) . s8(X) in int64, wrapped to int32; the last GEMM
    gives int32 C = acc + bias[m], ReLU max(C, 0)."""
    X = np.asarray(B).view(np.int8).astype(np.int64)
    n = len(As)
    get = lambda k, i, d=None: (epi.get(k) or [])[i] if i < len(epi.get(k) or []) else d
    for i, A in enumerate(As):
        acc = _wrap32(np.asarray(A).view(np.uint8).astype(np.int64) @ X)
        b = get("bias", i)
        if i == n - 1:
            c = _wrap32(acc + (0 if b is None else np.asarray(b, np.int64)[:, None]))
            if get("relu", i, 0):
                c = np.maximum(c, 0)
            return c.astype(np.int32)
        X = epi_int8(acc, b, get("scale", i), int(get("zero_point", i, 0) or 0), bool(get("relu", i, 0))).astype(np.int64)
    raise ValueError("empty chain")


def chain_float_epi(B: np.ndarray, As, epi: dict, bf16: bool) -> np.ndarray:
    """FP32 or BF16 chain (mathematical layout) with epilogue extensions, for
    tolerance checks: each GEMM in float64 rounded to fp32 (the accumulator), the
    extensions in fp32 (epi_float), BF16 intermediates rounded with bf16_rne."""
    conv = (lambda x: bf16_to_f32(x).astype(np.float64)) if bf16 else (lambda x: np.asarray(x, np.float64))
    X = conv(B)
    n = len(As)
    get = lambda k, i, d=None: (epi.get(k) or [])[i] if i < len(epi.get(k) or []) else d
    for i, A in enumerate(As):
        acc = (conv(A) @ X).astype(np.float32)
        y = epi_float(acc, get("bias", i), get("scale", i), bool(get("relu", i, 0)))
        if i == n - 1:
            return y
        X = bf16_to_f32(bf16_rne(y)).astype(np.float64) if bf16 else y.astype(np.float64)
    raise ValueError("empty chain")
