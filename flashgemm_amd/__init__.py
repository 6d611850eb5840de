"""flashgemm for MI355X: Python host mirror of the reference's C API.

Everything here is a thin ctypes layer over ``libflashgemm.so`` (built in-tree by
``python -m flashgemm_amd.build``); no computation happens in Python. Importing
works without the library; the first call raises ``RuntimeError`` when the
library is missing, so no path silently falls back to another implementation.

Two layers, mirroring include/flashgemm.h:

* the reference's entry points with the reference's names and argument order
  (multigemm/src/flashgemm.h:17-25, singlegemm/FP32_ID23.c:35-62), on host numpy
  arrays: ``flashgemm_set_thread_num``, ``flashgemm_multi_f32f32f32``,
  ``flashgemm_multi_bf16bf16f32``, ``flashgemm_multi_int8uint8int32``,
  ``LibShalom_set_thread_nums``, ``LibShalom_sgemm_mp``;
* the device API (``chain``, ``chain_epi``, ``chain_conv``, ``im2col``, ``sgemm``,
  ``sgemm_t``, ``f64_chain``, ``error_norms`` ...) on device buffers given as torch
  tensors or raw addresses, enqueued on a stream.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflashgemm.so")

DTYPE_F32, DTYPE_BF16, DTYPE_INT8 = 0, 1, 2
F32_MFMA, F32_BF16X3, F32_BF16X6 = 0, 1, 2
LAYOUT_MATH, LAYOUT_REFERENCE = 0, 1
FUSE_AUTO, FUSE_OFF, FUSE_ON = -1, 0, 1
(OK, ERR_INVALID, ERR_DOMAIN, ERR_NOMEM, ERR_HIP, ERR_BUDGET, ERR_NODEVICE,
 ERR_UNSUPPORTED) = 0, -1, -2, -3, -4, -5, -6, -7

STATUS = {0: "FG_OK", -1: "FG_ERR_INVALID", -2: "FG_ERR_DOMAIN", -3: "FG_ERR_NOMEM", -4: "FG_ERR_HIP",
          -5: "FG_ERR_BUDGET", -6: "FG_ERR_NODEVICE", -7: "FG_ERR_UNSUPPORTED"}


class FlashgemmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class ChainOpts(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_int), ("ref_threads", ctypes.c_int), ("f32_mode", ctypes.c_int),
                ("fuse", ctypes.c_int), ("tile_n", ctypes.c_int), ("splits", ctypes.c_int),
                ("a_packed", ctypes.c_int)]


class ChainReq(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("f32_mode", ctypes.c_int), ("C", ctypes.c_void_p),
                ("B", ctypes.c_void_p), ("N", ctypes.c_long), ("num_gemm", ctypes.c_int),
                ("A", ctypes.c_void_p * 8), ("M", ctypes.c_long * 8), ("K", ctypes.c_long * 8),
                ("a_packed", ctypes.c_int)]


class ConvGeom(ctypes.Structure):
    """fg_conv_geom: an NCHW image and the patches a convolution reads from it."""
    _fields_ = [(f, ctypes.c_long) for f in ("batch", "channels", "height", "width", "kernel_h", "kernel_w",
                                             "stride_h", "stride_w", "pad_h", "pad_w", "dilation_h",
                                             "dilation_w")]


def conv_geom(batch: int, channels: int, height: int, width: int, kernel_h: int, kernel_w: int, stride: int = 1,
              pad: int = 0, dilation: int = 1, stride_w: int | None = None, pad_w: int | None = None,
              dilation_w: int | None = None) -> ConvGeom:
    return ConvGeom(batch, channels, height, width, kernel_h, kernel_w, stride,
                    stride if stride_w is None else stride_w, pad, pad if pad_w is None else pad_w, dilation,
                    dilation if dilation_w is None else dilation_w)


class Epilogue(ctypes.Structure):
    """fg_epilogue: per-GEMM bias / scale (device pointers), INT8 zero point, ReLU."""
    _fields_ = [("bias", ctypes.c_void_p * 8), ("scale", ctypes.c_void_p * 8), ("zero_point", ctypes.c_int * 8),
                ("relu", ctypes.c_int * 8)]


def epilogue(bias=None, scale=None, zero_point=None, relu=None) -> Epilogue:
    """Build an fg_epilogue from per-GEMM lists (None entries: not set). bias and scale
    entries are device buffers (torch tensors or addresses)."""
    e = Epilogue()
    for i, b in enumerate(bias or []):
        e.bias[i] = _addr(b) if b is not None else None
    for i, sc in enumerate(scale or []):
        e.scale[i] = _addr(sc) if sc is not None else None
    for i, z in enumerate(zero_point or []):
        e.zero_point[i] = int(z)
    for i, r in enumerate(relu or []):
        e.relu[i] = int(bool(r))
    return e


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_inflight", ctypes.c_int),
                ("num_streams", ctypes.c_int), ("max_batch", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("per_gemm_baseline", ctypes.c_int)]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libflashgemm.so. Raises RuntimeError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m flashgemm_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, l, i = ctypes.c_void_p, ctypes.c_long, ctypes.c_int
        lp, vpp = ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_void_p)
        L.fg_chain.argtypes = [i, vp, vp, l, i, vpp, lp, lp, ctypes.POINTER(ChainOpts), vp]
        L.fg_chain_host.argtypes = [i, vp, vp, l, i, vpp, lp, lp, ctypes.POINTER(ChainOpts)]
        L.fg_chain_plan.argtypes = [i, l, i, lp, lp, ctypes.POINTER(ChainOpts), ctypes.POINTER(i),
                                    ctypes.POINTER(i)]
        L.fg_conv_dims.argtypes = [ctypes.POINTER(ConvGeom), lp, lp]
        L.fg_im2col.argtypes = [i, vp, vp, ctypes.POINTER(ConvGeom), vp]
        L.fg_chain_conv.argtypes = [i, vp, vp, ctypes.POINTER(ConvGeom), i, vpp, lp, lp,
                                    ctypes.POINTER(Epilogue), ctypes.POINTER(ChainOpts), vp]
        L.fg_chain_epi.argtypes = [i, vp, vp, l, i, vpp, lp, lp, ctypes.POINTER(Epilogue),
                                   ctypes.POINTER(ChainOpts), vp]
        L.fg_chain_conv_host.argtypes = [i, vp, vp, ctypes.POINTER(ConvGeom), i, vpp, lp, lp,
                                         ctypes.POINTER(ChainOpts)]
        L.fg_sgemm.argtypes = [vp, vp, vp, l, l, l, i, i, vp]
        L.fg_sgemm_t.argtypes = [i, i, vp, vp, vp, l, l, l, i, i, vp]
        L.fg_sgemm_noalloc.argtypes = [vp, vp, vp, l, l, l, i, i, vp]
        L.fg_sgemm_host.argtypes = [vp, vp, vp, l, l, l, i]
        L.fg_f64_chain.argtypes = [vp, vp, l, i, vpp, lp, lp, vp]
        L.fg_to_f64.argtypes = [i, vp, vp, l, vp]
        L.fg_error_norms.argtypes = [i, vp, vp, l, ctypes.POINTER(ctypes.c_double), vp]
        L.fg_measure_stream.argtypes = [l, i, i, ctypes.POINTER(ctypes.c_float), vp]
        L.fg_inject_alloc_failures.argtypes = [i]
        L.fg_init.argtypes = [i]
        L.fg_synchronize.argtypes = [vp]
        L.fg_last_error.restype = ctypes.c_char_p
        L.fg_status_string.restype = ctypes.c_char_p
        L.flashgemm_set_thread_num.argtypes = [i]
        L.LibShalom_set_thread_nums.argtypes = [i]
        L.LibShalom_sgemm_mp.argtypes = [i, i, vp, vp, vp, l, l, l]
        L.flashgemm_sgemm.argtypes = [i, i, vp, vp, vp, l, l, l]
        L.fg_packed_weight_bytes.argtypes = [i, i, l, l]
        L.fg_packed_weight_bytes.restype = ctypes.c_size_t
        L.fg_pack_weight.argtypes = [i, i, vp, l, l, vp, vp]
        L.fg_grouped_chain.argtypes = [i, ctypes.POINTER(ChainReq), vp]
        L.fg_grouped_fits.argtypes = [ctypes.POINTER(ChainReq)]
        L.fg_grouped_tile.argtypes = [ctypes.POINTER(ChainReq)]
        L.fg_grouped_levels.argtypes = [i, ctypes.POINTER(ChainReq), vp]
        L.fg_chain_plan_segments.argtypes = [i, l, i, lp, lp, ctypes.POINTER(ChainOpts), ctypes.POINTER(i),
                                             ctypes.POINTER(i), ctypes.POINTER(i)]
        L.fg_grouped_graph_create.argtypes = [i, i, ctypes.POINTER(vp)]
        L.fg_grouped_graph_launch.argtypes = [vp, i, ctypes.POINTER(ChainReq), vp]
        L.fg_grouped_graph_destroy.argtypes = [vp]
        L.fg_autotune_chain.argtypes = [i, i, l, i, lp, lp, vp, ctypes.POINTER(i), ctypes.POINTER(i),
                                        ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double)]
        L.fg_autotune_save.argtypes = [ctypes.c_char_p]
        L.fg_autotune_load.argtypes = [ctypes.c_char_p]
        L.fg_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
        L.fg_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.fg_server_add_weights.argtypes = [vp, i, i, i, vpp, lp, lp, ctypes.POINTER(i)]
        L.fg_server_submit.argtypes = [vp, i, l, vp, vp, ctypes.POINTER(l)]
        L.fg_server_drain.argtypes = [vp, ctypes.c_double]
        L.fg_server_run_trace.argtypes = [vp, i, ctypes.POINTER(i), lp, vpp, vpp, ctypes.POINTER(ctypes.c_double),
                                          ctypes.c_double]
        L.fg_server_latencies.argtypes = [vp, l, l, ctypes.POINTER(ctypes.c_double)]
        L.fg_server_cancel.argtypes = [vp, l]
        L.fg_server_num_requests.argtypes = [vp]
        L.fg_server_num_requests.restype = l
        L.fg_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check(status: int) -> None:
    if status != 0:
        raise FlashgemmError(status, lib().fg_last_error().decode(errors="replace"))


def last_error() -> str:
    return lib().fg_last_error().decode(errors="replace")


def opts(layout=LAYOUT_MATH, ref_threads=1, f32_mode=F32_MFMA, fuse=FUSE_AUTO, tile_n=0, splits=0,
         a_packed=0) -> ChainOpts:
    return ChainOpts(layout, ref_threads, f32_mode, fuse, tile_n, splits, a_packed)


def _addr(x) -> int:
    """Device or host address of a torch tensor, numpy array or int."""
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")


def _stream(stream) -> ctypes.c_void_p:
    if stream is None:
        return ctypes.c_void_p(0)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)  # torch.cuda.Stream


def _arrays(As, Ms=None, Ks=None):
    n = len(As)
    ptrs = (ctypes.c_void_p * n)(*[_addr(a) for a in As])
    Ms = Ms if Ms is not None else [a.shape[0] for a in As]
    Ks = Ks if Ks is not None else [a.shape[1] for a in As]
    return n, ptrs, (ctypes.c_long * n)(*Ms), (ctypes.c_long * n)(*Ks)


# ------------------------------------------------------------- device API --

def init(device: int = 0) -> None:
    _check(lib().fg_init(device))


def chain(dtype: int, C, B, As, N: int | None = None, Ms=None, Ks=None, o: ChainOpts | None = None,
          stream=None) -> None:
    """C = A_n ... A_1 B on device buffers (reference orientation, row-major)."""
    n, ptrs, M, K = _arrays(As, Ms, Ks)
    N = N if N is not None else B.shape[1]
    _check(lib().fg_chain(dtype, _addr(C), _addr(B), N, n, ptrs, M, K,
                          ctypes.byref(o) if o is not None else None, _stream(stream)))


def chain_host(dtype: int, C, B, As, o: ChainOpts | None = None) -> None:
    """Blocking chain on host numpy arrays."""
    n, ptrs, M, K = _arrays(As)
    _check(lib().fg_chain_host(dtype, _addr(C), _addr(B), B.shape[1], n, ptrs, M, K,
                               ctypes.byref(o) if o is not None else None))


def chain_plan(dtype: int, N: int, Ms, Ks, o: ChainOpts | None = None):
    n = len(Ms)
    fused, tn = ctypes.c_int(), ctypes.c_int()
    _check(lib().fg_chain_plan(dtype, N, n, (ctypes.c_long * n)(*Ms), (ctypes.c_long * n)(*Ks),
                               ctypes.byref(o) if o is not None else None, ctypes.byref(fused), ctypes.byref(tn)))
    return bool(fused.value), tn.value


def chain_plan_segments(dtype: int, N: int, Ms, Ks, o: ChainOpts | None = None):
    """The planner's segments: a list of (last GEMM index, tile width or 0 = unfused GEMM)."""
    n = len(Ms)
    ns = ctypes.c_int()
    ends, tiles = (ctypes.c_int * n)(), (ctypes.c_int * n)()
    _check(lib().fg_chain_plan_segments(dtype, N, n, (ctypes.c_long * n)(*Ms), (ctypes.c_long * n)(*Ks),
                                        ctypes.byref(o) if o is not None else None, ctypes.byref(ns), ends, tiles))
    return [(ends[s], tiles[s]) for s in range(ns.value)]


def conv_dims(g: ConvGeom):
    """(K1, N) of the patch matrix of geometry g."""
    K1, N = ctypes.c_long(), ctypes.c_long()
    _check(lib().fg_conv_dims(ctypes.byref(g), ctypes.byref(K1), ctypes.byref(N)))
    return K1.value, N.value


def im2col(dtype: int, B, X, g: ConvGeom, stream=None) -> None:
    """B (K1 x N) = patch matrix of the NCHW image X on device buffers."""
    _check(lib().fg_im2col(dtype, _addr(B), _addr(X), ctypes.byref(g), _stream(stream)))


def chain_conv(dtype: int, C, X, g: ConvGeom, As, Ms=None, Ks=None, o: ChainOpts | None = None,
               stream=None, epi: Epilogue | None = None) -> None:
    """C = A_n ... A_1 im2col(X) on device buffers; the fused chain gathers its B tile from X."""
    n, ptrs, M, K = _arrays(As, Ms, Ks)
    _check(lib().fg_chain_conv(dtype, _addr(C), _addr(X), ctypes.byref(g), n, ptrs, M, K,
                               ctypes.byref(epi) if epi is not None else None,
                               ctypes.byref(o) if o is not None else None, _stream(stream)))


def chain_conv_host(dtype: int, C, X, g: ConvGeom, As, o: ChainOpts | None = None) -> None:
    """Blocking convolution chain on host numpy arrays (X is NCHW)."""
    n, ptrs, M, K = _arrays(As)
    _check(lib().fg_chain_conv_host(dtype, _addr(C), _addr(X), ctypes.byref(g), n, ptrs, M, K,
                                    ctypes.byref(o) if o is not None else None))


def chain_epi(dtype: int, C, B, As, epi: Epilogue | None, N: int | None = None, Ms=None, Ks=None,
              o: ChainOpts | None = None, stream=None) -> None:
    """fg_chain with epilogue extensions (per-GEMM bias, scale, INT8 zero point, ReLU)."""
    n, ptrs, M, K = _arrays(As, Ms, Ks)
    N = N if N is not None else B.shape[1]
    _check(lib().fg_chain_epi(dtype, _addr(C), _addr(B), N, n, ptrs, M, K,
                              ctypes.byref(epi) if epi is not None else None,
                              ctypes.byref(o) if o is not None else None, _stream(stream)))


def sgemm(C, A, B, M: int, N: int, K: int, f32_mode: int = F32_MFMA, splits: int = 0, stream=None) -> None:
    _check(lib().fg_sgemm(_addr(C), _addr(A), _addr(B), M, N, K, f32_mode, splits, _stream(stream)))


def sgemm_t(transa: int, transb: int, C, A, B, M: int, N: int, K: int, f32_mode: int = F32_MFMA, splits: int = 0,
            stream=None) -> None:
    """C (M x N) = op(A) . op(B); transa: A stored K x M, transb: B stored N x K."""
    _check(lib().fg_sgemm_t(transa, transb, _addr(C), _addr(A), _addr(B), M, N, K, f32_mode, splits,
                            _stream(stream)))


NOALLOC_SCRATCH_BYTES = 32 << 20  # FG_NOALLOC_SCRATCH_BYTES


def sgemm_noalloc(C, A, B, M: int, N: int, K: int, f32_mode: int = F32_MFMA, splits: int = 0, stream=None) -> None:
    """fg_sgemm_noalloc: sgemm with its split-K slabs in the library's static device array
    (no allocation; calls overlapping on one GPU race)."""
    _check(lib().fg_sgemm_noalloc(_addr(C), _addr(A), _addr(B), M, N, K, f32_mode, splits, _stream(stream)))


def f64_chain(C, B, As, N: int | None = None, Ms=None, Ks=None, stream=None) -> None:
    n, ptrs, M, K = _arrays(As, Ms, Ks)
    N = N if N is not None else B.shape[1]
    _check(lib().fg_f64_chain(_addr(C), _addr(B), N, n, ptrs, M, K, _stream(stream)))


def to_f64(x_kind: int, x, out, n: int, stream=None) -> None:
    _check(lib().fg_to_f64(x_kind, _addr(x), _addr(out), n, _stream(stream)))


def error_norms(x_kind: int, x, ref, n: int, stream=None):
    """(sum (x-ref)^2, sum ref^2) in fp64; x_kind 0 fp32, 1 bf16, 2 int32, 3 fp64."""
    out = (ctypes.c_double * 2)()
    _check(lib().fg_error_norms(x_kind, _addr(x), _addr(ref), n, out, _stream(stream)))
    return out[0], out[1]


def inject_alloc_failures(count: int) -> None:
    """Make the library's next `count` device allocations fail as out of memory (tests of
    the error paths: the call needing one raises FlashgemmError with ERR_NOMEM); 0 clears."""
    _check(lib().fg_inject_alloc_failures(count))


def measure_stream(slice_kb: int, rounds: int, workgroups: int, stream=None) -> float:
    """Milliseconds for `workgroups` workgroups to stream 8 * 16 * (rounds + 1) KB each
    (the fused chains' weight-stream pattern, measurement only)."""
    ms = ctypes.c_float()
    _check(lib().fg_measure_stream(slice_kb, rounds, workgroups, ctypes.byref(ms), _stream(stream)))
    return ms.value


def packed_weight_bytes(dtype: int, M: int, K: int, f32_mode: int = F32_MFMA) -> int:
    """Bytes of one weight matrix in the fused kernels' packed layout."""
    return lib().fg_packed_weight_bytes(dtype, f32_mode, M, K)


def pack_weight(dtype: int, A, M: int, K: int, packed, f32_mode: int = F32_MFMA, stream=None) -> None:
    """Pack A (M x K, device) into `packed` for chains run with a_packed=1."""
    _check(lib().fg_pack_weight(dtype, f32_mode, _addr(A), M, K, _addr(packed), _stream(stream)))


def synchronize(stream=None) -> None:
    _check(lib().fg_synchronize(_stream(stream)))


# ------------------------------------------------ reference entry points --

def flashgemm_set_thread_num(num: int) -> None:
    lib().flashgemm_set_thread_num(num)


def _legacy(fn, C, B, N, num_gemm, triples):
    if len(triples) != 3 * num_gemm:
        raise ValueError("expected num_gemm triples (A_i, M_i, K_i)")
    args = [ctypes.c_void_p(_addr(C)), ctypes.c_void_p(_addr(B)), ctypes.c_long(N), ctypes.c_int(num_gemm)]
    for j in range(num_gemm):
        A, M, K = triples[3 * j:3 * j + 3]
        args += [ctypes.c_void_p(_addr(A)), ctypes.c_long(M), ctypes.c_long(K)]
    fn(*args)


def flashgemm_multi_f32f32f32(C, B, N: int, num_gemm: int, *triples) -> None:
    _legacy(lib().flashgemm_multi_f32f32f32, C, B, N, num_gemm, triples)


def flashgemm_multi_bf16bf16f32(C, B, N: int, num_gemm: int, *triples) -> None:
    _legacy(lib().flashgemm_multi_bf16bf16f32, C, B, N, num_gemm, triples)


def flashgemm_multi_int8uint8int32(C, B, N: int, num_gemm: int, *triples) -> None:
    _legacy(lib().flashgemm_multi_int8uint8int32, C, B, N, num_gemm, triples)


def LibShalom_set_thread_nums(num: int) -> None:
    lib().LibShalom_set_thread_nums(num)


def LibShalom_sgemm_mp(transa: int, transb: int, C, A, B, M: int, N: int, K: int) -> None:
    lib().LibShalom_sgemm_mp(transa, transb, _addr(C), _addr(A), _addr(B), M, N, K)


# ------------------------------------------------------ grouped launches --

def chain_req(dtype: int, C, B, As, N: int | None = None, f32_mode: int = F32_MFMA, Ms=None, Ks=None,
              a_packed: int = 0) -> ChainReq:
    r = ChainReq()
    r.dtype, r.f32_mode, r.a_packed = dtype, f32_mode, a_packed
    r.C, r.B = _addr(C), _addr(B)
    r.N = N if N is not None else B.shape[1]
    r.num_gemm = len(As)
    Ms = Ms if Ms is not None else [a.shape[0] for a in As]
    Ks = Ks if Ks is not None else [a.shape[1] for a in As]
    for i, a in enumerate(As):
        r.A[i], r.M[i], r.K[i] = _addr(a), Ms[i], Ks[i]
    return r


def grouped_chain(reqs, stream=None) -> None:
    arr = (ChainReq * len(reqs))(*reqs)
    _check(lib().fg_grouped_chain(len(reqs), arr, _stream(stream)))


def grouped_levels(reqs, stream=None) -> None:
    """Every request unfused, GEMM i of all of them in one launch (fg_grouped_levels)."""
    arr = (ChainReq * len(reqs))(*reqs)
    _check(lib().fg_grouped_levels(len(reqs), arr, _stream(stream)))


def grouped_fits(req: ChainReq) -> bool:
    rc = lib().fg_grouped_fits(ctypes.byref(req))
    if rc < 0:
        _check(rc)
    return rc == 1


def grouped_tile(req: ChainReq) -> int:
    """Tile width of the request in a grouped launch (32 or 64), 0 if it runs outside it."""
    rc = lib().fg_grouped_tile(ctypes.byref(req))
    if rc < 0:
        _check(rc)
    return rc


class GroupedGraph:
    """Captured graph of [table upload -> grouped kernel], replayed per batch."""

    def __init__(self, max_reqs: int, max_tiles: int):
        self.h = ctypes.c_void_p()
        _check(lib().fg_grouped_graph_create(max_reqs, max_tiles, ctypes.byref(self.h)))

    def launch(self, reqs, stream=None) -> None:
        arr = (ChainReq * len(reqs))(*reqs)
        _check(lib().fg_grouped_graph_launch(self.h, len(reqs), arr, _stream(stream)))

    def close(self) -> None:
        if self.h:
            lib().fg_grouped_graph_destroy(self.h)
            self.h = ctypes.c_void_p()


# --------------------------------------------------------------- autotune --

def autotune_chain(dtype: int, N: int, Ms, Ks, f32_mode: int = F32_MFMA, stream=None):
    n = len(Ms)
    f, t, s, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().fg_autotune_chain(dtype, f32_mode, N, n, (ctypes.c_long * n)(*Ms), (ctypes.c_long * n)(*Ks),
                                   _stream(stream), ctypes.byref(f), ctypes.byref(t), ctypes.byref(s),
                                   ctypes.byref(ms)))
    return {"fused": bool(f.value), "tile_n": t.value, "splits": s.value, "ms": ms.value}


def autotune_save(path: str) -> None:
    _check(lib().fg_autotune_save(path.encode()))


def autotune_load(path: str) -> None:
    _check(lib().fg_autotune_load(path.encode()))


def autotune_clear() -> None:
    lib().fg_autotune_clear()


def autotune_size() -> int:
    return lib().fg_autotune_size()


# ---------------------------------------------------------------- serving --

class Server:
    """Continuous-batching request loop over resident weight sets (device buffers)."""

    def __init__(self, budget_bytes: int | None = None, max_inflight: int | None = None, num_streams: int | None = None,
                 max_batch: int | None = None, use_graph: bool | None = None, per_gemm_baseline: bool = False):
        cfg = ServerConfig()
        lib().fg_server_config_default(ctypes.byref(cfg))
        if budget_bytes is not None:
            cfg.hbm_budget_bytes = budget_bytes
        if max_inflight is not None:
            cfg.max_inflight = max_inflight
        if num_streams is not None:
            cfg.num_streams = num_streams
        if max_batch is not None:
            cfg.max_batch = max_batch
        if use_graph is not None:
            cfg.use_graph = int(use_graph)
        cfg.per_gemm_baseline = int(per_gemm_baseline)
        self.h = ctypes.c_void_p()
        _check(lib().fg_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def add_weights(self, dtype: int, As, f32_mode: int = F32_MFMA) -> int:
        n, ptrs, M, K = _arrays(As)
        sid = ctypes.c_int()
        _check(lib().fg_server_add_weights(self.h, dtype, f32_mode, n, ptrs, M, K, ctypes.byref(sid)))
        return sid.value

    def submit(self, set_id: int, B, C, N: int | None = None) -> int:
        rid = ctypes.c_long()
        _check(lib().fg_server_submit(self.h, set_id, N if N is not None else B.shape[1], _addr(B), _addr(C),
                                      ctypes.byref(rid)))
        return rid.value

    def drain(self, timeout_s: float = 600.0) -> None:
        _check(lib().fg_server_drain(self.h, timeout_s))

    def cancel(self, req_id: int) -> bool:
        """True when the request was dropped before admission (its C is never written);
        False when it was already admitted, running or done (it completes normally)."""
        rc = lib().fg_server_cancel(self.h, req_id)
        if rc == ERR_INVALID and req_id >= 0 and req_id < lib().fg_server_num_requests(self.h):
            return False
        _check(rc)
        return True

    def run_trace(self, set_ids, Ns, Bs, Cs, arrivals, timeout_s: float = 600.0) -> None:
        n = len(set_ids)
        _check(lib().fg_server_run_trace(self.h, n, (ctypes.c_int * n)(*set_ids), (ctypes.c_long * n)(*Ns),
                                         (ctypes.c_void_p * n)(*[_addr(b) for b in Bs]),
                                         (ctypes.c_void_p * n)(*[_addr(c) for c in Cs]),
                                         (ctypes.c_double * n)(*arrivals), timeout_s))

    def num_requests(self) -> int:
        """Requests submitted to this server so far (warm-up and traces included)."""
        return lib().fg_server_num_requests(self.h)

    def latencies(self, first: int = 0, n: int | None = None):
        total = lib().fg_server_num_requests(self.h)
        n = total - first if n is None else n
        out = (ctypes.c_double * n)()
        _check(lib().fg_server_latencies(self.h, first, n, out))
        return list(out)

    def close(self) -> None:
        if self.h:
            lib().fg_server_destroy(self.h)
            self.h = ctypes.c_void_p()
