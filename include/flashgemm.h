/*
 * flashgemm.h - C ABI of libflashgemm (MI355X / gfx950, HIP).
 *
 * Part 1 keeps the reference's entry points with their exact names, signatures,
 * argument meaning and orientation, so code written against the reference links
 * against libflashgemm unchanged. Part 2 is the status-returning API over device
 * pointers and HIP streams that part 1 is built on.
 *
 * Orientation (all parts): a chain computes C = A_n ... A_2 A_1 B, B is K_1 x N,
 * A_i is M_i x K_i with K_{i+1} = M_i, C is M_n x N; everything row-major with
 * lda = K_i and ldb = ldc = N. N is the wide dimension. INT8 chains read A bytes as
 * unsigned and B bytes as signed, accumulate in int32 and saturate each
 * intermediate to int8; BF16 chains accumulate in fp32 and round each
 * intermediate to bf16 (round to nearest even); FP32 intermediates stay fp32.
 *
 * No torch or HIP types appear in the signatures: device pointers are plain
 * pointers and a stream is a `void *` holding a hipStream_t (NULL = default).
 */
#ifndef FLASHGEMM_H
#define FLASHGEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================== 1. reference ABI === */

/* Replaces multigemm/src/flashgemm.h:17 / flashgemm.c:6. Here it is the thread
 * count whose column-panel decomposition the REFERENCE intermediate layout follows
 * (see fg_chain_opts.layout); it does not start CPU threads. Default 1. */
extern int flashgemm_thread_num;

/* Replaces multigemm/src/flashgemm.h:19 (flashgemm.c:8-11). */
void flashgemm_set_thread_num(int num);

/* Replace multigemm/src/flashgemm.h:21, :23, :25 (flashgemm.c:24-424).
 * Host pointers; blocking. The variadic tail is num_gemm triples
 * (T *A_i, long M_i, long K_i), read as in flashgemm.c:54-62.
 * Like the reference, a K_i that is not a multiple of 2 prints
 * "K must be a multiple of 2" / "Dimension error for gemm i" and returns with C
 * untouched. INT8 and BF16 intermediates follow the reference's layout for
 * flashgemm_thread_num threads unless FLASHGEMM_LAYOUT=math is set; FP32 uses the
 * FP32 MFMA unless FLASHGEMM_F32_MODE=bf16x3 is set. */
void flashgemm_multi_f32f32f32(float *C, float *B, long N, int num_gemm, ...);
void flashgemm_multi_bf16bf16f32(float *C, uint16_t *B, long N, int num_gemm, ...);
void flashgemm_multi_int8uint8int32(int *C, uint8_t *B, long N, int num_gemm, ...);

/* Replaces singlegemm/src/set_TmTn.h:61-74. Accepted for compatibility; the GPU
 * needs no thread grid. */
void LibShalom_set_thread_nums(int num);

/* Replaces the singlegemm dispatcher singlegemm/FP32_ID23.c:35-62 (defined in the
 * reference's driver): C = A.B, FP32, A M x K, B K x N, C M x N, host pointers,
 * blocking. transb == 1 (NT) computes nothing, as in the reference. */
void LibShalom_sgemm_mp(int transa, int transb, float *C, float *A, float *B, long M, long N, long K);
/* Same call under this library's own name. */
void flashgemm_sgemm(int transa, int transb, float *C, float *A, float *B, long M, long N, long K);

/* ================================================= 2. status-code API === */

enum fg_status {
  FG_OK = 0,
  FG_ERR_INVALID = -1,    /* bad argument or shape */
  FG_ERR_DOMAIN = -2,     /* reference layout undefined for this N / thread count */
  FG_ERR_NOMEM = -3,      /* device allocation failed */
  FG_ERR_HIP = -4,        /* a HIP call failed (fg_last_error() has the text) */
  FG_ERR_BUDGET = -5,     /* request exceeds the serving HBM budget */
  FG_ERR_NODEVICE = -6,   /* no usable gfx950 device */
  FG_ERR_UNSUPPORTED = -7 /* option combination not implemented */
};

enum fg_dtype { FG_DTYPE_F32 = 0, FG_DTYPE_BF16 = 1, FG_DTYPE_INT8 = 2 };
/* FP32 arithmetic: FG_F32_MFMA (fp32 MFMA), FG_F32_BF16X3 (two-piece split,
 * x = hi + lo in bf16, three bf16 products per pair, fp32 accumulation), or
 * FG_F32_BF16X6 (three-piece split x = hi + mid + lo, the six products of order <= 2:
 * close to FP32 MFMA error; unfused, NN operands; runs as one BF16 GEMM over 6K). */
enum fg_f32_mode { FG_F32_MFMA = 0, FG_F32_BF16X3 = 1, FG_F32_BF16X6 = 2 };
enum fg_layout { FG_LAYOUT_MATH = 0, FG_LAYOUT_REFERENCE = 1 };
enum fg_fuse { FG_FUSE_AUTO = -1, FG_FUSE_OFF = 0, FG_FUSE_ON = 1 };

typedef struct fg_chain_opts {
  int layout;      /* FG_LAYOUT_MATH (default) or FG_LAYOUT_REFERENCE */
  int ref_threads; /* thread count the REFERENCE layout follows (>= 1)   */
  int f32_mode;    /* FG_F32_MFMA (default), FG_F32_BF16X3, FG_F32_BF16X6 */
  int fuse;        /* FG_FUSE_AUTO (default), FG_FUSE_OFF, FG_FUSE_ON    */
  int tile_n;      /* fused tile width: 0 = planner, or 32, 64, 96, 128, 256 (96 not for INT8/BF16) */
  int splits;      /* split-K of unfused GEMMs: 0 = planner              */
  int a_packed;    /* 1: A[i] come from fg_pack_weight (fused path only) */
} fg_chain_opts;

void fg_chain_opts_default(fg_chain_opts *opts);
const char *fg_status_string(int status);
/* Text of the last error on the calling thread ("" if none). */
const char *fg_last_error(void);
/* Select the device (default 0) and check it is gfx950. Called implicitly. */
int fg_init(int device);
int fg_synchronize(void *stream);

/* Chain on device pointers, enqueued on `stream`, asynchronous. A[i] are device
 * pointers. opts may be NULL (defaults). Library scratch (spill intermediates,
 * split-K partials, column sums) comes from a per-stream cache owned by the
 * library; calls on one stream are ordered, calls on different streams use
 * different scratch. */
int fg_chain(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
             const long *K, const fg_chain_opts *opts, void *stream);

/* ---- epilogue extensions (no reference counterpart: the reference only
 * saturates its INT8 intermediate to int8, multigemm/src/kernel_int8.h:329-385, and
 * rounds its BF16 intermediate to bf16, kernel_bf16.h:356-390). Per GEMM i, row m
 * (output channel m of GEMM i):
 *   INT8 intermediate: a = acc + bias[m] (int32, wrapping); t = rne(fp32(a) * scale[m])
 *     (half-to-even) when scale is set, else t = a; q = sat8(t + zero_point);
 *     ReLU: q = max(q, zero_point). The next GEMM reads q.
 *   INT8 output (int32 C): c = acc + bias[m]; ReLU: max(c, 0). No scale, no zero point.
 *   FP32 / BF16: y = acc * scale[m], then y + bias[m] (fp32, two roundings); ReLU:
 *     max(y, 0) (NaN -> 0); a BF16 intermediate is then rounded to bf16.
 * Every field at zero / NULL gives the reference's epilogue. bias, scale: device
 * pointers to M_i entries (bias int32 for INT8, fp32 otherwise). ---- */
typedef struct fg_epilogue {
  const void *bias[8];
  const float *scale[8];
  int zero_point[8]; /* INT8 intermediates only, -128..127 */
  int relu[8];
} fg_epilogue;

/* fg_chain with epilogue extensions (epi may be NULL). FG_ERR_UNSUPPORTED for a
 * zero point outside INT8 intermediates or a scale on the INT8 int32 output. */
int fg_chain_epi(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                 const long *K, const fg_epilogue *epi, const fg_chain_opts *opts, void *stream);

/* Same with host pointers; blocking (copies in, runs, copies out). */
int fg_chain_host(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                  const long *K, const fg_chain_opts *opts);

/* FP32 single GEMM C = A.B on device pointers (f32_mode as in fg_chain_opts);
 * splits = 0 lets the planner choose split-K. */
int fg_sgemm(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode, int splits,
             void *stream);
int fg_sgemm_host(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode);
/* C = op(A) . op(B) with op(A) M x K and op(B) K x N: transa = 1 reads A stored
 * K x M (row-major, lda = M), transb = 1 reads B stored N x K (ldb = K); C is M x N.
 * The reference's dispatcher takes the same flags but computes NN only (NT is an
 * empty stub, singlegemm/FP32_ID23.c:39-43); LibShalom_sgemm_mp keeps that
 * behaviour, this call computes every combination. */
int fg_sgemm_t(int transa, int transb, float *C, const float *A, const float *B, long M, long N, long K,
               int f32_mode, int splits, void *stream);

/* fg_sgemm that allocates nothing. No counterpart in the reference (whose drivers
 * allocate per call); for callers that cannot call hipMalloc: code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation. Same plan, kernels
 * and bits as fg_sgemm; its split-K slabs live in a static device array of
 * FG_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of library
 * scratch. A plan whose slabs (splits x M x N x 4 bytes) exceed that returns
 * FG_ERR_INVALID. Every call on one GPU shares the array: calls that overlap on one
 * GPU (other streams, other threads) race. */
#define FG_NOALLOC_SCRATCH_BYTES (32u << 20)
int fg_sgemm_noalloc(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode, int splits,
                     void *stream);

/* Weights in the fused kernels' packed layout: MFMA fragment order, zero padded,
 * INT8 bytes pre-biased to signed, FP32 for FG_F32_BF16X3 pre-split into bf16 hi/lo.
 * The reference packs A inside every call (multigemm/src/PACK_x86.h:616-647,
 * flashgemm.c:85-105); fg_chain does the same unless opts->a_packed says A[i] are
 * already packed (for weights that stay resident). packed must hold
 * fg_packed_weight_bytes bytes; the layout depends on dtype and f32_mode. */
size_t fg_packed_weight_bytes(int dtype, int f32_mode, long M, long K);
int fg_pack_weight(int dtype, int f32_mode, const void *A, long M, long K, void *packed, void *stream);

/* Plan inspection: writes the planner's decision for a chain without running it.
 * fused = 1 if the whole chain runs in one fused launch, tile_n its tile width. */
int fg_chain_plan(int dtype, long N, int num_gemm, const long *M, const long *K, const fg_chain_opts *opts,
                  int *fused, int *tile_n);
/* The full plan (no reference counterpart: the reference runs every chain panel by
 * panel on chip, multigemm/src/flashgemm.c:123-148). A chain is cut into segments of
 * consecutive GEMMs: segment s ends with GEMM seg_end[s] (0-based, inclusive) and runs
 * fused at tile width seg_tile[s] (its intermediates on chip; a segment that ends
 * before the chain does writes the next GEMM's operand to HBM), or, when seg_tile[s]
 * is 0, as one unfused GEMM. seg_end and seg_tile hold num_gemm entries. */
int fg_chain_plan_segments(int dtype, long N, int num_gemm, const long *M, const long *K, const fg_chain_opts *opts,
                           int *num_segments, int *seg_end, int *seg_tile);

/* ---- convolution front end: B produced from an image (im2col). The reference's
 * chains are convolution layers written as GEMMs: its test shapes take K_1 =
 * channels * 3 * 3 or * 5 * 5 and N = height * width (multigemm/test/test_f32_3.c:14-48),
 * and its callers lay the patches out as B before the call. Here the fused chain
 * gathers the patches straight from the image while it stages its B tile, so B
 * never exists in HBM. ----
 *
 * X is batch x channels x height x width, row-major (NCHW), of the chain's dtype
 * (fp32, raw bf16 or int8). B[k][n] = X[b][c][iy][ix] with
 *   k = (c * kernel_h + ky) * kernel_w + kx,   n = (b * out_h + oy) * out_w + ox,
 *   iy = oy * stride_h - pad_h + ky * dilation_h,  ix = ox * stride_w - pad_w + kx * dilation_w,
 * and 0 where (iy, ix) falls outside the image (the row order of torch's unfold).
 * out_h = (height + 2 pad_h - dilation_h (kernel_h - 1) - 1) / stride_h + 1, out_w alike;
 * K_1 = channels * kernel_h * kernel_w, N = batch * out_h * out_w. */
typedef struct fg_conv_geom {
  long batch, channels, height, width;
  long kernel_h, kernel_w;
  long stride_h, stride_w;     /* >= 1 */
  long pad_h, pad_w;           /* >= 0 */
  long dilation_h, dilation_w; /* >= 1 */
} fg_conv_geom;

/* Output size of the patch matrix: *K1 = channels * kernel_h * kernel_w, *N =
 * batch * out_h * out_w (either pointer may be NULL). FG_ERR_INVALID for a bad
 * geometry (non-positive sizes, empty output, an image of 2^31 elements or more). */
int fg_conv_dims(const fg_conv_geom *g, long *K1, long *N);
/* B = im2col(X) on device pointers: B is K1 x N row-major of the dtype. */
int fg_im2col(int dtype, void *B, const void *X, const fg_conv_geom *g, void *stream);
/* C = A_n ... A_1 im2col(X) on device pointers; K[0] must equal K1 of g and C is
 * M_n x N, epi as in fg_chain_epi (may be NULL). The fused chain gathers its B tile
 * from X; when the chain runs unfused (options, LDS, the reference layout) B is
 * formed in library scratch first. */
int fg_chain_conv(int dtype, void *C, const void *X, const fg_conv_geom *g, int num_gemm, const void *const *A,
                  const long *M, const long *K, const fg_epilogue *epi, const fg_chain_opts *opts, void *stream);
/* Same with host pointers (no epilogue extensions); blocking, like fg_chain_host. */
int fg_chain_conv_host(int dtype, void *C, const void *X, const fg_conv_geom *g, int num_gemm, const void *const *A,
                       const long *M, const long *K, const fg_chain_opts *opts);

/* ---- grouped launches (no reference counterpart: the reference runs one chain
 * per call). Many chains of uneven shapes and mixed dtypes, mathematical layout,
 * all pointers device pointers. ---- */
typedef struct fg_chain_req {
  int dtype;             /* FG_DTYPE_*                        */
  int f32_mode;          /* FG_F32_* for FP32 requests        */
  void *C;               /* M_n x N, fp32 (int32 for INT8)    */
  const void *B;         /* K_1 x N                           */
  long N;
  int num_gemm;          /* 1..8                              */
  const void *A[8];      /* A_i, M_i x K_i                    */
  long M[8], K[8];
  int a_packed;          /* 1: A[i] come from fg_pack_weight  */
} fg_chain_req;

/* All requests in one persistent launch: each request's columns are cut into tiles
 * (32 columns wide, or 64 where the autotune table's plan for the request's shape
 * bucket says so) that run the whole chain on chip. Requests whose intermediates do
 * not fit on chip at that width, or whose tuned plan is not one fused launch at 32 or
 * 64 columns, run with their own plan (fg_chain) on the same stream after it.
 * Asynchronous. */
int fg_grouped_chain(int count, const fg_chain_req *reqs, void *stream);
/* 1 if the request runs on chip in a grouped launch, 0 if not, < 0 on error. */
int fg_grouped_fits(const fg_chain_req *req);
/* The grouped launch's tile width for the request (32 or 64), 0 if it runs outside
 * the grouped launch, < 0 on error. */
int fg_grouped_tile(const fg_chain_req *req);
/* Every request unfused, level by level: GEMM i of all requests (i = 1 .. the longest
 * chain, any mix of dtypes and modes) runs in one launch, 128 x 128 output tiles of
 * all of them in one grid; a GEMM whose tiles alone would not fill the chip is split
 * over k as fg_chain's planner splits a lone GEMM (the split depends on its shape
 * only), its slabs added in slab order by one reduction launch per level.
 * Intermediates go to the stream's scratch in the next operand's type (as fg_chain's
 * unfused path writes them). A request's result has the same bits in any batch (and
 * alone); INT8 results equal every other path's. For batches of requests too large for
 * the on-chip grouped launch: a few launches per level instead of one or more per GEMM
 * per request. Weights as given (a_packed must be 0). Asynchronous. */
int fg_grouped_levels(int count, const fg_chain_req *reqs, void *stream);

/* A captured hipGraph of [descriptor-table upload -> grouped kernel], replayed per
 * batch. Four captured copies rotate so a batch can be prepared while earlier ones
 * run. Every request of a replayed batch must fit on chip (else
 * FG_ERR_UNSUPPORTED); at most max_reqs requests and max_tiles tiles. */
typedef struct fg_grouped_graph fg_grouped_graph;
int fg_grouped_graph_create(int max_reqs, int max_tiles, fg_grouped_graph **out);
int fg_grouped_graph_launch(fg_grouped_graph *graph, int count, const fg_chain_req *reqs, void *stream);
void fg_grouped_graph_destroy(fg_grouped_graph *graph);

/* ---- serving engine (no reference counterpart: the reference is a blocking
 * library call). Continuous batching of chain requests against resident weight
 * sets: requests are admitted in arrival order while a slot is free and their
 * bytes (B + C) plus the registered weights fit the HBM budget; a scheduler thread
 * packs every admitted request into the next grouped launch on one of num_streams
 * streams, so work of requests arriving while others run overlaps on the device.
 * per_gemm_baseline = 1 runs each request alone with one launch per GEMM instead.
 * All buffers are device pointers owned by the caller. The engine's streams do not
 * wait for other streams: B must hold its data, and no other work may still write
 * B or C, when a request is submitted; C is valid once fg_server_drain returns. ---- */
typedef struct fg_server_config {
  size_t hbm_budget_bytes; /* default 200 GiB                      */
  int max_inflight;        /* request slots, default 1024          */
  int num_streams;         /* 1..8, default 2                      */
  int max_batch;           /* requests per launch, default 512     */
  int use_graph;           /* replay captured graphs, default 1    */
  int per_gemm_baseline;   /* default 0                            */
} fg_server_config;
typedef struct fg_server fg_server;

void fg_server_config_default(fg_server_config *cfg);
int fg_server_create(const fg_server_config *cfg, fg_server **out);
/* Register a resident weight chain (device pointers A_i, M_i x K_i). */
int fg_server_add_weights(fg_server *srv, int dtype, int f32_mode, int num_gemm, const void *const *A,
                          const long *M, const long *K, int *set_id);
/* Enqueue C = A_n ... A_1 B for weight set `set_id`; B is K_1 x N, C is M_n x N. */
int fg_server_submit(fg_server *srv, int set_id, long N, const void *B, void *C, long *req_id);
/* Wait until every submitted request has completed (or the timeout passes). */
int fg_server_drain(fg_server *srv, double timeout_s);
/* Submit n requests at the given arrival offsets (seconds from the call), then drain. */
int fg_server_run_trace(fg_server *srv, int n, const int *set_ids, const long *Ns, const void *const *Bs,
                        void *const *Cs, const double *arrival_s, double timeout_s);
/* Cancel request req_id while it still waits for a slot or budget (not yet admitted to
 * a launch): it is dropped, its C is never written, and fg_server_drain no longer
 * waits for it. FG_OK when cancelled; FG_ERR_INVALID for an unknown id or a request
 * already admitted, running or done (it then completes normally). */
int fg_server_cancel(fg_server *srv, long req_id);
/* Latency (completion - arrival, seconds) of requests [first, first + n); -1 for a
 * request not completed yet, -2 for a cancelled one. */
int fg_server_latencies(fg_server *srv, long first, long n, double *out);
long fg_server_num_requests(fg_server *srv);
void fg_server_destroy(fg_server *srv);

/* ---- autotuner (replaces the reference's static blocking cost model,
 * singlegemm/src/set_block.h:8-200 and set_TmTn.h:10-58). Times the runnable plans
 * of one chain shape on seeded random data: the whole chain fused at each tile width
 * that fits LDS, every cut of the chain into fused segments (the fusion depth: which
 * intermediates stay on chip and which spill to HBM), and every GEMM unfused with and
 * without split-K. The fastest plan is kept in a table keyed by the shape bucket
 * (dtype / FP32 mode, chain length, ceil(log2) of N and of every M_i and K_i); the
 * planner of fg_chain and the grouped launch (tile width, or whether a request joins
 * it at all) consult it when opts leave the choice open and the plan can run the
 * exact shape (fg_chain_plan_segments shows the result). fused / tile_n report
 * whether the best plan is the whole chain in one fused launch. Outputs may be NULL. ---- */
int fg_autotune_chain(int dtype, int f32_mode, long N, int num_gemm, const long *M, const long *K, void *stream,
                      int *fused, int *tile_n, int *splits, double *ms);
int fg_autotune_save(const char *path); /* "flashgemm-autotune v2 gfx950" text table */
int fg_autotune_load(const char *path); /* merges entries; refuses bad entries and
                                           other versions, leaving the table as it was */
int fg_autotune_clear(void);
int fg_autotune_size(void);

/* ---- FP64 error oracle (measurement only; never on a product path) ---- */
/* C = A_n ... A_1 B in fp64 on device pointers. */
int fg_f64_chain(double *C, const double *B, long N, int num_gemm, const double *const *A, const long *M,
                 const long *K, void *stream);
/* Convert n elements of x (x_kind: 0 fp32, 1 bf16, 2 int32, 3 fp64) to fp64. */
int fg_to_f64(int x_kind, const void *x, double *out, long n, void *stream);
/* out[0] = sum (x - ref)^2, out[1] = sum ref^2 in fp64, fixed reduction order.
 * out is a host pointer; the call synchronizes the stream. */
int fg_error_norms(int x_kind, const void *x, const double *ref, long n, double *out, void *stream);

/* ---- Allocation-failure injection (tests of the error paths; no counterpart in the
 * reference, whose calls abort on a failed posix_memalign) ----
 * The next `count` device allocations the library makes (per-stream scratch, grouped
 * launch tables, autotune buffers, a server's packed weight copies) fail as if the
 * device were out of memory: the call that needed one returns FG_ERR_NOMEM and leaves
 * nothing half-built, so the same call succeeds once the injection is spent.
 * count <= 0 clears it. Returns FG_OK. */
int fg_inject_alloc_failures(int count);

/* ---- L2 -> CU stream rate (measurement only; no counterpart in the reference) ----
 * `workgroups` workgroups of 8 waves; wave w of each streams its own slice_kb KB slice
 * of a region the workgroups share (8 slices), 1 KB per load (16 B per lane), 16
 * loads in flight, 16 * (rounds + 1) loads per wave: the access pattern of the fused
 * chains' weight stream. *ms = elapsed time of the launch (HIP events on `stream`);
 * bytes per workgroup = 8 * 16 * (rounds + 1) KB. */
int fg_measure_stream(long slice_kb, int rounds, int workgroups, float *ms, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FLASHGEMM_H */
