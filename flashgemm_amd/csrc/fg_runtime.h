// Host-runtime helpers shared by api.cpp and the grouped / serving translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "fg_internal.h"

namespace fg {

// Execution plan of a chain: consecutive segments of GEMMs, each either one fused
// launch at tile width tile[s] (its intermediates on chip; a segment that ends before
// the chain does writes the next operand to HBM) or, with tile[s] == 0, one unfused
// GEMM. One segment with a tile is the whole chain on chip.
struct ChainPlan {
  int nseg;
  int end[CHAIN_MAXG];   // last GEMM (inclusive) of segment s
  int tile[CHAIN_MAXG];  // fused tile width; 0: a single unfused GEMM
  int splits;            // split-K of the unfused GEMMs (0: planner)
};
// True when `p` can run this exact shape: segments cover the chain in order, unfused
// segments hold one GEMM, and every fused segment fits LDS at its tile width.
bool plan_valid(int kind, long N, int n, const long *M, const long *K, const ChainPlan &p);
// Model-based plan (no table): the whole chain fused when it fits; otherwise the
// segmentation with the lowest modelled time (fused runs where they fit, spills between).
void plan_segments(int kind, long N, int n, const long *M, const long *K, ChainPlan *p);
// Autotune table (autotune.cpp): the measured plan of the chain's shape bucket.
bool autotune_lookup_plan(int kind, long N, int n, const long *M, const long *K, ChainPlan *p);
// Runs a plan on device pointers (mathematical layout, plain epilogues): fused
// segments pack their weights unless a_packed (then every segment must be fused).
int run_plan(int kind, int dtype, const ChainPlan &p, void *C, const void *B, long N, int n, const void *const *A,
             const long *M, const long *K, int a_packed, hipStream_t s);

int set_error(int status, const char *fmt, ...);
int hip_status(hipError_t e, const char *what);
int ensure_init();
int num_cus();
// Named per-stream scratch buffer of at least `bytes` (grow-only).
// Device allocation through the library's allocator: the next n calls fail as if the
// device were out of memory after fg_inject_alloc_failures(n).
hipError_t dev_alloc(void **p, size_t bytes);
int scratch(hipStream_t s, const char *name, size_t bytes, void **out);
// Frees the scratch buffers and the grouped-table ring of stream s (s synchronized).
void release_stream_scratch(hipStream_t s);
void release_stream_ring(hipStream_t s);
// Upper bound of the scratch a chain takes outside a grouped launch (api.cpp).
size_t chain_scratch_bound(int dtype, int f32_mode, long N, int n, const long *M, const long *K, bool single_split);
// fg_sgemm with its split-K slabs in a caller-given workspace (api.cpp).
int sgemm_ws(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode, int splits, void *stream,
             void *ws, size_t ws_bytes);
int kind_of(int dtype, int f32_mode);
// FG_ERR_INVALID (with the message set) unless f32_mode is one of FG_F32_*.
int check_f32_mode(int f32_mode);
int plan_splits(long M, long N, long K, int kind);
int plan_tile_n(int kind, long N, int num_gemm, const long *M, const long *K, bool epi = false);

}  // namespace fg
