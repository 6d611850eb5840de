// Host-runtime helpers shared by api.cpp and the grouped / serving translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace fg {

int set_error(int status, const char *fmt, ...);
int hip_status(hipError_t e, const char *what);
int ensure_init();
int num_cus();
// Named per-stream scratch buffer of at least `bytes` (grow-only).
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
int plan_splits(long M, long N, long K, int kind);
int plan_tile_n(int kind, long N, int num_gemm, const long *M, const long *K, bool epi = false);
// Autotune table lookup (autotune.cpp): true when the shape has a measured decision.
bool autotune_lookup(int kind, long N, int n, const long *M, const long *K, int *fused, int *tile_n, int *splits);

}  // namespace fg
