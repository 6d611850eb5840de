// fg_sgemm_noalloc: the FP32 GEMM entry point for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc). It has no counterpart in the reference, whose drivers allocate per call.
//
// The split-K slabs that fg_sgemm takes from the stream's scratch cache live here in a
// file-scope device array of FG_NOALLOC_SCRATCH_BYTES. Every GPU the library runs on
// holds its own copy until the process ends. In host code `scratch` names a host-side
// stand-in, not a device address: its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is passed to
// the kernels. The kernels overwrite every slab element they later read, so nothing
// needs clearing between calls; all calls on one GPU share the array, so calls that
// overlap on one GPU (different streams, or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/flashgemm.h"
#include "fg_runtime.h"

namespace {

static __device__ __align__(256) char scratch[FG_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES];  // device address of `scratch`, per device (null: not looked up yet)

int scratch_ptr(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fg::hip_status(e, "hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES) return fg::set_error(FG_ERR_NODEVICE, "device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fg::hip_status(e, "hipGetSymbolAddress(scratch)");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return FG_OK;
}

}  // namespace

extern "C" int fg_sgemm_noalloc(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode,
                                int splits, void *stream) {
  int rc;
  if ((rc = fg::ensure_init())) return rc;
  void *ws = nullptr;
  if ((rc = scratch_ptr(&ws))) return rc;
  return fg::sgemm_ws(C, A, B, M, N, K, f32_mode, splits, stream, ws, FG_NOALLOC_SCRATCH_BYTES);
}
