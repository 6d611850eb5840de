// C ABI of libflashgemm: the reference's entry points (part 1 of flashgemm.h) and
// the status-code API over device pointers (part 2).
//
// The reference driver (multigemm/src/flashgemm.c:24-424) partitions N over OpenMP
// threads, packs A, and runs the chain panel by panel on the host. Here the same
// call becomes: validate, plan (one fused launch when every intermediate panel fits
// LDS, otherwise one tiled launch per GEMM with the intermediate in HBM), enqueue on
// a HIP stream. Host-pointer calls copy in and out around that.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flashgemm.h"
#include "fg_internal.h"
#include "fg_runtime.h"

using namespace fg;

// ------------------------------------------------------------ error state --
namespace fg {
static thread_local std::string g_last_error;

int set_error(int status, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

int hip_status(hipError_t e, const char *what) {
  if (e == hipSuccess) return FG_OK;
  if (e == hipErrorOutOfMemory) return set_error(FG_ERR_NOMEM, "%s: %s", what, hipGetErrorString(e));
  return set_error(FG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------- device ---
static std::mutex g_mu;
static int g_device = -1;
static int g_num_cu = 256;

int ensure_init() {
  if (g_device >= 0) return FG_OK;
  return fg_init(0);
}

int num_cus() { return g_num_cu; }

// ------------------------------------------------------- device allocation --
// Every device allocation the library makes goes through dev_alloc, so a test can
// make the next n of them fail as if the device were out of memory
// (fg_inject_alloc_failures) and check that each caller reports FG_ERR_NOMEM and
// leaves its state consistent.
static std::atomic<int> g_inject_alloc{0};

hipError_t dev_alloc(void **p, size_t bytes) {
  int n = g_inject_alloc.load();
  while (n > 0)
    if (g_inject_alloc.compare_exchange_weak(n, n - 1)) return hipErrorOutOfMemory;
  return hipMalloc(p, bytes);
}

// ------------------------------------------------------- scratch per stream --
// Each stream owns named device buffers that only grow. Work on one stream is
// ordered, so reusing its buffers across calls is safe; different streams never
// share a buffer.
struct StreamScratch {
  std::map<std::string, std::pair<void *, size_t>> bufs;
};
static std::map<hipStream_t, StreamScratch> g_scratch;

int scratch(hipStream_t s, const char *name, size_t bytes, void **out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto &slot = g_scratch[s].bufs[name];
  if (slot.second < bytes) {
    const size_t old = slot.second;
    if (slot.first) {
      // the old buffer may still be read by queued work on this stream
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) return hip_status(e, "hipStreamSynchronize");
      (void)hipFree(slot.first);
      slot.first = nullptr;
      slot.second = 0;
    }
    // grow geometrically (at least double, whole MiB past one): a growth synchronizes
    // the stream, and workloads whose sizes creep up (a serving engine's batches)
    // would otherwise stall on one growth after another
    size_t cap = bytes < 256 ? 256 : bytes;
    if (cap < 2 * old) cap = 2 * old;
    if (cap > (1u << 20)) cap = (cap + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    hipError_t e = dev_alloc(&slot.first, cap);
    if (e != hipSuccess) {
      slot.first = nullptr;
      return hip_status(e, "hipMalloc(scratch)");
    }
    slot.second = cap;
  }
  *out = slot.first;
  return FG_OK;
}

// Frees every scratch buffer of stream s (the caller has synchronized s; called
// before a stream the library created is destroyed, so its buffers do not outlive it).
void release_stream_scratch(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_scratch.find(s);
  if (it == g_scratch.end()) return;
  for (auto &kv : it->second.bufs)
    if (kv.second.first) (void)hipFree(kv.second.first);
  g_scratch.erase(it);
}

size_t dtype_bytes(int dtype) { return dtype == FG_DTYPE_F32 ? 4 : dtype == FG_DTYPE_BF16 ? 2 : 1; }

int kind_of(int dtype, int f32_mode) {
  if (dtype == FG_DTYPE_BF16) return K_BF16;
  if (dtype == FG_DTYPE_INT8) return K_I8;
  return f32_mode == FG_F32_BF16X3 ? K_F32X3 : (f32_mode == FG_F32_BF16X6 ? K_F32X6 : K_F32);
}

int check_f32_mode(int f32_mode) {
  if (f32_mode < FG_F32_MFMA || f32_mode > FG_F32_BF16X6) return set_error(FG_ERR_INVALID, "f32_mode %d", f32_mode);
  return FG_OK;
}

// ----------------------------------------------------------------- planner --
int plan_splits(long M, long N, long K, int kind) {
  long tiles = ((M + 127) / 128) * ((N + 127) / 128);
  if (tiles >= (long)num_cus() * 3 / 4) return 1;
  long bk = (kind == K_I8) ? 128 : (kind == K_BF16 ? 64 : 32);
  long want = (num_cus() + tiles - 1) / tiles;
  long maxs = K / (4 * bk);
  if (maxs < 1) maxs = 1;
  long s = want < maxs ? want : maxs;
  return (int)(s < 1 ? 1 : s);
}

// Estimated cycles of a fused plan on one CU. Tiles run in rounds of one per CU. A
// tile costs about 10^4 cycles of staging, barriers and epilogues plus, per GEMM, the
// larger of two bounds: the busiest wave's MFMA issue (passes x k-steps x its row x
// column blocks, at the MFMA's cycles per block times the two waves that share a
// SIMD; INT8/BF16 measured ~1.5x the MFMA rate beside the weight stream) and the
// weight stream itself: ~52 bytes per cycle per CU from L2 (measured for the pattern)
// with all waves on all columns; with the waves split over G column groups every
// group reads all of A_i, and measured conv-shaped chains fit half that rate.
static double fused_cost(int kind, long N, int num_gemm, const long *M, const long *K, int tn) {
  const int groups = chain_fused_wave_groups(kind, tn);
  const int rows = chain_fused_rows_per_wave(kind);
  const long pass = (long)(8 / groups) * rows;
  const int blocks = (rows / 32) * (tn / 32 / groups);
  const long kc = kind == K_F32 ? 8 : (kind == K_I8 ? 32 : 16);
  const double unit = kind == K_F32 ? 4 * 64 * 2 : (kind == K_F32X3 ? 3 * 32 * 2 : 32 * 2 * 1.5);
  const double esz = kind == K_I8 ? 1 : (kind == K_BF16 ? 2 : 4);  // packed bytes per weight element
  double t = 1e4;
  for (int i = 0; i < num_gemm; ++i) {
    const double mfma = (double)((M[i] + pass - 1) / pass) * ((K[i] + kc - 1) / kc) * blocks * unit;
    const double stream = (double)groups * ((M[i] + rows - 1) / rows * rows) * ((K[i] + kc - 1) / kc * kc) * esz /
                          (groups > 1 ? 26.0 : 52.0);
    t += mfma > stream ? mfma : stream;
  }
  const long tiles = (N + tn - 1) / tn;
  const long rounds = (tiles + num_cus() - 1) / num_cus();
  return (double)rounds * t;
}

// Upper bound of the per-stream scratch one chain can take when it does not run in a
// grouped launch: the two spill buffers of the unfused path, the largest split-K slab
// set (one slab per split, skinny FP32 GEMMs included; none when single_split), the
// INT8 column sums, and the packed weights of a fused run.
size_t chain_scratch_bound(int dtype, int f32_mode, long N, int n, const long *M, const long *K, bool single_split) {
  const int kind = kind_of(dtype, f32_mode);
  long maxm = 0;
  for (int i = 0; i + 1 < n; ++i) maxm = M[i] > maxm ? M[i] : maxm;
  size_t bytes = (size_t)2 * maxm * N * dtype_bytes(dtype), part = 0, packed = 0;
  for (int i = 0; i < n; ++i) {
    long sk = 1;
    if (!single_split) {
      long kper;
      sk = (kind == K_F32 && skinny_f32_shape(M[i], N)) ? skinny_f32_splits(M[i], N, K[i], &kper)
                                                         : plan_splits(M[i], N, K[i], kind);
    }
    if (sk > 1) part = std::max(part, (size_t)sk * M[i] * N * 4);
    packed += (packed_a_bytes(kind, M[i], K[i]) + 255) & ~(size_t)255;
  }
  return bytes + part + packed + (kind == K_I8 ? (size_t)N * 4 : 0);
}

// Fused tile width: the cheapest tile (fused_cost) that fits LDS; within 3% the one
// with more workgroups. 0 when none fits.
int plan_tile_n(int kind, long N, int num_gemm, const long *M, const long *K, bool epi) {
  if (num_gemm > chain_fused_max_gemms() || !chain_fused_shape_ok(kind, N, num_gemm, K)) return 0;
  int best = 0;
  double best_cost = 0;
  for (int c : {32, 64, 96, 128, 256}) {  // narrow first: ties keep more workgroups
    if (!chain_fused_tile_ok(kind, c) || chain_fused_lds_bytes(kind, M, K, num_gemm, c, epi) > LDS_BYTES_MAX)
      continue;
    const double cost = fused_cost(kind, N, num_gemm, M, K, c);
    if (best == 0 || cost < best_cost * 0.97) {
      best = c;
      best_cost = cost;
    }
  }
  return best;
}

// ---------------------------------------------------------------- segments --
// A dependent kernel launch costs ~2.5 us back to back (~6000 cycles at 2.4 GHz).
constexpr double LAUNCH_CYCLES = 6000.0;
// HBM per CU and clock at 8 TB/s over 256 CUs at 2.4 GHz.
constexpr double HBM_BYTES_PER_CU_CLK = 13.0;

static double elem_bytes(int kind) { return kind == K_I8 ? 1 : (kind == K_BF16 ? 2 : 4); }

// Modelled time (chip cycles) of one unfused GEMM: MFMA at ~60 % of the kind's peak
// or its operand and output bytes at HBM speed, whichever is longer, plus a launch.
static double gemm_cycles(int kind, long M, long N, long K, bool last) {
  const double rate = kind == K_F32 ? 256.0 : (kind == K_F32X3 ? 4096.0 / 3 : (kind == K_BF16 ? 4096.0 : 8192.0));
  const double e = elem_bytes(kind);
  const double flops = 2.0 * M * N * K, bytes = e * ((double)M * K + (double)K * N) + (last ? 4.0 : e) * M * N;
  const double mfma = flops / (num_cus() * rate * 0.6), hbm = bytes / (num_cus() * HBM_BYTES_PER_CU_CLK);
  return (mfma > hbm ? mfma : hbm) + LAUNCH_CYCLES;
}

// Modelled time of a fused segment: fused_cost plus its input and output at HBM speed
// (the fused kernel stages them before its first and after its last MFMA) and a launch.
static double segment_cycles(int kind, long N, int n, const long *M, const long *K, int tn, bool last) {
  const double e = elem_bytes(kind);
  const double io = e * K[0] * N + (last ? 4.0 : e) * M[n - 1] * N;
  return fused_cost(kind, N, n, M, K, tn) + io / (num_cus() * HBM_BYTES_PER_CU_CLK) + LAUNCH_CYCLES;
}

bool plan_valid(int kind, long N, int n, const long *M, const long *K, const ChainPlan &p) {
  if (p.nseg < 1 || p.nseg > n || p.splits < 0) return false;
  int start = 0;
  for (int s = 0; s < p.nseg; ++s) {
    const int end = p.end[s], len = end - start + 1;
    if (end < start || end >= n) return false;
    if (p.tile[s] == 0) {
      if (len != 1) return false;
    } else if (!chain_fused_tile_ok(kind, p.tile[s]) || !chain_fused_shape_ok(kind, N, len, K + start) ||
               chain_fused_lds_bytes(kind, M + start, K + start, len, p.tile[s]) > LDS_BYTES_MAX) {
      return false;
    }
    start = end + 1;
  }
  return start == n;
}

// Shortest path over GEMM boundaries: segment [i, j) is one unfused GEMM (j = i + 1)
// or, for two GEMMs or more, a fused run at plan_tile_n when it fits LDS. The fewest
// segments win (every boundary spills an intermediate to HBM and back, which is what
// fusion saves), the modelled time breaks ties; the autotuner measures the rest.
void plan_segments(int kind, long N, int n, const long *M, const long *K, ChainPlan *p) {
  constexpr double PER_SEGMENT = 1e15;
  double best[CHAIN_MAXG + 1];
  int from[CHAIN_MAXG + 1], tile[CHAIN_MAXG + 1];
  best[0] = 0;
  for (int j = 1; j <= n; ++j) {
    best[j] = 1e300;
    for (int i = 0; i < j; ++i) {
      const int len = j - i;
      const bool last = j == n;
      double c = 1e300;
      int t = 0;
      if (len == 1) c = gemm_cycles(kind, M[i], N, K[i], last);
      const int tn = len >= 2 && chain_fused_shape_ok(kind, N, len, K + i) ? plan_tile_n(kind, N, len, M + i, K + i)
                                                                         : 0;
      if (tn > 0) {
        const double cf = segment_cycles(kind, N, len, M + i, K + i, tn, last);
        if (cf < c) {
          c = cf;
          t = tn;
        }
      }
      c += PER_SEGMENT;
      if (best[i] + c < best[j]) {
        best[j] = best[i] + c;
        from[j] = i;
        tile[j] = t;
      }
    }
  }
  int ends[CHAIN_MAXG], tiles[CHAIN_MAXG], ns = 0;
  for (int j = n; j > 0; j = from[j]) {
    ends[ns] = j - 1;
    tiles[ns] = tile[j];
    ++ns;
  }
  p->nseg = ns;
  for (int s = 0; s < ns; ++s) {  // back to chain order
    p->end[s] = ends[ns - 1 - s];
    p->tile[s] = tiles[ns - 1 - s];
  }
  p->splits = 0;
}

}  // namespace fg

// ================================================================ part 2 ===

extern "C" {

void fg_chain_opts_default(fg_chain_opts *o) {
  o->layout = FG_LAYOUT_MATH;
  o->ref_threads = 1;
  o->f32_mode = FG_F32_MFMA;
  o->fuse = FG_FUSE_AUTO;
  o->tile_n = 0;
  o->splits = 0;
  o->a_packed = 0;
}

size_t fg_packed_weight_bytes(int dtype, int f32_mode, long M, long K) {
  if (dtype < FG_DTYPE_F32 || dtype > FG_DTYPE_INT8 || M <= 0 || K <= 0) return 0;
  if (kind_of(dtype, f32_mode) == K_F32X6) return 0;  // unfused only: no packed weights
  return packed_a_bytes(kind_of(dtype, f32_mode), M, K);
}

int fg_pack_weight(int dtype, int f32_mode, const void *A, long M, long K, void *packed, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (dtype < FG_DTYPE_F32 || dtype > FG_DTYPE_INT8 || !A || !packed || M <= 0 || K <= 0)
    return set_error(FG_ERR_INVALID, "pack_weight: dtype=%d M=%ld K=%ld", dtype, M, K);
  if ((rc = check_f32_mode(f32_mode))) return rc;
  if (kind_of(dtype, f32_mode) == K_F32X6)
    return set_error(FG_ERR_UNSUPPORTED, "the three-piece split runs unfused: no packed weights");
  PackJob J{};
  J.n = 1;
  J.A[0] = A;
  J.M[0] = M;
  J.K[0] = K;
  J.out[0] = packed;
  return hip_status(launch_pack_a(kind_of(dtype, f32_mode), J, (hipStream_t)stream), "pack weight");
}

const char *fg_status_string(int s) {
  switch (s) {
    case FG_OK: return "ok";
    case FG_ERR_INVALID: return "invalid argument";
    case FG_ERR_DOMAIN: return "reference layout undefined for this shape";
    case FG_ERR_NOMEM: return "out of device memory";
    case FG_ERR_HIP: return "HIP error";
    case FG_ERR_BUDGET: return "over HBM budget";
    case FG_ERR_NODEVICE: return "no gfx950 device";
    case FG_ERR_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char *fg_last_error(void) { return fg::g_last_error.c_str(); }

int fg_init(int device) {
  std::lock_guard<std::mutex> lk(fg::g_mu);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= device || device < 0)
    return set_error(FG_ERR_NODEVICE, "no HIP device %d (%d visible)", device, n);
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, device);
  if (e != hipSuccess) return hip_status(e, "hipGetDeviceProperties");
  if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return set_error(FG_ERR_NODEVICE, "device %d is %s, not gfx950", device, p.gcnArchName);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_status(e, "hipSetDevice");
  fg::g_device = device;
  fg::g_num_cu = p.multiProcessorCount;
  return FG_OK;
}

int fg_synchronize(void *stream) {
  return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

// check_a: the weight pointers are part of the call (not for plan inspection).
static int validate_chain(int dtype, long N, int num_gemm, const void *const *A, const long *M, const long *K,
                          bool check_a = true) {
  if (dtype < FG_DTYPE_F32 || dtype > FG_DTYPE_INT8) return set_error(FG_ERR_INVALID, "dtype %d", dtype);
  if (N <= 0) return set_error(FG_ERR_INVALID, "N = %ld", N);
  if (num_gemm < 1 || num_gemm > 8) return set_error(FG_ERR_INVALID, "num_gemm = %d (1..8)", num_gemm);
  if (!M || !K) return set_error(FG_ERR_INVALID, "M or K is NULL");
  for (int i = 0; i < num_gemm; ++i) {
    if (check_a && (!A || !A[i])) return set_error(FG_ERR_INVALID, "A[%d] is NULL", i);
    if (M[i] <= 0 || K[i] <= 0) return set_error(FG_ERR_INVALID, "gemm %d: M=%ld K=%ld", i + 1, M[i], K[i]);
    if (i > 0 && K[i] != M[i - 1])
      return set_error(FG_ERR_INVALID, "gemm %d: K=%ld but the previous M=%ld", i + 1, K[i], M[i - 1]);
  }
  return FG_OK;
}

// Reference panels for `threads` threads; FG_ERR_DOMAIN when a panel width is not
// 16 or 32 (outside the reference's domain, where its layout is undefined).
static int ref_panels(long N, int threads, RefPanels *p) {
  if (threads < 1) return set_error(FG_ERR_INVALID, "ref_threads = %d", threads);
  long groups = (N + 15) / 16;
  long nb = (groups % threads == 0) ? (groups / threads) * 16 : ((groups + threads - 1) / threads) * 16;
  long ne = N % nb;
  long num_n = N / nb + (ne == 0 ? 0 : 1);
  for (long id = 0; id < num_n; ++id) {
    long NB = (id < num_n - 1) ? nb : (ne > 0 ? ne : nb);
    long tail = NB % 32;
    if (tail != 0 && tail != 16)
      return set_error(FG_ERR_DOMAIN, "reference panel of width %ld (N=%ld, %d threads)", tail, N, threads);
  }
  p->nb = nb;
  p->ne = ne;
  p->num_n = num_n;
  p->N = N;
  return FG_OK;
}

// Decision order: explicit options, then the autotune table, then the heuristics
// (widest tile that fits and still fills the CUs; split-K for skinny GEMMs).
// epi: the chain has epilogue extensions (their per-row tables take LDS too).
// The reference layout runs fused (the epilogues write the reference's permuted
// intermediates, *refl = 1) when every reference panel is a 32-aligned block and every
// intermediate M_i a multiple of 32 (chain_refl_ok), without epilogue extensions or an
// image B (img); otherwise it runs unfused with a gather after each GEMM.
// The plan: one fused segment (the whole chain on chip), or, when it does not fit,
// the segments of plan_segments (mathematical layout, plain epilogues, matrix B) or one
// unfused GEMM per segment (the reference layout's gather path, epilogue extensions,
// options that ask for it). An autotune entry for the chain's shape bucket decides
// first when the options leave the choice open and it can run this exact shape.
static int plan_chain(int dtype, long N, int num_gemm, const long *M, const long *K, const fg_chain_opts *opts,
                      ChainPlan *pl, bool epi = false, bool img = false, int *refl_out = nullptr) {
  fg_chain_opts o;
  if (opts) o = *opts; else fg_chain_opts_default(&o);
  int kind = kind_of(dtype, o.f32_mode);
  if (kind == K_F32X6) {  // the three-piece split runs unfused, one BF16 GEMM per chain GEMM
    if (o.fuse == FG_FUSE_ON || o.tile_n > 0) return set_error(FG_ERR_UNSUPPORTED, "the three-piece split runs unfused");
    if (refl_out) *refl_out = 0;
    pl->splits = o.splits;
    pl->nseg = num_gemm;
    for (int i = 0; i < num_gemm; ++i) {
      pl->end[i] = i;
      pl->tile[i] = 0;
    }
    return FG_OK;
  }
  const bool any_ref = o.layout == FG_LAYOUT_REFERENCE && dtype != FG_DTYPE_F32;
  bool refl = false;
  if (any_ref && !epi && !img) {
    RefPanels p;
    refl = ref_panels(N, o.ref_threads, &p) == FG_OK && chain_refl_ok(kind, p, num_gemm, M);
  }
  if (refl_out) *refl_out = refl;
  const bool ref = any_ref && !refl;  // the reference layout through the unfused gather path
  pl->splits = o.splits;
  if (!any_ref && !epi && !img && o.fuse == FG_FUSE_AUTO && o.tile_n == 0 && o.splits == 0) {
    // an entry that cannot run this exact shape (a tile width the kind lacks, an
    // arena beyond LDS) is ignored: the heuristics below decide instead
    ChainPlan t;
    if (autotune_lookup_plan(kind, N, num_gemm, M, K, &t) && plan_valid(kind, N, num_gemm, M, K, t)) {
      *pl = t;
      return FG_OK;
    }
  }
  int tn = 0;
  if (!ref && o.fuse != FG_FUSE_OFF) {
    if (o.tile_n > 0) {
      if (chain_fused_tile_ok(kind, o.tile_n) && chain_fused_shape_ok(kind, N, num_gemm, K) &&
          chain_fused_lds_bytes(kind, M, K, num_gemm, o.tile_n, epi) <= LDS_BYTES_MAX)
        tn = o.tile_n;
    } else {
      tn = plan_tile_n(kind, N, num_gemm, M, K, epi);
    }
  }
  if (o.fuse == FG_FUSE_ON && tn == 0)
    return set_error(FG_ERR_UNSUPPORTED, "fused chain requested but %s",
                     ref ? "the reference layout runs unfused" : "no tile width fits LDS");
  if (tn > 0) {
    pl->nseg = 1;
    pl->end[0] = num_gemm - 1;
    pl->tile[0] = tn;
    return FG_OK;
  }
  if (!ref && !epi && !img && o.fuse == FG_FUSE_AUTO && o.tile_n == 0 &&
      chain_fused_shape_ok(kind, N, num_gemm, K)) {
    const int splits = pl->splits;
    plan_segments(kind, N, num_gemm, M, K, pl);
    pl->splits = splits;
    return FG_OK;
  }
  pl->nseg = num_gemm;
  for (int i = 0; i < num_gemm; ++i) {
    pl->end[i] = i;
    pl->tile[i] = 0;
  }
  return FG_OK;
}

static bool whole_fused(const ChainPlan &p) { return p.nseg == 1 && p.tile[0] > 0; }

int fg_chain_plan(int dtype, long N, int num_gemm, const long *M, const long *K, const fg_chain_opts *opts,
                  int *fused, int *tile_n) {
  ChainPlan p;
  int rc = plan_chain(dtype, N, num_gemm, M, K, opts, &p);
  if (rc) return rc;
  *fused = whole_fused(p);
  *tile_n = *fused ? p.tile[0] : 0;
  return FG_OK;
}

int fg_chain_plan_segments(int dtype, long N, int num_gemm, const long *M, const long *K, const fg_chain_opts *opts,
                           int *num_segments, int *seg_end, int *seg_tile) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, nullptr, M, K, false))) return rc;
  ChainPlan p;
  if ((rc = plan_chain(dtype, N, num_gemm, M, K, opts, &p))) return rc;
  *num_segments = p.nseg;
  for (int s = 0; s < p.nseg; ++s) {
    seg_end[s] = p.end[s];
    seg_tile[s] = p.tile[s];
  }
  return FG_OK;
}

// One GEMM (+ INT8 column sums, + split-K reduction) of the unfused path. With
// epilogue extensions (ex != nullptr) the GEMM always writes fp32 / int32 slabs and
// the reduction pass applies the extensions while it forms the output.
// trans (FP32 kinds): bit 0 A stored K x M, bit 1 B stored N x K.
// ws: a caller-owned workspace of ws_bytes for the split-K slabs (FP32 and BF16
// kinds), used instead of the stream's scratch so that nothing is allocated; a plan
// that needs more than ws_bytes is refused with FG_ERR_INVALID.
static int run_gemm(int kind, int epi, const void *A, const void *B, void *C, long M, long N, long K, int splits,
                    hipStream_t s, const EpiOne *ex = nullptr, int trans = 0, void *ws = nullptr,
                    size_t ws_bytes = 0, int kgroups = 0) {
  int *colsum = nullptr;
  int rc;
  if (kind == K_F32X6) {
    // three-piece split: one BF16 GEMM over k' = 6K on the split operands, fp32 out
    // (the FP32 kinds' requantization keeps fp32)
    if (trans) return set_error(FG_ERR_UNSUPPORTED, "the three-piece split takes NN operands only");
    if (ws) return set_error(FG_ERR_UNSUPPORTED, "the three-piece split needs scratch for its split operands");
    uint16_t *A6 = nullptr, *B6 = nullptr;
    if ((rc = scratch(s, "x6a", (size_t)M * 6 * K * 2, (void **)&A6))) return rc;
    if ((rc = scratch(s, "x6b", (size_t)6 * K * N * 2, (void **)&B6))) return rc;
    if ((rc = hip_status(launch_split3((const float *)A, M, K, A6, (const float *)B, N, B6, s), "split3"))) return rc;
    return run_gemm(K_BF16, epi == EPI_REQUANT ? EPI_STORE : epi, A6, B6, C, M, N, 6 * K, splits, s, ex);
  }
  auto partials = [&](size_t bytes, void **p) -> int {
    if (!ws) return scratch(s, "partials", bytes, p);
    if (bytes > ws_bytes)
      return set_error(FG_ERR_INVALID, "split-K slabs need %zu bytes, the workspace holds %zu (M=%ld N=%ld K=%ld)",
                       bytes, ws_bytes, M, N, K);
    *p = ws;
    return FG_OK;
  };
  if (kind == K_F32 && epi != EPI_PARTIAL && splits <= 0 && trans == 0 && skinny_f32_shape(M, N)) {
    // small M or N: 64 x 64 tiles, split-K over the grid
    long kper = 0;
    const int sk = skinny_f32_splits(M, N, K, &kper);
    float *out = (float *)C;
    if ((sk > 1 || ex) && (rc = partials((size_t)sk * M * N * 4, (void **)&out))) return rc;
    if ((rc = hip_status(launch_skinny_f32((const float *)A, (const float *)B, out, M, N, K, kper, sk, s),
                         "skinny gemm")))
      return rc;
    if (sk == 1 && !ex) return FG_OK;
    return hip_status(launch_splitk_reduce(kind, EPI_STORE, out, sk, M, N, nullptr, C, ex, s), "split-K reduce");
  }
  if (splits <= 0) splits = plan_splits(M, N, K, kind);
  if (splits <= 1 && !ex) {  // INT8: the kernel sums its own B columns (colsum == nullptr)
    GemmArgs g{kind, epi, A, B, C, nullptr, M, N, K, 1, trans, kgroups};
    return hip_status(launch_gemm(g, s), "gemm");
  }
  if (kind == K_I8) {  // split-K slabs see part of K: the reduction adds the full column sums
    if ((rc = scratch(s, "colsum", sizeof(int) * N, (void **)&colsum))) return rc;
    if ((rc = hip_status(launch_colsum_i8((const int8_t *)B, K, N, colsum, s), "colsum"))) return rc;
  }
  void *part = nullptr;
  if ((rc = partials((size_t)splits * M * N * 4, &part))) return rc;
  GemmArgs g{kind, EPI_PARTIAL, A, B, part, colsum, M, N, K, splits, trans};
  if ((rc = hip_status(launch_gemm(g, s), "gemm(split-K)"))) return rc;
  // the kernel may use fewer slabs than asked when K is short
  long bk = (kind == K_I8) ? 128 : (kind == K_BF16 ? 64 : 32);
  long kper = ((K + splits - 1) / splits + bk - 1) / bk * bk;
  int used = (int)((K + kper - 1) / kper);
  return hip_status(launch_splitk_reduce(kind, epi, part, used, M, N, colsum, C, ex, s), "split-K reduce");
}

// Arguments of a fused launch (A[i] as given; run_fused packs them unless a_packed).
static ChainArgs chain_args(int kind, const void *B, void *C, long N, int num_gemm, const void *const *A,
                            const long *M, const long *K, int tile_n) {
  ChainArgs c{};
  c.kind = kind;
  c.num_gemm = num_gemm;
  for (int i = 0; i < num_gemm; ++i) {
    c.A[i] = A[i];
    c.M[i] = M[i];
    c.K[i] = K[i];
  }
  c.B = B;
  c.C = C;
  c.N = N;
  c.tile_n = tile_n;
  return c;
}

// Fused launch: unless the weights come packed, pack every A_i into MFMA fragment
// order first (per-stream scratch, one launch for the whole chain). FP32 weights that
// need no padding are read as they are (chain_raw_weights_ok): no pack launch.
static int run_fused(ChainArgs c, int a_packed, hipStream_t s) {
  int rc;
  if (!a_packed && chain_raw_weights_ok(c)) {
    c.raw_a = 1;
  } else if (!a_packed) {
    size_t off[CHAIN_MAXG], total = 0;
    for (int i = 0; i < c.num_gemm; ++i) {
      off[i] = total;
      total += (packed_a_bytes(c.kind, c.M[i], c.K[i]) + 255) & ~(size_t)255;
    }
    char *pk = nullptr;
    if ((rc = scratch(s, "apack", total, (void **)&pk))) return rc;
    PackJob J{};
    J.n = c.num_gemm;
    for (int i = 0; i < c.num_gemm; ++i) {
      J.A[i] = c.A[i];
      J.M[i] = c.M[i];
      J.K[i] = c.K[i];
      J.out[i] = pk + off[i];
      c.A[i] = pk + off[i];
    }
    if ((rc = hip_status(launch_pack_a(c.kind, J, s), "pack A"))) return rc;
  }
  return hip_status(launch_chain_fused(c, s), "fused chain");
}

}  // extern "C"

namespace fg {
// Segment s writes its output (the next operand, requantized as the unfused epilogue
// does) to spill buffer s % 2 and reads its input from the other one (or B).
int run_plan(int kind, int dtype, const ChainPlan &p, void *C, const void *B, long N, int n, const void *const *A,
             const long *M, const long *K, int a_packed, hipStream_t s) {
  int rc;
  if (!plan_valid(kind, N, n, M, K, p)) return set_error(FG_ERR_INVALID, "chain plan does not fit this chain");
  for (int q = 0; q < p.nseg; ++q)
    if (a_packed && p.tile[q] == 0) return set_error(FG_ERR_UNSUPPORTED, "packed weights need the fused path");
  const size_t e = dtype_bytes(dtype);
  long maxm = 0;
  for (int q = 0; q + 1 < p.nseg; ++q) maxm = M[p.end[q]] > maxm ? M[p.end[q]] : maxm;
  void *buf[2] = {nullptr, nullptr};
  if (p.nseg > 1)
    for (int b = 0; b < 2; ++b)
      if ((rc = scratch(s, b ? "spill1" : "spill0", (size_t)maxm * N * e, &buf[b]))) return rc;
  const void *X = B;
  int g0 = 0;
  for (int q = 0; q < p.nseg; ++q) {
    const int g1 = p.end[q];
    const bool last = g1 == n - 1;
    void *Y = last ? C : buf[q & 1];
    if (p.tile[q] > 0) {
      ChainArgs c = chain_args(kind, X, Y, N, g1 - g0 + 1, A + g0, M + g0, K + g0, p.tile[q]);
      c.out_requant = !last;
      ChainParams P;
      make_chain_params(c, &P);
      if (!chain_fused_aligned(P)) return set_error(FG_ERR_UNSUPPORTED, "fused segment needs B aligned to 4 elements");
      if ((rc = run_fused(c, a_packed, s))) return rc;
    } else if ((rc = run_gemm(kind, last ? EPI_STORE : EPI_REQUANT, A[g0], X, Y, M[g0], N, K[g0], p.splits, s))) {
      return rc;
    }
    X = Y;
    g0 = g1 + 1;
  }
  return FG_OK;
}
}  // namespace fg

extern "C" {

// Epilogue extensions from the public struct; *out = nullptr when every field is at
// its default (the reference's epilogues, the plain kernels).
static int epi_params(int dtype, int num_gemm, const fg_epilogue *e, EpiParams *p, const EpiParams **out) {
  *out = nullptr;
  if (!e) return FG_OK;
  bool any = false;
  for (int i = 0; i < num_gemm; ++i) {
    const bool last = i == num_gemm - 1;
    if (dtype == FG_DTYPE_INT8 && !last && (e->zero_point[i] < -128 || e->zero_point[i] > 127))
      return set_error(FG_ERR_INVALID, "epilogue: zero_point[%d] = %d (-128..127)", i, e->zero_point[i]);
    if ((dtype != FG_DTYPE_INT8 || last) && e->zero_point[i] != 0)
      return set_error(FG_ERR_UNSUPPORTED, "epilogue: zero_point[%d] applies to INT8 intermediates only", i);
    if (dtype == FG_DTYPE_INT8 && last && e->scale[i])
      return set_error(FG_ERR_UNSUPPORTED, "epilogue: the INT8 chain's int32 output takes no scale");
    p->g[i] = EpiOne{e->bias[i], e->scale[i], e->zero_point[i], e->relu[i] != 0};
    any = any || e->bias[i] || e->scale[i] || e->zero_point[i] || e->relu[i];
  }
  if (any) *out = p;
  return FG_OK;
}

static int chain_impl(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                      const long *K, const EpiParams *ep, const fg_chain_opts *opts, void *stream);

int fg_chain(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
             const long *K, const fg_chain_opts *opts, void *stream) {
  return chain_impl(dtype, C, B, N, num_gemm, A, M, K, nullptr, opts, stream);
}

int fg_chain_epi(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                 const long *K, const fg_epilogue *epi, const fg_chain_opts *opts, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, A, M, K))) return rc;
  EpiParams p{};
  const EpiParams *ep;
  if ((rc = epi_params(dtype, num_gemm, epi, &p, &ep))) return rc;
  return chain_impl(dtype, C, B, N, num_gemm, A, M, K, ep, opts, stream);
}

static int chain_impl(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                      const long *K, const EpiParams *ep, const fg_chain_opts *opts, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, A, M, K))) return rc;
  if (!B || !C) return set_error(FG_ERR_INVALID, "B or C is NULL");
  fg_chain_opts o;
  if (opts) o = *opts; else fg_chain_opts_default(&o);
  if ((rc = check_f32_mode(o.f32_mode))) return rc;
  hipStream_t s = (hipStream_t)stream;
  int kind = kind_of(dtype, o.f32_mode);
  bool ref = o.layout == FG_LAYOUT_REFERENCE && dtype != FG_DTYPE_F32;
  RefPanels panels{};
  if (ref && (rc = ref_panels(N, o.ref_threads, &panels))) return rc;

  ChainPlan pl;
  int refl = 0;
  if ((rc = plan_chain(dtype, N, num_gemm, M, K, &o, &pl, ep != nullptr, false, &refl))) return rc;
  const int splits = pl.splits;
  const bool b_aligned = N % 4 == 0 && ((uintptr_t)B % (4 * dtype_bytes(dtype))) == 0;
  if (whole_fused(pl)) {
    ChainArgs c = chain_args(kind, B, C, N, num_gemm, A, M, K, pl.tile[0]);
    c.epi = ep;
    c.refl = refl;
    ChainParams P;
    make_chain_params(c, &P);
    if (chain_fused_aligned(P)) return run_fused(c, o.a_packed, s);
    if (o.fuse == FG_FUSE_ON || o.a_packed)
      return set_error(FG_ERR_UNSUPPORTED, "fused chain requested but B is not aligned to 4 elements");
    // misaligned B: run unfused
  } else if (pl.nseg < num_gemm && (b_aligned || pl.tile[0] == 0)) {
    // fused runs where they fit, spills between (mathematical layout, plain epilogues)
    return run_plan(kind, dtype, pl, C, B, N, num_gemm, A, M, K, o.a_packed, s);
  }
  if (o.a_packed) return set_error(FG_ERR_UNSUPPORTED, "packed weights need the fused path");

  // Unfused: GEMM i writes its requantized output to HBM; the reference layout adds
  // a gather that re-reads it the way the reference's next GEMM does.
  size_t ebytes = (kind == K_I8) ? 1 : (kind == K_BF16 ? 2 : 4);
  long maxm = 0;
  for (int i = 0; i + 1 < num_gemm; ++i) maxm = M[i] > maxm ? M[i] : maxm;
  void *buf[3] = {nullptr, nullptr, nullptr};
  if (num_gemm > 1) {
    for (int b = 0; b < (ref ? 3 : 2); ++b) {
      char name[16];
      snprintf(name, sizeof name, "spill%d", b);
      if ((rc = scratch(s, name, (size_t)maxm * N * ebytes, &buf[b]))) return rc;
    }
  }
  const void *X = B;
  for (int i = 0; i < num_gemm; ++i) {
    bool last = i == num_gemm - 1;
    void *Y = last ? C : buf[i & 1];
    if ((rc = run_gemm(kind, last ? EPI_STORE : EPI_REQUANT, A[i], X, Y, M[i], N, K[i], splits, s,
                       ep ? &ep->g[i] : nullptr)))
      return rc;
    if (!last && ref) {
      if ((rc = hip_status(launch_ref_permute(kind == K_I8, Y, buf[2], M[i], panels, s), "ref permute")))
        return rc;
      // move the permuted copy into the ping-pong slot the next GEMM reads
      void *tmp = buf[i & 1];
      buf[i & 1] = buf[2];
      buf[2] = tmp;
      Y = buf[i & 1];
    }
    X = Y;
  }
  return FG_OK;
}

int fg_sgemm(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode, int splits,
             void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0)
    return set_error(FG_ERR_INVALID, "sgemm: M=%ld N=%ld K=%ld", M, N, K);
  if ((rc = check_f32_mode(f32_mode))) return rc;
  const int kind = kind_of(FG_DTYPE_F32, f32_mode);
  return run_gemm(kind, EPI_STORE, A, B, C, M, N, K, splits, (hipStream_t)stream, nullptr, 0, nullptr, 0, 1);
}

}  // extern "C"

namespace fg {
// fg_sgemm with the split-K slabs in the workspace ws (ws_bytes) instead of the
// stream's scratch: the same plan and kernels, so the same bits, and no allocation.
int sgemm_ws(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode, int splits, void *stream,
             void *ws, size_t ws_bytes) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0)
    return set_error(FG_ERR_INVALID, "sgemm: M=%ld N=%ld K=%ld", M, N, K);
  if ((rc = check_f32_mode(f32_mode))) return rc;
  const int kind = kind_of(FG_DTYPE_F32, f32_mode);
  return run_gemm(kind, EPI_STORE, A, B, C, M, N, K, splits, (hipStream_t)stream, nullptr, 0, ws, ws_bytes, 1);
}
}  // namespace fg

extern "C" {

int fg_sgemm_t(int transa, int transb, float *C, const float *A, const float *B, long M, long N, long K,
               int f32_mode, int splits, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0)
    return set_error(FG_ERR_INVALID, "sgemm: M=%ld N=%ld K=%ld", M, N, K);
  if ((transa != 0 && transa != 1) || (transb != 0 && transb != 1))
    return set_error(FG_ERR_INVALID, "sgemm: transa=%d transb=%d (0 or 1)", transa, transb);
  if ((rc = check_f32_mode(f32_mode))) return rc;
  const int kind = kind_of(FG_DTYPE_F32, f32_mode);
  return run_gemm(kind, EPI_STORE, A, B, C, M, N, K, splits, (hipStream_t)stream, nullptr, transa | (transb << 1),
                  nullptr, 0, 1);
}

// ------------------------------------------------------- convolution front end --

// Checks a convolution geometry and fills the kernels' ImgGeom. The padded image
// must hold fewer than 2^31 elements (every offset the kernels form is 32-bit), and
// K1 and N must stay below 2^31.
static int conv_geom(const fg_conv_geom *g, ImgGeom *ig, long *K1, long *N) {
  constexpr long lim = 1L << 31;
  if (!g) return set_error(FG_ERR_INVALID, "conv geometry is NULL");
  if (g->batch < 1 || g->channels < 1 || g->height < 1 || g->width < 1 || g->kernel_h < 1 || g->kernel_w < 1 ||
      g->stride_h < 1 || g->stride_w < 1 || g->pad_h < 0 || g->pad_w < 0 || g->dilation_h < 1 ||
      g->dilation_w < 1 || g->batch >= lim || g->channels >= lim || g->height >= lim || g->width >= lim ||
      g->kernel_h >= lim || g->kernel_w >= lim || g->stride_h >= lim || g->stride_w >= lim || g->pad_h >= lim ||
      g->pad_w >= lim || g->dilation_h >= lim || g->dilation_w >= lim)
    return set_error(FG_ERR_INVALID, "conv geometry: a size is out of range");
  const long ph = g->height + 2 * g->pad_h, pw = g->width + 2 * g->pad_w;
  const long eh = g->dilation_h * (g->kernel_h - 1) + 1, ew = g->dilation_w * (g->kernel_w - 1) + 1;
  if (ph < eh || pw < ew)
    return set_error(FG_ERR_INVALID, "conv geometry: the %ldx%ld kernel extent exceeds the padded %ldx%ld image", eh,
                     ew, ph, pw);
  const long oh = (ph - eh) / g->stride_h + 1, ow = (pw - ew) / g->stride_w + 1;
  // products checked one factor at a time against 2^31
  auto below = [](long a, long b) { return a < lim && b < lim && (b == 0 || a < lim / b); };
  if (!below(ph, pw) || !below(ph * pw, g->channels) || !below(ph * pw * g->channels, g->batch))
    return set_error(FG_ERR_INVALID, "conv geometry: the padded image has 2^31 elements or more");
  if (!below(oh, ow) || !below(oh * ow, g->batch) || !below(g->kernel_h, g->kernel_w) ||
      !below(g->kernel_h * g->kernel_w, g->channels))
    return set_error(FG_ERR_INVALID, "conv geometry: K1 or N reaches 2^31");
  if (ig) {
    ig->C = (int)g->channels;
    ig->H = (int)g->height;
    ig->W = (int)g->width;
    ig->kh = (int)g->kernel_h;
    ig->kw = (int)g->kernel_w;
    ig->sh = (int)g->stride_h;
    ig->sw = (int)g->stride_w;
    ig->ph = (int)g->pad_h;
    ig->pw = (int)g->pad_w;
    ig->dh = (int)g->dilation_h;
    ig->dw = (int)g->dilation_w;
    ig->OH = (int)oh;
    ig->OW = (int)ow;
    ig->vec4 = 0;
    ig->nq = 0;
  }
  if (K1) *K1 = g->channels * g->kernel_h * g->kernel_w;
  if (N) *N = g->batch * oh * ow;
  return FG_OK;
}

int fg_conv_dims(const fg_conv_geom *g, long *K1, long *N) { return conv_geom(g, nullptr, K1, N); }

int fg_im2col(int dtype, void *B, const void *X, const fg_conv_geom *g, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (dtype < FG_DTYPE_F32 || dtype > FG_DTYPE_INT8) return set_error(FG_ERR_INVALID, "dtype %d", dtype);
  if (!B || !X) return set_error(FG_ERR_INVALID, "im2col: B or X is NULL");
  ImgGeom ig;
  long K1, N;
  if ((rc = conv_geom(g, &ig, &K1, &N))) return rc;
  return hip_status(launch_im2col((int)dtype_bytes(dtype), X, B, ig, K1, N, (hipStream_t)stream), "im2col");
}

int fg_chain_conv(int dtype, void *C, const void *X, const fg_conv_geom *g, int num_gemm, const void *const *A,
                  const long *M, const long *K, const fg_epilogue *epi, const fg_chain_opts *opts, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  ImgGeom ig;
  long K1, N;
  if ((rc = conv_geom(g, &ig, &K1, &N))) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, A, M, K))) return rc;
  {
    const long elems = g->batch * g->channels * g->height * g->width;
    ig.nq = (int)(elems / 4);
    ig.vec4 = dtype == FG_DTYPE_INT8 && ig.sw == 1 && ig.OW % 4 == 0 && elems % 4 == 0 && ((uintptr_t)X & 3) == 0;
  }
  if (K[0] != K1) return set_error(FG_ERR_INVALID, "conv chain: K[0]=%ld but the patches have K1=%ld rows", K[0], K1);
  if (!X || !C) return set_error(FG_ERR_INVALID, "conv chain: X or C is NULL");
  EpiParams p{};
  const EpiParams *ep;
  if ((rc = epi_params(dtype, num_gemm, epi, &p, &ep))) return rc;
  fg_chain_opts o;
  if (opts) o = *opts; else fg_chain_opts_default(&o);
  hipStream_t s = (hipStream_t)stream;
  const int kind = kind_of(dtype, o.f32_mode);
  // The gathering kernel loads single elements, so the fused path takes any N: plan
  // with N rounded up to whole 4-column groups.
  ChainPlan pl;
  if ((rc = plan_chain(dtype, (N + 3) / 4 * 4, num_gemm, M, K, &o, &pl, ep != nullptr, true))) return rc;
  if (whole_fused(pl)) {
    ChainArgs c = chain_args(kind, X, C, N, num_gemm, A, M, K, pl.tile[0]);
    c.img = &ig;
    c.epi = ep;
    return run_fused(c, o.a_packed, s);
  }
  // unfused (options, LDS or the reference layout): form B, then the matrix chain
  void *B;
  if ((rc = scratch(s, "im2col", (size_t)K1 * N * dtype_bytes(dtype), &B))) return rc;
  if ((rc = hip_status(launch_im2col((int)dtype_bytes(dtype), X, B, ig, K1, N, s), "im2col"))) return rc;
  return chain_impl(dtype, C, B, N, num_gemm, A, M, K, ep, &o, stream);
}

// -------------------------------------------------------- host-pointer forms --
// The host forms share one stream and its scratch ("host_*" buffers): one mutex
// serializes them.
static std::mutex g_host_mu;

static hipStream_t host_stream() {
  static hipStream_t s = nullptr;
  static std::once_flag once;
  std::call_once(once, [] { (void)hipStreamCreate(&s); });
  return s;
}

int fg_chain_host(int dtype, void *C, const void *B, long N, int num_gemm, const void *const *A, const long *M,
                  const long *K, const fg_chain_opts *opts) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, A, M, K))) return rc;
  if (!B || !C) return set_error(FG_ERR_INVALID, "B or C is NULL");
  std::lock_guard<std::mutex> lk(g_host_mu);
  hipStream_t s = host_stream();
  size_t e = dtype_bytes(dtype);
  void *dB, *dC, *dA[8];
  if ((rc = scratch(s, "host_B", (size_t)K[0] * N * e, &dB))) return rc;
  if ((rc = scratch(s, "host_C", (size_t)M[num_gemm - 1] * N * 4, &dC))) return rc;
  for (int i = 0; i < num_gemm; ++i) {
    char name[16];
    snprintf(name, sizeof name, "host_A%d", i);
    if ((rc = scratch(s, name, (size_t)M[i] * K[i] * e, &dA[i]))) return rc;
    if ((rc = hip_status(hipMemcpyAsync(dA[i], A[i], (size_t)M[i] * K[i] * e, hipMemcpyHostToDevice, s), "H2D A")))
      return rc;
  }
  if ((rc = hip_status(hipMemcpyAsync(dB, B, (size_t)K[0] * N * e, hipMemcpyHostToDevice, s), "H2D B"))) return rc;
  if ((rc = fg_chain(dtype, dC, dB, N, num_gemm, (const void *const *)dA, M, K, opts, s))) return rc;
  if ((rc = hip_status(hipMemcpyAsync(C, dC, (size_t)M[num_gemm - 1] * N * 4, hipMemcpyDeviceToHost, s), "D2H C")))
    return rc;
  return hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int fg_chain_conv_host(int dtype, void *C, const void *X, const fg_conv_geom *g, int num_gemm, const void *const *A,
                       const long *M, const long *K, const fg_chain_opts *opts) {
  int rc;
  if ((rc = ensure_init())) return rc;
  long K1, N;
  if ((rc = conv_geom(g, nullptr, &K1, &N))) return rc;
  if ((rc = validate_chain(dtype, N, num_gemm, A, M, K))) return rc;
  if (!X || !C) return set_error(FG_ERR_INVALID, "conv chain: X or C is NULL");
  std::lock_guard<std::mutex> lk(g_host_mu);
  hipStream_t s = host_stream();
  const size_t e = dtype_bytes(dtype), xbytes = (size_t)g->batch * g->channels * g->height * g->width * e;
  void *dX, *dC, *dA[8];
  if ((rc = scratch(s, "host_X", xbytes, &dX))) return rc;
  if ((rc = scratch(s, "host_C", (size_t)M[num_gemm - 1] * N * 4, &dC))) return rc;
  for (int i = 0; i < num_gemm; ++i) {
    char name[16];
    snprintf(name, sizeof name, "host_A%d", i);
    if ((rc = scratch(s, name, (size_t)M[i] * K[i] * e, &dA[i]))) return rc;
    if ((rc = hip_status(hipMemcpyAsync(dA[i], A[i], (size_t)M[i] * K[i] * e, hipMemcpyHostToDevice, s), "H2D A")))
      return rc;
  }
  if ((rc = hip_status(hipMemcpyAsync(dX, X, xbytes, hipMemcpyHostToDevice, s), "H2D X"))) return rc;
  if ((rc = fg_chain_conv(dtype, dC, dX, g, num_gemm, (const void *const *)dA, M, K, nullptr, opts, s))) return rc;
  if ((rc = hip_status(hipMemcpyAsync(C, dC, (size_t)M[num_gemm - 1] * N * 4, hipMemcpyDeviceToHost, s), "D2H C")))
    return rc;
  return hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int fg_sgemm_host(float *C, const float *A, const float *B, long M, long N, long K, int f32_mode) {
  const void *As[1] = {A};
  long Ms[1] = {M}, Ks[1] = {K};
  fg_chain_opts o;
  fg_chain_opts_default(&o);
  o.f32_mode = f32_mode;
  o.fuse = FG_FUSE_OFF;  // a single GEMM is the tiled kernel (split-K when skinny)
  return fg_chain_host(FG_DTYPE_F32, C, B, N, 1, As, Ms, Ks, &o);
}

// ------------------------------------------------------------- FP64 oracle --
int fg_f64_chain(double *C, const double *B, long N, int num_gemm, const double *const *A, const long *M,
                 const long *K, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if ((rc = validate_chain(FG_DTYPE_F32, N, num_gemm, (const void *const *)A, M, K))) return rc;
  hipStream_t s = (hipStream_t)stream;
  long maxm = 0;
  for (int i = 0; i + 1 < num_gemm; ++i) maxm = M[i] > maxm ? M[i] : maxm;
  void *buf[2] = {nullptr, nullptr};
  if (num_gemm > 1)
    for (int b = 0; b < 2; ++b)
      if ((rc = scratch(s, b ? "f64_1" : "f64_0", (size_t)maxm * N * 8, &buf[b]))) return rc;
  const double *X = B;
  for (int i = 0; i < num_gemm; ++i) {
    double *Y = (i == num_gemm - 1) ? C : (double *)buf[i & 1];
    if ((rc = hip_status(launch_f64_gemm(A[i], X, Y, M[i], N, K[i], s), "f64 gemm"))) return rc;
    X = Y;
  }
  return FG_OK;
}

int fg_to_f64(int x_kind, const void *x, double *out, long n, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (x_kind < 0 || x_kind > 3 || n <= 0) return set_error(FG_ERR_INVALID, "to_f64 kind=%d n=%ld", x_kind, n);
  return hip_status(launch_to_f64(x_kind, x, out, n, (hipStream_t)stream), "to_f64");
}

int fg_error_norms(int x_kind, const void *x, const double *ref, long n, double *out, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (x_kind < 0 || x_kind > 3 || n <= 0) return set_error(FG_ERR_INVALID, "error_norms kind=%d n=%ld", x_kind, n);
  hipStream_t s = (hipStream_t)stream;
  void *ws;
  if ((rc = scratch(s, "err", sizeof(double) * (2 + 2 * 1024), &ws))) return rc;
  if ((rc = hip_status(launch_err_norms(x_kind, x, ref, n, (double *)ws, s), "error norms"))) return rc;
  if ((rc = hip_status(hipMemcpyAsync(out, ws, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H norms")))
    return rc;
  return hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int fg_inject_alloc_failures(int count) {
  fg::g_inject_alloc.store(count > 0 ? count : 0);
  return FG_OK;
}

int fg_measure_stream(long slice_kb, int rounds, int workgroups, float *ms, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (slice_kb < 1 || slice_kb > 65536 || rounds < 1 || workgroups < 1 || workgroups > 65536 || !ms)
    return set_error(FG_ERR_INVALID, "measure_stream slice_kb=%ld rounds=%d workgroups=%d", slice_kb, rounds,
                     workgroups);
  hipStream_t s = (hipStream_t)stream;
  void *region, *out;
  if ((rc = scratch(s, "stream_region", (size_t)8 * slice_kb * 1024, &region))) return rc;
  if ((rc = scratch(s, "stream_out", (size_t)workgroups * 512 * sizeof(unsigned), &out))) return rc;
  hipEvent_t e0, e1;
  if ((rc = hip_status(hipEventCreate(&e0), "hipEventCreate"))) return rc;
  if ((rc = hip_status(hipEventCreate(&e1), "hipEventCreate"))) {
    (void)hipEventDestroy(e0);
    return rc;
  }
  rc = hip_status(hipEventRecord(e0, s), "hipEventRecord");
  if (!rc) rc = hip_status(launch_stream_probe(region, (unsigned)slice_kb, rounds, workgroups, (unsigned *)out, s),
                           "stream probe");
  if (!rc) rc = hip_status(hipEventRecord(e1, s), "hipEventRecord");
  if (!rc) rc = hip_status(hipEventSynchronize(e1), "hipEventSynchronize");
  if (!rc) rc = hip_status(hipEventElapsedTime(ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// ================================================================ part 1 ===

int flashgemm_thread_num = 1;

void flashgemm_set_thread_num(int num) { flashgemm_thread_num = num; }

static int env_f32_mode() {
  const char *v = getenv("FLASHGEMM_F32_MODE");
  return (v && strcmp(v, "bf16x3") == 0) ? FG_F32_BF16X3 : FG_F32_MFMA;
}

static int env_layout() {
  const char *v = getenv("FLASHGEMM_LAYOUT");
  return (v && strcmp(v, "math") == 0) ? FG_LAYOUT_MATH : FG_LAYOUT_REFERENCE;
}

// Reads the variadic triples the way multigemm/src/flashgemm.c:54-68 does, with the
// same K % 2 check and messages; returns false when the call must stop.
static bool read_triples(va_list ap, int num_gemm, std::vector<const void *> &A, std::vector<long> &M,
                         std::vector<long> &K) {
  if (num_gemm < 1) return false;
  A.resize(num_gemm);
  M.resize(num_gemm);
  K.resize(num_gemm);
  for (int i = 0; i < num_gemm; ++i) {
    A[i] = va_arg(ap, const void *);
    M[i] = va_arg(ap, long);
    K[i] = va_arg(ap, long);
    if (K[i] % 2 != 0) {
      printf("K must be a multiple of 2\n");
      printf("Dimension error for gemm %d\n", i + 1);
      fflush(stdout);
      return false;
    }
  }
  return true;
}

static void legacy_chain(int dtype, void *C, const void *B, long N, int num_gemm, std::vector<const void *> &A,
                         std::vector<long> &M, std::vector<long> &K) {
  fg_chain_opts o;
  fg_chain_opts_default(&o);
  o.f32_mode = env_f32_mode();
  o.layout = (dtype == FG_DTYPE_F32) ? FG_LAYOUT_MATH : env_layout();
  o.ref_threads = flashgemm_thread_num > 0 ? flashgemm_thread_num : 1;
  int rc = fg_chain_host(dtype, C, B, N, num_gemm, A.data(), M.data(), K.data(), &o);
  if (rc == FG_ERR_DOMAIN) {
    fprintf(stderr, "flashgemm: %s; computing the mathematical chain instead\n", fg_last_error());
    o.layout = FG_LAYOUT_MATH;
    rc = fg_chain_host(dtype, C, B, N, num_gemm, A.data(), M.data(), K.data(), &o);
  }
  if (rc != FG_OK) fprintf(stderr, "flashgemm: %s (%s)\n", fg_status_string(rc), fg_last_error());
}

void flashgemm_multi_f32f32f32(float *C, float *B, long N, int num_gemm, ...) {
  std::vector<const void *> A;
  std::vector<long> M, K;
  va_list ap;
  va_start(ap, num_gemm);
  bool ok = read_triples(ap, num_gemm, A, M, K);
  va_end(ap);
  if (ok) legacy_chain(FG_DTYPE_F32, C, B, N, num_gemm, A, M, K);
}

void flashgemm_multi_bf16bf16f32(float *C, uint16_t *B, long N, int num_gemm, ...) {
  std::vector<const void *> A;
  std::vector<long> M, K;
  va_list ap;
  va_start(ap, num_gemm);
  bool ok = read_triples(ap, num_gemm, A, M, K);
  va_end(ap);
  if (ok) legacy_chain(FG_DTYPE_BF16, C, B, N, num_gemm, A, M, K);
}

void flashgemm_multi_int8uint8int32(int *C, uint8_t *B, long N, int num_gemm, ...) {
  std::vector<const void *> A;
  std::vector<long> M, K;
  va_list ap;
  va_start(ap, num_gemm);
  bool ok = read_triples(ap, num_gemm, A, M, K);
  va_end(ap);
  if (ok) legacy_chain(FG_DTYPE_INT8, C, B, N, num_gemm, A, M, K);
}

void LibShalom_set_thread_nums(int num) { (void)num; }

void flashgemm_sgemm(int transa, int transb, float *C, float *A, float *B, long M, long N, long K) {
  (void)transa;
  if (transb == 1) return;  // NT: the reference computes nothing (FP32_ID23.c:39-43)
  int rc = fg_sgemm_host(C, A, B, M, N, K, env_f32_mode());
  if (rc != FG_OK) fprintf(stderr, "flashgemm: %s (%s)\n", fg_status_string(rc), fg_last_error());
}

void LibShalom_sgemm_mp(int transa, int transb, float *C, float *A, float *B, long M, long N, long K) {
  flashgemm_sgemm(transa, transb, C, A, B, M, N, K);
}

}  // extern "C"
