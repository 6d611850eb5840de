// Autotuner: picks, per (kind, chain shape, N bucket), whether a chain runs fused on
// chip and at which tile width, or unfused with which split-K, by timing the
// candidates on the device. The reference chooses its blocking with a static cost
// model (singlegemm/src/set_block.h:8-17,40-200: per-micro-kernel costs, thread
// grids from set_TmTn.h:10-58); here the choice is measured. Decisions live in a
// table the planner consults, persisted as a versioned text file keyed by arch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flashgemm.h"
#include "fg_internal.h"
#include "fg_runtime.h"

using namespace fg;

namespace {

struct Key {
  int kind, num_gemm, nbucket;
  long M[CHAIN_MAXG], K[CHAIN_MAXG];
  bool operator<(const Key &o) const { return memcmp(this, &o, sizeof(Key)) < 0; }
};

struct Choice {
  int fused, tile_n, splits;
  double ms;
};

std::mutex g_mu;
std::map<Key, Choice> g_table;

int bucket(long N) { return N <= 1 ? 0 : (int)std::floor(std::log2((double)N)); }

Key make_key(int kind, long N, int n, const long *M, const long *K) {
  Key k;
  memset(&k, 0, sizeof k);
  k.kind = kind;
  k.num_gemm = n;
  k.nbucket = bucket(N);
  for (int i = 0; i < n && i < CHAIN_MAXG; ++i) {
    k.M[i] = M[i];
    k.K[i] = K[i];
  }
  return k;
}

}  // namespace

namespace fg {
// Planner hook: true and the decision when the table has this shape.
bool autotune_lookup(int kind, long N, int n, const long *M, const long *K, int *fused, int *tile_n, int *splits) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(make_key(kind, N, n, M, K));
  if (it == g_table.end()) return false;
  *fused = it->second.fused;
  *tile_n = it->second.tile_n;
  *splits = it->second.splits;
  return true;
}
}  // namespace fg

extern "C" {

int fg_autotune_chain(int dtype, int f32_mode, long N, int num_gemm, const long *M, const long *K, void *stream,
                      int *fused_out, int *tile_n_out, int *splits_out, double *ms_out) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (dtype < 0 || dtype > 2 || N <= 0 || num_gemm < 1 || num_gemm > CHAIN_MAXG)
    return set_error(FG_ERR_INVALID, "autotune arguments");
  for (int i = 0; i < num_gemm; ++i)
    if (M[i] <= 0 || K[i] <= 0 || (i > 0 && K[i] != M[i - 1])) return set_error(FG_ERR_INVALID, "autotune shapes");
  hipStream_t s = (hipStream_t)stream;
  const size_t e = dtype == FG_DTYPE_F32 ? 4 : dtype == FG_DTYPE_BF16 ? 2 : 1;
  // operands: zero-filled device buffers (timing only; zeros keep every path finite)
  std::vector<void *> bufs;
  auto alloc = [&](size_t bytes, void **p) -> int {
    hipError_t er = hipMalloc(p, bytes);
    if (er != hipSuccess) return hip_status(er, "hipMalloc(autotune)");
    bufs.push_back(*p);
    return hip_status(hipMemsetAsync(*p, 0, bytes, s), "hipMemsetAsync");
  };
  void *B, *C, *A[CHAIN_MAXG];
  rc = alloc((size_t)K[0] * N * e, &B);
  if (!rc) rc = alloc((size_t)M[num_gemm - 1] * N * 4, &C);
  for (int i = 0; i < num_gemm && !rc; ++i) rc = alloc((size_t)M[i] * K[i] * e, &A[i]);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = hip_status(hipEventCreate(&e0), "hipEventCreate");
  if (!rc) rc = hip_status(hipEventCreate(&e1), "hipEventCreate");

  int kind = kind_of(dtype, f32_mode);
  Choice best{0, 0, 0, 1e30};
  struct Cand { int fuse, tile_n, splits; };
  std::vector<Cand> cands;
  for (int tn : {32, 64, 96, 128, 256})
    if (chain_fused_tile_ok(kind, tn) && chain_fused_lds_bytes(kind, M, K, num_gemm, tn) <= LDS_BYTES_MAX)
      cands.push_back({FG_FUSE_ON, tn, 0});
  cands.push_back({FG_FUSE_OFF, 0, 1});
  cands.push_back({FG_FUSE_OFF, 0, 0});  // planner's split-K
  for (const Cand &c : cands) {
    if (rc) break;
    fg_chain_opts o;
    fg_chain_opts_default(&o);
    o.f32_mode = f32_mode;
    o.fuse = c.fuse;
    o.tile_n = c.tile_n;
    o.splits = c.splits;
    auto run = [&]() { return fg_chain(dtype, C, B, N, num_gemm, (const void *const *)A, M, K, &o, stream); };
    int r2 = FG_OK;
    for (int w = 0; w < 2 && r2 == FG_OK; ++w) r2 = run();
    if (r2 != FG_OK) continue;  // candidate not runnable: skip it
    const int reps = 5;
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < reps && r2 == FG_OK; ++i) r2 = run();
    (void)hipEventRecord(e1, s);
    if ((rc = hip_status(hipEventSynchronize(e1), "hipEventSynchronize"))) break;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (r2 == FG_OK && ms < best.ms) best = {c.fuse == FG_FUSE_ON, c.tile_n, c.splits, ms};
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  for (void *p : bufs) (void)hipFree(p);
  if (rc) return rc;
  if (best.ms >= 1e30) return set_error(FG_ERR_UNSUPPORTED, "no runnable candidate");
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[make_key(kind, N, num_gemm, M, K)] = best;
  }
  if (fused_out) *fused_out = best.fused;
  if (tile_n_out) *tile_n_out = best.tile_n;
  if (splits_out) *splits_out = best.splits;
  if (ms_out) *ms_out = best.ms;
  return FG_OK;
}

int fg_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.clear();
  return FG_OK;
}

int fg_autotune_size(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return (int)g_table.size();
}

// Text format, one decision per line after a header:
//   flashgemm-autotune v1 <arch>
//   <kind> <num_gemm> <nbucket> <M_1..M_n> <K_1..K_n> <fused> <tile_n> <splits> <ms>
int fg_autotune_save(const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) return set_error(FG_ERR_INVALID, "cannot write %s", path);
  fprintf(f, "flashgemm-autotune v1 gfx950\n");
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_table) {
    const Key &k = kv.first;
    fprintf(f, "%d %d %d", k.kind, k.num_gemm, k.nbucket);
    for (int i = 0; i < k.num_gemm; ++i) fprintf(f, " %ld", k.M[i]);
    for (int i = 0; i < k.num_gemm; ++i) fprintf(f, " %ld", k.K[i]);
    fprintf(f, " %d %d %d %.6f\n", kv.second.fused, kv.second.tile_n, kv.second.splits, kv.second.ms);
  }
  fclose(f);
  return FG_OK;
}

int fg_autotune_load(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(FG_ERR_INVALID, "cannot read %s", path);
  char magic[64] = {0}, ver[16] = {0}, arch[32] = {0};
  if (fscanf(f, "%63s %15s %31s", magic, ver, arch) != 3 || strcmp(magic, "flashgemm-autotune") != 0 ||
      strcmp(ver, "v1") != 0 || strcmp(arch, "gfx950") != 0) {
    fclose(f);
    return set_error(FG_ERR_INVALID, "%s: not a v1 gfx950 autotune table", path);
  }
  std::map<Key, Choice> loaded;
  int rc = FG_OK;
  while (true) {
    Key k;
    memset(&k, 0, sizeof k);
    if (fscanf(f, "%d %d %d", &k.kind, &k.num_gemm, &k.nbucket) != 3) break;
    if (k.kind < 0 || k.kind > 3 || k.num_gemm < 1 || k.num_gemm > CHAIN_MAXG) {
      rc = set_error(FG_ERR_INVALID, "%s: bad entry", path);
      break;
    }
    bool ok = true;
    for (int i = 0; i < k.num_gemm && ok; ++i) ok = fscanf(f, "%ld", &k.M[i]) == 1;
    for (int i = 0; i < k.num_gemm && ok; ++i) ok = fscanf(f, "%ld", &k.K[i]) == 1;
    Choice c;
    ok = ok && fscanf(f, "%d %d %d %lf", &c.fused, &c.tile_n, &c.splits, &c.ms) == 4;
    // a fused entry names a tile width the fused kernels have for this kind
    if (!ok || (c.fused && !chain_fused_tile_ok(k.kind, c.tile_n)) || (!c.fused && c.splits < 0)) {
      rc = set_error(FG_ERR_INVALID, "%s: bad entry", path);
      break;
    }
    loaded[k] = c;
  }
  fclose(f);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : loaded) g_table[kv.first] = kv.second;
  return FG_OK;
}

}  // extern "C"
