// Autotuner: picks, per (kind, shape bucket), how a chain runs: the tile width of a
// fused launch and the fusion depth (which GEMMs share a fused segment and which
// intermediates spill to HBM between segments), by timing the candidates on the
// device. The reference chooses its blocking with a static cost model
// (singlegemm/src/set_block.h:8-17,40-200: per-micro-kernel costs, thread grids from
// set_TmTn.h:10-58) and always runs the whole chain per panel
// (multigemm/src/flashgemm.c:123-148); here the choice is measured.
//
// Keys are shape buckets: the chain length and ceil(log2) of N and of every M_i and
// K_i, so one measurement serves every chain of about the same shape (the serving
// workload's 64 weight sets share a few dozen buckets). A bucket's plan is used for a
// shape only if it can run that exact shape (plan_valid); otherwise the model decides.
// Decisions live in a table the planner (fg_chain) and the grouped launch consult,
// persisted as a versioned text file keyed by arch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/flashgemm.h"
#include "fg_internal.h"
#include "fg_runtime.h"

using namespace fg;

namespace {

struct Key {
  int kind, num_gemm, nb;
  int mb[CHAIN_MAXG], kb[CHAIN_MAXG];
  bool operator<(const Key &o) const { return memcmp(this, &o, sizeof(Key)) < 0; }
};

struct Entry {
  ChainPlan plan;
  double ms;
};

std::mutex g_mu;
std::map<Key, Entry> g_table;

int bucket(long x) { return x <= 1 ? 0 : (int)std::ceil(std::log2((double)x)); }

Key make_key(int kind, long N, int n, const long *M, const long *K) {
  Key k;
  memset(&k, 0, sizeof k);
  k.kind = kind;
  k.num_gemm = n;
  k.nb = bucket(N);
  for (int i = 0; i < n && i < CHAIN_MAXG; ++i) {
    k.mb[i] = bucket(M[i]);
    k.kb[i] = bucket(K[i]);
  }
  return k;
}

// Structural check of a stored plan (the shape-specific check is plan_valid).
bool plan_well_formed(int kind, int n, const ChainPlan &p) {
  if (p.nseg < 1 || p.nseg > n || p.splits < 0) return false;
  int start = 0;
  for (int s = 0; s < p.nseg; ++s) {
    if (p.end[s] < start || p.end[s] >= n) return false;
    if (p.tile[s] == 0 ? p.end[s] != start : !chain_fused_tile_ok(kind, p.tile[s])) return false;
    start = p.end[s] + 1;
  }
  return start == n;
}

bool same_plan(const ChainPlan &a, const ChainPlan &b) {
  if (a.nseg != b.nseg || a.splits != b.splits) return false;
  for (int s = 0; s < a.nseg; ++s)
    if (a.end[s] != b.end[s] || a.tile[s] != b.tile[s]) return false;
  return true;
}

// Candidate plans of a chain: the whole chain fused at every tile width that fits;
// every cut into segments (fused runs at the planner's tile width where they fit,
// single GEMMs unfused); every GEMM unfused with the planner's split-K and with none.
std::vector<ChainPlan> candidates(int kind, long N, int n, const long *M, const long *K) {
  std::vector<ChainPlan> out;
  auto add = [&](const ChainPlan &p) {
    if (!plan_valid(kind, N, n, M, K, p)) return;
    for (const ChainPlan &q : out)
      if (same_plan(p, q)) return;
    out.push_back(p);
  };
  for (int tn : {32, 64, 96, 128, 256}) {
    ChainPlan p{};
    p.nseg = 1;
    p.end[0] = n - 1;
    p.tile[0] = tn;
    add(p);
  }
  std::vector<ChainPlan> cuts;
  for (unsigned mask = 0; mask < (1u << (n - 1)); ++mask) {  // bit i: a boundary after GEMM i
    ChainPlan p{};
    int start = 0;
    bool ok = true;
    for (int g = 0; g < n && ok; ++g) {
      if (g < n - 1 && !(mask >> g & 1u)) continue;
      const int len = g - start + 1;
      int tn = 0;
      if (len > 1) {
        tn = chain_fused_shape_ok(kind, N, len, K + start) ? plan_tile_n(kind, N, len, M + start, K + start) : 0;
        ok = tn > 0;
      }
      p.end[p.nseg] = g;
      p.tile[p.nseg] = tn;
      ++p.nseg;
      start = g + 1;
    }
    if (ok) cuts.push_back(p);
  }
  std::stable_sort(cuts.begin(), cuts.end(), [](const ChainPlan &a, const ChainPlan &b) { return a.nseg < b.nseg; });
  if (cuts.size() > 48) cuts.resize(48);
  for (const ChainPlan &p : cuts) add(p);
  for (int sp : {0, 1}) {
    ChainPlan p{};
    p.nseg = n;
    for (int g = 0; g < n; ++g) {
      p.end[g] = g;
      p.tile[g] = 0;
    }
    p.splits = sp;
    add(p);
  }
  return out;
}

}  // namespace

namespace fg {
bool autotune_lookup_plan(int kind, long N, int n, const long *M, const long *K, ChainPlan *p) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(make_key(kind, N, n, M, K));
  if (it == g_table.end()) return false;
  *p = it->second.plan;
  return true;
}
}  // namespace fg

extern "C" {

int fg_autotune_chain(int dtype, int f32_mode, long N, int num_gemm, const long *M, const long *K, void *stream,
                      int *fused_out, int *tile_n_out, int *splits_out, double *ms_out) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (dtype < 0 || dtype > 2 || N <= 0 || num_gemm < 1 || num_gemm > CHAIN_MAXG || !M || !K)
    return set_error(FG_ERR_INVALID, "autotune arguments");
  for (int i = 0; i < num_gemm; ++i)
    if (M[i] <= 0 || K[i] <= 0 || (i > 0 && K[i] != M[i - 1])) return set_error(FG_ERR_INVALID, "autotune shapes");
  hipStream_t s = (hipStream_t)stream;
  const int esz = dtype == FG_DTYPE_F32 ? 4 : dtype == FG_DTYPE_BF16 ? 2 : 1;
  if ((rc = check_f32_mode(f32_mode))) return rc;
  const int kind = kind_of(dtype, f32_mode);
  if (kind == K_F32X6) return set_error(FG_ERR_UNSUPPORTED, "the three-piece split has one plan (unfused)");
  const std::vector<ChainPlan> cands = candidates(kind, N, num_gemm, M, K);
  if (cands.empty()) return set_error(FG_ERR_UNSUPPORTED, "no runnable plan");
  // operands: seeded random data of the chain's dtype (timing only)
  std::vector<void *> bufs;
  auto alloc = [&](size_t elems, size_t bytes_per, void **p, unsigned long long seed) -> int {
    hipError_t er = dev_alloc(p, elems * bytes_per);
    if (er != hipSuccess) return hip_status(er, "hipMalloc(autotune)");
    bufs.push_back(*p);
    return seed ? hip_status(launch_fill_random(esz, *p, (long)elems, seed, s), "fill") : FG_OK;
  };
  void *B = nullptr, *C = nullptr, *A[CHAIN_MAXG] = {};
  rc = alloc((size_t)K[0] * N, esz, &B, 1);
  if (!rc) rc = alloc((size_t)M[num_gemm - 1] * N, 4, &C, 0);
  for (int i = 0; i < num_gemm && !rc; ++i) rc = alloc((size_t)M[i] * K[i], esz, &A[i], 2 + i);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = hip_status(hipEventCreate(&e0), "hipEventCreate");
  if (!rc) rc = hip_status(hipEventCreate(&e1), "hipEventCreate");

  Entry best{};
  best.ms = 1e30;
  for (const ChainPlan &p : cands) {
    if (rc) break;
    auto run = [&]() { return run_plan(kind, dtype, p, C, B, N, num_gemm, (const void *const *)A, M, K, 0, s); };
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
    if (r2 == FG_OK && ms < best.ms) best = Entry{p, ms};
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
  const bool whole = best.plan.nseg == 1 && best.plan.tile[0] > 0;
  if (fused_out) *fused_out = whole;
  if (tile_n_out) *tile_n_out = whole ? best.plan.tile[0] : 0;
  if (splits_out) *splits_out = best.plan.splits;
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
//   flashgemm-autotune v2 <arch>
//   <kind> <num_gemm> <N bucket> <M_1..M_n buckets> <K_1..K_n buckets>
//       <segments> <end_1> <tile_1> ... <end_s> <tile_s> <splits> <ms>
// (buckets are ceil(log2); a segment's tile 0 is one unfused GEMM).
int fg_autotune_save(const char *path) {
  FILE *f = fopen(path, "w");
  if (!f) return set_error(FG_ERR_INVALID, "cannot write %s", path);
  fprintf(f, "flashgemm-autotune v2 gfx950\n");
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_table) {
    const Key &k = kv.first;
    const ChainPlan &p = kv.second.plan;
    fprintf(f, "%d %d %d", k.kind, k.num_gemm, k.nb);
    for (int i = 0; i < k.num_gemm; ++i) fprintf(f, " %d", k.mb[i]);
    for (int i = 0; i < k.num_gemm; ++i) fprintf(f, " %d", k.kb[i]);
    fprintf(f, " %d", p.nseg);
    for (int s = 0; s < p.nseg; ++s) fprintf(f, " %d %d", p.end[s], p.tile[s]);
    fprintf(f, " %d %.6f\n", p.splits, kv.second.ms);
  }
  fclose(f);
  return FG_OK;
}

// Every entry is checked (kind, chain length, buckets, segments in order, a tile
// width the fused kernels have for the kind, single GEMMs for unfused segments); a
// file with any bad entry is refused whole and the table is left as it was.
int fg_autotune_load(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return set_error(FG_ERR_INVALID, "cannot read %s", path);
  char magic[64] = {0}, ver[16] = {0}, arch[32] = {0};
  if (fscanf(f, "%63s %15s %31s", magic, ver, arch) != 3 || strcmp(magic, "flashgemm-autotune") != 0 ||
      strcmp(ver, "v2") != 0 || strcmp(arch, "gfx950") != 0) {
    fclose(f);
    return set_error(FG_ERR_INVALID, "%s: not a v2 gfx950 autotune table", path);
  }
  std::map<Key, Entry> loaded;
  int rc = FG_OK, line = 1;
  while (true) {
    Key k;
    memset(&k, 0, sizeof k);
    if (fscanf(f, "%d %d %d", &k.kind, &k.num_gemm, &k.nb) != 3) break;
    ++line;
    bool ok = k.kind >= 0 && k.kind <= 3 && k.num_gemm >= 1 && k.num_gemm <= CHAIN_MAXG && k.nb >= 0 && k.nb < 63;
    for (int i = 0; i < k.num_gemm && ok; ++i) ok = fscanf(f, "%d", &k.mb[i]) == 1 && k.mb[i] >= 0 && k.mb[i] < 63;
    for (int i = 0; i < k.num_gemm && ok; ++i) ok = fscanf(f, "%d", &k.kb[i]) == 1 && k.kb[i] >= 0 && k.kb[i] < 63;
    Entry e{};
    ok = ok && fscanf(f, "%d", &e.plan.nseg) == 1 && e.plan.nseg >= 1 && e.plan.nseg <= k.num_gemm;
    for (int s = 0; s < e.plan.nseg && ok; ++s) ok = fscanf(f, "%d %d", &e.plan.end[s], &e.plan.tile[s]) == 2;
    ok = ok && fscanf(f, "%d %lf", &e.plan.splits, &e.ms) == 2 && plan_well_formed(k.kind, k.num_gemm, e.plan);
    if (!ok) {
      rc = set_error(FG_ERR_INVALID, "%s: bad entry on line %d", path, line);
      break;
    }
    loaded[k] = e;
  }
  fclose(f);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : loaded) g_table[kv.first] = kv.second;
  return FG_OK;
}

}  // extern "C"
