// Grouped launches: many chains of uneven shapes and mixed dtypes in one persistent
// kernel, directly or replayed from a captured hipGraph.
//
// The reference runs one chain per call with an OpenMP team (multigemm/src/
// flashgemm.c:79-148) and has no batching; this is the GPU counterpart of running
// many such calls at once. Each request's columns are cut into tiles of the
// grouped tile width of its kind; the tile list is sorted by decreasing work so the
// static round-robin over workgroups balances, and every tile runs the whole chain
// on chip. Requests whose intermediates do not fit LDS at that width are not part
// of the grouped launch (fg_grouped_chain runs them unfused after it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/flashgemm.h"
#include "fg_internal.h"
#include "fg_runtime.h"

using namespace fg;

namespace {

struct Packed {
  std::vector<char> bytes;  // GroupTable + ChainParams[] + int2[]
  size_t lds = 0;
  int items = 0;
  bool wide = false;  // some request runs 64-column tiles
};

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// Validate a request and build its descriptor; fits = intermediates fit on chip.
int describe(const fg_chain_req &r, ChainParams *P, int *tile_n, size_t *lds, bool *fits) {
  if (r.dtype < FG_DTYPE_F32 || r.dtype > FG_DTYPE_INT8) return set_error(FG_ERR_INVALID, "dtype %d", r.dtype);
  if (r.num_gemm < 1 || r.num_gemm > CHAIN_MAXG || r.N <= 0 || !r.B || !r.C)
    return set_error(FG_ERR_INVALID, "request: num_gemm=%d N=%ld", r.num_gemm, r.N);
  for (int i = 0; i < r.num_gemm; ++i) {
    if (!r.A[i] || r.M[i] <= 0 || r.K[i] <= 0) return set_error(FG_ERR_INVALID, "request gemm %d", i + 1);
    if (i > 0 && r.K[i] != r.M[i - 1]) return set_error(FG_ERR_INVALID, "request gemm %d: K != previous M", i + 1);
  }
  if (int rc = check_f32_mode(r.f32_mode)) return rc;
  int kind = kind_of(r.dtype, r.f32_mode);
  if (kind == K_F32X6) return set_error(FG_ERR_UNSUPPORTED, "the three-piece split runs through fg_chain only");
  ChainArgs c{};
  c.kind = kind;
  c.num_gemm = r.num_gemm;
  for (int i = 0; i < r.num_gemm; ++i) {
    c.A[i] = r.A[i];
    c.M[i] = r.M[i];
    c.K[i] = r.K[i];
  }
  c.B = r.B;
  c.C = r.C;
  c.N = r.N;
  // Tile width: 32 columns, unless the autotune table has a plan for the request's
  // shape bucket that runs this exact shape: the whole chain fused at 32 or 64
  // columns sets the width; any other plan (wider tiles, segments that spill, unfused
  // GEMMs) keeps the request out of the grouped launch, so it runs with that plan.
  *tile_n = grouped_tile_n(kind);
  bool tuned_out = false;
  ChainPlan tp;
  if (autotune_lookup_plan(kind, r.N, r.num_gemm, r.M, r.K, &tp) && plan_valid(kind, r.N, r.num_gemm, r.M, r.K, tp)) {
    if (tp.nseg == 1 && (tp.tile[0] == 32 || tp.tile[0] == 64)) *tile_n = tp.tile[0];
    else tuned_out = true;
  }
  c.tile_n = *tile_n;
  *lds = chain_fused_lds_bytes(kind, r.M, r.K, r.num_gemm, *tile_n);
  make_chain_params(c, P);
  P->gtnb = *tile_n / 32;
  *fits = !tuned_out && *lds <= LDS_BYTES_MAX && chain_fused_shape_ok(kind, r.N, r.num_gemm, r.K) &&
          chain_fused_aligned(*P);
  return FG_OK;
}

// Weights of requests that are not pre-packed are packed on stream s into its
// "gpack" scratch before the grouped kernel runs: each distinct (kind, A, M, K) once,
// one launch per kind and PACK_MAX matrices. P[j].A then point at the packed copies.
int pack_weights(const fg_chain_req *reqs, const std::vector<int> &src, std::vector<ChainParams> &P,
                 hipStream_t s) {
  struct W {
    int kind;
    const void *A;
    long M, K;
    size_t off;
  };
  std::vector<W> ws;
  std::vector<std::vector<int>> slot(P.size());  // P[j].A[i] -> ws index
  size_t total = 0;
  for (size_t j = 0; j < P.size(); ++j) {
    if (reqs[src[j]].a_packed) continue;
    slot[j].assign(P[j].num_gemm, -1);
    for (int i = 0; i < P[j].num_gemm; ++i) {
      int found = -1;
      for (size_t w = 0; w < ws.size() && found < 0; ++w)
        if (ws[w].kind == P[j].kind && ws[w].A == P[j].A[i] && ws[w].M == P[j].M[i] && ws[w].K == P[j].K[i])
          found = (int)w;
      if (found < 0) {
        found = (int)ws.size();
        ws.push_back({P[j].kind, P[j].A[i], P[j].M[i], P[j].K[i], total});
        total += (packed_a_bytes(P[j].kind, P[j].M[i], P[j].K[i]) + 255) & ~(size_t)255;
      }
      slot[j][i] = found;
    }
  }
  if (ws.empty()) return FG_OK;
  char *buf = nullptr;
  int rc = scratch(s, "gpack", total, (void **)&buf);
  if (rc) return rc;
  for (int kind = 0; kind < 4; ++kind) {
    PackJob J{};
    for (size_t w = 0; w <= ws.size(); ++w) {
      if (J.n == PACK_MAX || (w == ws.size() && J.n > 0)) {
        if ((rc = hip_status(launch_pack_a(kind, J, s), "pack weights"))) return rc;
        J.n = 0;
      }
      if (w == ws.size() || ws[w].kind != kind) continue;
      J.A[J.n] = ws[w].A;
      J.M[J.n] = ws[w].M;
      J.K[J.n] = ws[w].K;
      J.out[J.n] = buf + ws[w].off;
      ++J.n;
    }
  }
  for (size_t j = 0; j < P.size(); ++j)
    for (size_t i = 0; i < slot[j].size(); ++i) P[j].A[i] = buf + ws[slot[j][i]].off;
  return FG_OK;
}

// Pack the requests that fit into a table (packing their weights on stream s when
// needed); `rest` receives the indices of the others.
int pack(int count, const fg_chain_req *reqs, hipStream_t s, Packed *out, std::vector<int> *rest) {
  std::vector<ChainParams> P;
  std::vector<int> src;  // request index of P[j]
  struct Item {
    int req, tile;
    double cost;
  };
  std::vector<Item> items;
  size_t lds = 0;
  for (int r = 0; r < count; ++r) {
    ChainParams p;
    int tn;
    size_t l;
    bool fits;
    int rc = describe(reqs[r], &p, &tn, &l, &fits);
    if (rc) return rc;
    if (!fits) {
      if (rest) rest->push_back(r);
      continue;
    }
    lds = std::max(lds, l);
    double per_col = 0;
    for (int i = 0; i < p.num_gemm; ++i) per_col += (double)p.M[i] * p.K[i];
    int idx = (int)P.size();
    P.push_back(p);
    src.push_back(r);
    long tiles = (reqs[r].N + tn - 1) / tn;
    for (long t = 0; t < tiles; ++t) {
      long w = std::min<long>(tn, reqs[r].N - t * tn);
      items.push_back({idx, (int)t, per_col * (double)std::max<long>(w, 32)});
    }
  }
  std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
  if (int rc = pack_weights(reqs, src, P, s)) return rc;
  GroupTable h{};
  h.num_reqs = (int)P.size();
  h.num_items = (int)items.size();
  h.reqs_off = (int)align16(sizeof(GroupTable));
  h.items_off = (int)align16(h.reqs_off + P.size() * sizeof(ChainParams));
  out->bytes.assign(h.items_off + items.size() * sizeof(int2), 0);
  memcpy(out->bytes.data(), &h, sizeof h);
  if (!P.empty()) memcpy(out->bytes.data() + h.reqs_off, P.data(), P.size() * sizeof(ChainParams));
  int2 *dst = (int2 *)(out->bytes.data() + h.items_off);
  for (size_t i = 0; i < items.size(); ++i) dst[i] = make_int2(items[i].req, items[i].tile);
  out->lds = lds;
  out->items = (int)items.size();
  out->wide = std::any_of(P.begin(), P.end(), [](const ChainParams &p) { return p.gtnb == 2; });
  return FG_OK;
}

// Pinned staging + device table, reused once the event recorded after its last use
// has completed.
struct Slot {
  void *host = nullptr;
  void *dev = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
};

int slot_reserve(Slot &s, size_t bytes) {
  if (!s.done) {
    hipError_t e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_status(e, "hipEventCreate");
  }
  hipError_t e = hipEventSynchronize(s.done);  // previous use finished
  if (e != hipSuccess) return hip_status(e, "hipEventSynchronize");
  if (s.cap >= bytes) return FG_OK;
  if (s.host) (void)hipHostFree(s.host);
  if (s.dev) (void)hipFree(s.dev);
  s.host = s.dev = nullptr;
  s.cap = 0;
  size_t cap = std::max<size_t>(bytes, 64 * 1024);
  if ((e = hipHostMalloc(&s.host, cap, hipHostMallocDefault)) != hipSuccess) return hip_status(e, "hipHostMalloc");
  if ((e = dev_alloc(&s.dev, cap)) != hipSuccess) {
    s.dev = nullptr;
    return hip_status(e, "hipMalloc(group table)");
  }
  s.cap = cap;
  return FG_OK;
}

constexpr int RING = 4;
struct StreamRing {
  Slot slot[RING];
  int next = 0;
};
std::mutex g_ring_mu;
std::vector<std::pair<hipStream_t, StreamRing *>> g_rings;

StreamRing *ring_for(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_ring_mu);
  for (auto &p : g_rings)
    if (p.first == s) return p.second;
  g_rings.push_back({s, new StreamRing});
  return g_rings.back().second;
}

int grouped_grid(int items) { return std::max(1, std::min(items, num_cus())); }

}  // namespace

namespace fg {
void release_stream_ring(hipStream_t s) {
  StreamRing *R = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ring_mu);
    for (size_t i = 0; i < g_rings.size(); ++i)
      if (g_rings[i].first == s) {
        R = g_rings[i].second;
        g_rings.erase(g_rings.begin() + i);
        break;
      }
  }
  if (!R) return;
  for (Slot &sl : R->slot) {
    if (sl.done) (void)hipEventSynchronize(sl.done);
    if (sl.host) (void)hipHostFree(sl.host);
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  delete R;
}
}  // namespace fg

struct fg_grouped_graph {
  struct G {
    Slot slot;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t copy = nullptr;
  } g[RING];
  int next = 0;
  size_t table_cap = 0;
};

extern "C" {

int fg_grouped_chain(int count, const fg_chain_req *reqs, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (count < 0 || (count > 0 && !reqs)) return set_error(FG_ERR_INVALID, "count=%d", count);
  if (count == 0) return FG_OK;
  hipStream_t s = (hipStream_t)stream;
  Packed pk;
  std::vector<int> rest;
  if ((rc = pack(count, reqs, s, &pk, &rest))) return rc;
  if (pk.items > 0) {
    StreamRing *R = ring_for(s);
    Slot &sl = R->slot[R->next];
    R->next = (R->next + 1) % RING;
    if ((rc = slot_reserve(sl, pk.bytes.size()))) return rc;
    memcpy(sl.host, pk.bytes.data(), pk.bytes.size());
    if ((rc = hip_status(hipMemcpyAsync(sl.dev, sl.host, pk.bytes.size(), hipMemcpyHostToDevice, s), "H2D table")))
      return rc;
    if ((rc = hip_status(launch_grouped(sl.dev, grouped_grid(pk.items), pk.lds, pk.wide, s), "grouped launch"))) return rc;
    if ((rc = hip_status(hipEventRecord(sl.done, s), "hipEventRecord"))) return rc;
  }
  for (int r : rest) {
    const fg_chain_req &q = reqs[r];
    fg_chain_opts o;
    fg_chain_opts_default(&o);
    o.f32_mode = q.f32_mode;
    o.a_packed = q.a_packed;
    if ((rc = fg_chain(q.dtype, q.C, q.B, q.N, q.num_gemm, q.A, q.M, q.K, &o, stream))) return rc;
  }
  return FG_OK;
}

int fg_grouped_levels(int count, const fg_chain_req *reqs, void *stream) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (count < 0 || (count > 0 && !reqs)) return set_error(FG_ERR_INVALID, "count=%d", count);
  if (count == 0) return FG_OK;
  hipStream_t s = (hipStream_t)stream;
  // validate; two ping-pong intermediates per request, in the next operand's type
  std::vector<size_t> off(count), slot(count);
  size_t total = 0;
  int levels = 0;
  for (int r = 0; r < count; ++r) {
    const fg_chain_req &q = reqs[r];
    if (q.dtype < FG_DTYPE_F32 || q.dtype > FG_DTYPE_INT8) return set_error(FG_ERR_INVALID, "request %d: dtype", r);
    if (q.num_gemm < 1 || q.num_gemm > CHAIN_MAXG || q.N <= 0 || !q.B || !q.C)
      return set_error(FG_ERR_INVALID, "request %d: num_gemm=%d N=%ld", r, q.num_gemm, q.N);
    if (q.a_packed) return set_error(FG_ERR_UNSUPPORTED, "request %d: packed weights need the fused path", r);
    if ((rc = check_f32_mode(q.f32_mode))) return rc;
    if (kind_of(q.dtype, q.f32_mode) == K_F32X6)
      return set_error(FG_ERR_UNSUPPORTED, "request %d: the three-piece split runs through fg_chain only", r);
    long maxm = 0;
    for (int i = 0; i < q.num_gemm; ++i) {
      if (!q.A[i] || q.M[i] <= 0 || q.K[i] <= 0) return set_error(FG_ERR_INVALID, "request %d gemm %d", r, i + 1);
      if (i > 0 && q.K[i] != q.M[i - 1])
        return set_error(FG_ERR_INVALID, "request %d gemm %d: K != previous M", r, i + 1);
      if (i + 1 < q.num_gemm) maxm = std::max(maxm, q.M[i]);
    }
    const size_t e = q.dtype == FG_DTYPE_INT8 ? 1 : q.dtype == FG_DTYPE_BF16 ? 2 : 4;
    off[r] = total;
    slot[r] = (maxm * q.N * e + 255) & ~(size_t)255;
    total += 2 * slot[r];
    levels = std::max(levels, q.num_gemm);
  }
  char *spill = nullptr;
  if (total > 0 && (rc = scratch(s, "levels", total, (void **)&spill))) return rc;
  // Per level: every GEMM's descriptor, split over k by its shape alone (plan_splits:
  // the split a lone launch of that GEMM would use to fill the chip), so a request's
  // bits never depend on the batch; split GEMMs write slabs that the level's reduction
  // adds in slab order.
  std::vector<std::vector<GemmDesc>> lev(levels);
  std::vector<std::vector<ReduceDesc>> red(levels);
  std::vector<size_t> slab_bytes(levels, 0);
  for (int L = 0; L < levels; ++L) {
    for (int r = 0; r < count; ++r) {
      const fg_chain_req &q = reqs[r];
      if (q.num_gemm <= L) continue;
      const bool last = L == q.num_gemm - 1;
      char *x0 = spill ? spill + off[r] : nullptr, *x1 = x0 ? x0 + slot[r] : nullptr;
      const void *Bin = L == 0 ? q.B : ((L - 1) & 1 ? x1 : x0);
      void *Cout = last ? q.C : (L & 1 ? x1 : x0);
      const int kind = kind_of(q.dtype, q.f32_mode), epi = last ? EPI_STORE : EPI_REQUANT;
      GemmDesc g;
      gemm_desc(&g, kind, epi, q.A[L], Bin, Cout, q.M[L], q.N, q.K[L]);
      const int sp = gemm_desc_split(&g, plan_splits(q.M[L], q.N, q.K[L], kind), nullptr);
      if (sp > 1) {
        // slab offset for now; made a pointer once the level's slabs are allocated
        g.C = (void *)(uintptr_t)slab_bytes[L];
        red[L].push_back({(const void *)(uintptr_t)slab_bytes[L], Cout, q.M[L] * q.N, sp, kind, epi, 0});
        slab_bytes[L] += ((size_t)sp * q.M[L] * q.N * 4 + 255) & ~(size_t)255;
      }
      lev[L].push_back(g);
    }
  }
  const size_t slab_max = *std::max_element(slab_bytes.begin(), slab_bytes.end());
  char *slabs = nullptr;
  if (slab_max > 0 && (rc = scratch(s, "lvslab", slab_max, (void **)&slabs))) return rc;
  // tables: a level's tiles with the longest k-slices first (kper over the kind's MFMA
  // rate), then its reduction
  auto tile_cost = [](const GemmDesc &d) {
    const double rate = d.kind == K_I8 ? 32 : d.kind == K_BF16 ? 16 : d.kind == K_F32X3 ? 16.0 / 3 : 1;
    return (double)d.kper / rate;
  };
  struct Part {
    bool reduce;
    int items;
    size_t off;
  };
  std::vector<Part> parts;
  std::vector<char> bytes;
  auto put_table = [&](const void *hdr, size_t hsz, const void *descs, size_t dbytes, const std::vector<int2> &items,
                       bool reduce) {
    const size_t at = (bytes.size() + 255) & ~(size_t)255;
    const size_t doff = align16(hsz), ioff = align16(doff + dbytes);
    bytes.resize(at + ioff + items.size() * sizeof(int2), 0);
    memcpy(bytes.data() + at, hdr, hsz);
    memcpy(bytes.data() + at + doff, descs, dbytes);
    memcpy(bytes.data() + at + ioff, items.data(), items.size() * sizeof(int2));
    parts.push_back({reduce, (int)items.size(), at});
  };
  for (int L = 0; L < levels; ++L) {
    std::vector<GemmDesc> &d = lev[L];
    for (GemmDesc &g : d)
      if (g.splits > 1) g.C = slabs + (uintptr_t)g.C;
    std::stable_sort(d.begin(), d.end(),
                     [&](const GemmDesc &a, const GemmDesc &b) { return tile_cost(a) > tile_cost(b); });
    std::vector<int2> items;
    for (size_t j = 0; j < d.size(); ++j) {
      const long units = (d[j].M + LEVEL_TILE - 1) / LEVEL_TILE * d[j].tiles_n * d[j].splits;
      for (long t = 0; t < units; ++t) items.push_back(make_int2((int)j, (int)t));
    }
    LevelTable h{};
    h.num_descs = (int)d.size();
    h.num_items = (int)items.size();
    h.descs_off = (int)align16(sizeof(LevelTable));
    h.items_off = (int)align16(h.descs_off + d.size() * sizeof(GemmDesc));
    put_table(&h, sizeof h, d.data(), d.size() * sizeof(GemmDesc), items, false);
    if (red[L].empty()) continue;
    std::vector<int2> ritems;
    for (size_t j = 0; j < red[L].size(); ++j) {
      ReduceDesc &rd = red[L][j];
      rd.P = slabs + (uintptr_t)rd.P;
      const long chunks = (rd.MN + 255) / 256;
      for (long c = 0; c < chunks; ++c) ritems.push_back(make_int2((int)j, (int)c));
    }
    ReduceTable rh{};
    rh.num_descs = (int)red[L].size();
    rh.num_items = (int)ritems.size();
    rh.descs_off = (int)align16(sizeof(ReduceTable));
    rh.items_off = (int)align16(rh.descs_off + red[L].size() * sizeof(ReduceDesc));
    put_table(&rh, sizeof rh, red[L].data(), red[L].size() * sizeof(ReduceDesc), ritems, true);
  }
  StreamRing *R = ring_for(s);
  Slot &sl = R->slot[R->next];
  R->next = (R->next + 1) % RING;
  if ((rc = slot_reserve(sl, bytes.size()))) return rc;
  memcpy(sl.host, bytes.data(), bytes.size());
  if ((rc = hip_status(hipMemcpyAsync(sl.dev, sl.host, bytes.size(), hipMemcpyHostToDevice, s), "H2D level tables")))
    return rc;
  for (const Part &p : parts) {
    const char *t = (const char *)sl.dev + p.off;
    if ((rc = hip_status(p.reduce ? launch_grouped_reduce(t, p.items, s) : launch_grouped_gemm(t, p.items, s),
                         p.reduce ? "grouped level reduction" : "grouped level launch")))
      return rc;
  }
  return hip_status(hipEventRecord(sl.done, s), "hipEventRecord");
}

int fg_grouped_fits(const fg_chain_req *req) {
  ChainParams p;
  int tn;
  size_t l;
  bool fits;
  int rc = describe(*req, &p, &tn, &l, &fits);
  if (rc) return rc;
  return fits ? 1 : 0;
}

int fg_grouped_tile(const fg_chain_req *req) {
  ChainParams p;
  int tn;
  size_t l;
  bool fits;
  int rc = describe(*req, &p, &tn, &l, &fits);
  if (rc) return rc;
  return fits ? tn : 0;
}

int fg_grouped_graph_create(int max_reqs, int max_tiles, fg_grouped_graph **out) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (max_reqs < 1 || max_tiles < 1 || !out) return set_error(FG_ERR_INVALID, "graph sizes");
  auto *G = new fg_grouped_graph;
  G->table_cap = align16(sizeof(GroupTable)) + align16((size_t)max_reqs * sizeof(ChainParams)) +
                 (size_t)max_tiles * sizeof(int2);
  hipStream_t cs;
  if ((rc = hip_status(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate"))) {
    delete G;
    return rc;
  }
  for (int i = 0; i < RING && rc == FG_OK; ++i) {
    auto &g = G->g[i];
    if ((rc = slot_reserve(g.slot, G->table_cap))) break;
    memset(g.slot.host, 0, g.slot.cap);
    // capture: table upload, then the grouped kernel (full LDS, one wave of CUs)
    if ((rc = hip_status(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture"))) break;
    hipError_t e1 = hipMemcpyAsync(g.slot.dev, g.slot.host, sizeof(GroupTable), hipMemcpyHostToDevice, cs);
    hipError_t e2 = launch_grouped(g.slot.dev, num_cus(), LDS_BYTES_MAX, false, cs);
    hipError_t e3 = hipStreamEndCapture(cs, &g.graph);
    if ((rc = hip_status(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3), "capture"))) break;
    size_t n = 0;
    if ((rc = hip_status(hipGraphGetNodes(g.graph, nullptr, &n), "hipGraphGetNodes"))) break;
    std::vector<hipGraphNode_t> nodes(n);
    if ((rc = hip_status(hipGraphGetNodes(g.graph, nodes.data(), &n), "hipGraphGetNodes"))) break;
    for (auto nd : nodes) {
      hipGraphNodeType t;
      if (hipGraphNodeGetType(nd, &t) == hipSuccess && t == hipGraphNodeTypeMemcpy) g.copy = nd;
    }
    if (!g.copy) {
      rc = set_error(FG_ERR_HIP, "captured graph has no copy node");
      break;
    }
    rc = hip_status(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  (void)hipStreamDestroy(cs);
  if (rc) {
    fg_grouped_graph_destroy(G);
    return rc;
  }
  *out = G;
  return FG_OK;
}

int fg_grouped_graph_launch(fg_grouped_graph *G, int count, const fg_chain_req *reqs, void *stream) {
  int rc;
  if (!G || count < 0 || (count > 0 && !reqs)) return set_error(FG_ERR_INVALID, "graph launch");
  Packed pk;
  std::vector<int> rest;
  if ((rc = pack(count, reqs, (hipStream_t)stream, &pk, &rest))) return rc;
  if (!rest.empty())
    return set_error(FG_ERR_UNSUPPORTED, "request %d does not fit on chip at the grouped tile width", rest[0]);
  if (pk.bytes.size() > G->table_cap) return set_error(FG_ERR_INVALID, "batch exceeds the graph's table");
  auto &g = G->g[G->next];
  G->next = (G->next + 1) % RING;
  if ((rc = slot_reserve(g.slot, G->table_cap))) return rc;  // waits for this slot's last replay
  memcpy(g.slot.host, pk.bytes.data(), pk.bytes.size());
  if (pk.wide) {
    // the captured kernel runs 32-column tiles only: a batch with 64-column requests
    // is launched directly from the same slot
    hipStream_t s = (hipStream_t)stream;
    if ((rc = hip_status(hipMemcpyAsync(g.slot.dev, g.slot.host, pk.bytes.size(), hipMemcpyHostToDevice, s),
                         "H2D table")))
      return rc;
    if ((rc = hip_status(launch_grouped(g.slot.dev, num_cus(), LDS_BYTES_MAX, true, s), "grouped launch")))
      return rc;
    return hip_status(hipEventRecord(g.slot.done, s), "hipEventRecord");
  }
  if ((rc = hip_status(hipGraphExecMemcpyNodeSetParams1D(g.exec, g.copy, g.slot.dev, g.slot.host,
                                                         pk.bytes.size(), hipMemcpyHostToDevice),
                       "update copy node")))
    return rc;
  if ((rc = hip_status(hipGraphLaunch(g.exec, (hipStream_t)stream), "hipGraphLaunch"))) return rc;
  return hip_status(hipEventRecord(g.slot.done, (hipStream_t)stream), "hipEventRecord");
}

void fg_grouped_graph_destroy(fg_grouped_graph *G) {
  if (!G) return;
  for (auto &g : G->g) {
    if (g.slot.done) (void)hipEventSynchronize(g.slot.done);
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.slot.host) (void)hipHostFree(g.slot.host);
    if (g.slot.dev) (void)hipFree(g.slot.dev);
    if (g.slot.done) (void)hipEventDestroy(g.slot.done);
  }
  delete G;
}

}  // extern "C"
