// Serving engine: continuous batching of chain requests against device-resident
// weight sets. No reference counterpart (the reference is a blocking library call);
// this is the request loop around the grouped chain launches.
//
//   * weight sets are registered once and stay resident (charged to the HBM budget);
//   * requests carry device buffers B (K_1 x N) and C (M_n x N) and are admitted,
//     in arrival order, while a slot is free and their bytes fit the budget;
//   * a scheduler thread packs every admitted request into the next grouped launch
//     (replayed from a captured graph, or launched directly) on one of several
//     streams, so requests that arrive while others run are admitted and their work
//     overlaps on the device; requests too large for the grouped tile run unfused
//     on the same stream, level by level (fg_grouped_levels: GEMM i of all of them
//     in one launch);
//   * completion is observed through events; latency = completion - arrival.
// In per-GEMM mode the same scheduler runs each request alone with one launch per
// GEMM (the plain-kernel baseline).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/flashgemm.h"
#include "fg_internal.h"
#include "fg_runtime.h"

using namespace fg;
using Clock = std::chrono::steady_clock;

namespace {

struct WeightSet {
  int dtype, f32_mode, num_gemm;
  void *A[CHAIN_MAXG];       // the caller's weights (unfused and per-GEMM runs)
  void *packed[CHAIN_MAXG];  // packed once at registration (grouped fused runs)
  long M[CHAIN_MAXG], K[CHAIN_MAXG];
  size_t bytes;              // raw + packed bytes, charged to the budget
};

struct Request {
  long id;
  int set;
  long N;
  const void *B;
  void *C;
  size_t bytes;
  Clock::time_point arrival, done;
  int status;
  bool cancelled;
};

// The requests of one part of a launched batch, completed when `ev` has: the on-chip
// grouped launch and the level launch of one batch are two parts on one stream, so
// the first part's requests retire without waiting for the second's work.
struct Batch {
  std::vector<long> ids;
  hipEvent_t ev;
  int stream;
  bool last;  // the batch's last part: its stream is free for another batch after it
};

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

}  // namespace

struct fg_server {
  fg_server_config cfg;
  std::vector<WeightSet> sets;
  size_t weight_bytes = 0;

  std::mutex mu;
  std::condition_variable cv;       // drain(): a batch completed
  std::condition_variable work_cv;  // scheduler: a request was submitted, or stop
  std::deque<long> pending;          // submitted, not admitted
  std::vector<long> admitted;        // admitted, not launched
  std::vector<Request> reqs;         // by id
  std::vector<Batch> inflight;
  size_t used_bytes = 0;
  int used_slots = 0;
  long completed = 0;
  int first_error = FG_OK;
  std::string error_text;

  std::vector<hipStream_t> streams;
  std::vector<fg_grouped_graph *> graphs;
  long graph_tiles = 0;  // tile capacity of each captured graph's table
  std::vector<hipEvent_t> free_events;
  int next_stream = 0;
  std::atomic<bool> stop{false};
  std::thread worker;
  Clock::time_point t0 = Clock::now();
};

namespace {

// One scheduler pass. Returns true if it did anything.
bool step(fg_server *S) {
  bool did = false;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    // 1. retire completed batches
    for (size_t i = 0; i < S->inflight.size();) {
      hipError_t q = hipEventQuery(S->inflight[i].ev);
      if (q == hipErrorNotReady) {
        ++i;
        continue;
      }
      auto now = Clock::now();
      for (long id : S->inflight[i].ids) {
        Request &r = S->reqs[id];
        r.done = now;
        if (q != hipSuccess) r.status = FG_ERR_HIP;
        S->used_bytes -= r.bytes;
        S->used_slots -= 1;
        S->completed += 1;
      }
      S->free_events.push_back(S->inflight[i].ev);
      S->inflight.erase(S->inflight.begin() + i);
      did = true;
    }
    // 2. admit in arrival order while a slot and the budget allow
    while (!S->pending.empty()) {
      Request &r = S->reqs[S->pending.front()];
      if (S->used_slots >= S->cfg.max_inflight) break;
      if (S->used_bytes + r.bytes > S->cfg.hbm_budget_bytes - S->weight_bytes) break;
      S->used_slots += 1;
      S->used_bytes += r.bytes;
      S->admitted.push_back(r.id);
      S->pending.pop_front();
      did = true;
    }
    if (did) S->cv.notify_all();
  }
  // 3. launch everything admitted on the next stream (at most num_streams batches
  //    of this engine in flight, so a batch waits for a stream rather than queueing
  //    behind a long one)
  std::vector<long> batch;
  // Descriptors of the batch, built under the lock: submit() and add_weights() may
  // grow the request and weight-set vectors while this thread launches.
  std::vector<fg_chain_req> q;
  int sidx = -1;
  hipEvent_t ev = nullptr, ev2 = nullptr;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    const int busy = (int)std::count_if(S->inflight.begin(), S->inflight.end(), [](const Batch &b) { return b.last; });
    if (S->admitted.empty() || busy >= S->cfg.num_streams) return did;
    int take = std::min<int>((int)S->admitted.size(), S->cfg.max_batch);
    batch.assign(S->admitted.begin(), S->admitted.begin() + take);
    S->admitted.erase(S->admitted.begin(), S->admitted.begin() + take);
    sidx = S->next_stream;
    S->next_stream = (S->next_stream + 1) % S->cfg.num_streams;
    for (hipEvent_t *e : {&ev, &ev2})
      if (!S->free_events.empty()) {
        *e = S->free_events.back();
        S->free_events.pop_back();
      }
    q.resize(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
      const Request &r = S->reqs[batch[i]];
      const WeightSet &w = S->sets[r.set];
      fg_chain_req &x = q[i];
      memset(&x, 0, sizeof x);
      x.dtype = w.dtype;
      x.f32_mode = w.f32_mode;
      x.C = r.C;
      x.B = r.B;
      x.N = r.N;
      x.num_gemm = w.num_gemm;
      for (int g = 0; g < w.num_gemm; ++g) {
        x.A[g] = w.A[g];
        x.M[g] = w.M[g];
        x.K[g] = w.K[g];
      }
      if (!S->cfg.per_gemm_baseline && fg_grouped_fits(&x) == 1) {
        for (int g = 0; g < w.num_gemm; ++g) x.A[g] = w.packed[g];
        x.a_packed = 1;
      }
    }
  }
  for (hipEvent_t *e : {&ev, &ev2})
    if (!*e) (void)hipEventCreateWithFlags(e, hipEventDisableTiming);
  hipStream_t st = S->streams[sidx];
  int rc = FG_OK;
  std::vector<long> head_ids;  // the first part (on-chip grouped launch) when both run
  if (S->cfg.per_gemm_baseline) {
    for (const fg_chain_req &x : q) {
      fg_chain_opts o;
      fg_chain_opts_default(&o);
      o.f32_mode = x.f32_mode;
      o.fuse = FG_FUSE_OFF;
      o.splits = 1;
      rc = fg_chain(x.dtype, x.C, x.B, x.N, x.num_gemm, x.A, x.M, x.K, &o, st);
      if (rc) break;
    }
  } else {
    std::vector<fg_chain_req> fit, rest;
    std::vector<long> rest_ids;
    long tiles = 0;
    for (size_t i = 0; i < q.size(); ++i) {
      const fg_chain_req &x = q[i];
      (x.a_packed ? fit : rest).push_back(x);
      (x.a_packed ? head_ids : rest_ids).push_back(batch[i]);
      if (x.a_packed) tiles += (x.N + grouped_tile_n(kind_of(x.dtype, x.f32_mode)) - 1) /
                               grouped_tile_n(kind_of(x.dtype, x.f32_mode));
    }
    if (!fit.empty()) {
      // a batch of very wide requests can hold more tiles than the captured graph's
      // table: it is launched directly instead
      if (S->cfg.use_graph && tiles <= S->graph_tiles)
        rc = fg_grouped_graph_launch(S->graphs[sidx], (int)fit.size(), fit.data(), st);
      else
        rc = fg_grouped_chain((int)fit.size(), fit.data(), st);
      if (!rc && !rest.empty()) rc = hip_status(hipEventRecord(ev2, st), "hipEventRecord");
    }
    // the others run unfused, level by level: GEMM i of all of them in one launch, so
    // their tiles fill the chip together instead of a few CUs per request and one
    // launch per GEMM
    if (!rc && !rest.empty()) rc = fg_grouped_levels((int)rest.size(), rest.data(), st);
    if (fit.empty() || rest.empty()) head_ids.clear();  // one part
    else batch = rest_ids;
  }
  if (!rc) rc = hip_status(hipEventRecord(ev, st), "hipEventRecord");
  std::lock_guard<std::mutex> lk(S->mu);
  if (rc) {
    if (S->first_error == FG_OK) {
      S->first_error = rc;
      S->error_text = fg_last_error();
    }
    for (long id : batch) S->reqs[id].status = rc;
    for (long id : head_ids) S->reqs[id].status = rc;
  }
  if (!head_ids.empty()) S->inflight.push_back({head_ids, ev2, sidx, false});
  else S->free_events.push_back(ev2);
  S->inflight.push_back({batch, ev, sidx, true});
  return true;
}

// The scheduler sleeps while the engine is idle (nothing submitted, nothing on the
// device) until a submit or stop wakes it. While batches are in flight it keeps
// polling their events, yielding between passes: a timed wait between polls (10 us,
// timer slack 1 us) delayed each completion by its sleep and, at 20,000 requests/s,
// raised the p99 latency from 3.5 to 14 ms.
void worker_main(fg_server *S) {
  (void)hipSetDevice(0);
  while (!S->stop.load()) {
    if (step(S)) continue;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->stop.load()) break;
    if (S->inflight.empty() && S->admitted.empty() && S->pending.empty()) {
      S->work_cv.wait(lk, [&] { return S->stop.load() || !S->pending.empty(); });
    } else {
      lk.unlock();
      std::this_thread::yield();
    }
  }
}

}  // namespace

extern "C" {

void fg_server_config_default(fg_server_config *c) {
  c->hbm_budget_bytes = (size_t)200 << 30;
  c->max_inflight = 1024;
  c->num_streams = 2;
  c->max_batch = 512;
  c->use_graph = 1;
  c->per_gemm_baseline = 0;
}

int fg_server_create(const fg_server_config *cfg, fg_server **out) {
  int rc;
  if ((rc = ensure_init())) return rc;
  if (!out) return set_error(FG_ERR_INVALID, "out is NULL");
  auto *S = new fg_server;
  if (cfg) S->cfg = *cfg; else fg_server_config_default(&S->cfg);
  if (S->cfg.num_streams < 1 || S->cfg.num_streams > 8 || S->cfg.max_inflight < 1 || S->cfg.max_batch < 1) {
    delete S;
    return set_error(FG_ERR_INVALID, "server config");
  }
  for (int i = 0; i < S->cfg.num_streams && !rc; ++i) {
    hipStream_t s;
    rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    if (!rc) S->streams.push_back(s);
    if (!rc && S->cfg.use_graph && !S->cfg.per_gemm_baseline) {
      fg_grouped_graph *g = nullptr;
      S->graph_tiles = (long)S->cfg.max_batch * 160;
      rc = fg_grouped_graph_create(S->cfg.max_batch, (int)S->graph_tiles, &g);
      S->graphs.push_back(g);
    }
  }
  if (rc) {
    fg_server_destroy(S);
    return rc;
  }
  S->worker = std::thread(worker_main, S);
  *out = S;
  return FG_OK;
}

int fg_server_add_weights(fg_server *S, int dtype, int f32_mode, int num_gemm, const void *const *A, const long *M,
                          const long *K, int *set_id) {
  if (!S || dtype < 0 || dtype > 2 || num_gemm < 1 || num_gemm > CHAIN_MAXG || !A || !set_id)
    return set_error(FG_ERR_INVALID, "add_weights");
  WeightSet w{};
  w.dtype = dtype;
  w.f32_mode = f32_mode;
  w.num_gemm = num_gemm;
  size_t e = dtype == FG_DTYPE_F32 ? 4 : dtype == FG_DTYPE_BF16 ? 2 : 1;
  const bool pack = !S->cfg.per_gemm_baseline;
  for (int i = 0; i < num_gemm; ++i) {
    if (!A[i] || M[i] <= 0 || K[i] <= 0 || (i > 0 && K[i] != M[i - 1]))
      return set_error(FG_ERR_INVALID, "weight shapes");
    w.A[i] = (void *)A[i];
    w.M[i] = M[i];
    w.K[i] = K[i];
    w.bytes += (size_t)M[i] * K[i] * e;
    if (pack) w.bytes += fg_packed_weight_bytes(dtype, f32_mode, M[i], K[i]);
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->weight_bytes + w.bytes > S->cfg.hbm_budget_bytes)
      return set_error(FG_ERR_BUDGET, "weights exceed the HBM budget");
    S->weight_bytes += w.bytes;  // reserved while packing
  }
  int rc = FG_OK;
  if (pack) {
    // pack once on the first server stream; registration waits for it
    hipStream_t st = S->streams[0];
    for (int i = 0; i < num_gemm && !rc; ++i) {
      rc = hip_status(dev_alloc(&w.packed[i], fg_packed_weight_bytes(dtype, f32_mode, M[i], K[i])),
                      "hipMalloc(packed weights)");
      if (!rc) rc = fg_pack_weight(dtype, f32_mode, A[i], M[i], K[i], w.packed[i], st);
    }
    if (!rc) rc = hip_status(hipStreamSynchronize(st), "hipStreamSynchronize");
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (rc) {
    for (int i = 0; i < num_gemm; ++i)
      if (w.packed[i]) (void)hipFree(w.packed[i]);
    S->weight_bytes -= w.bytes;
    return rc;
  }
  *set_id = (int)S->sets.size();
  S->sets.push_back(w);
  return FG_OK;
}

namespace {
// Latency is measured from `arrival`: the submit call's entry, or a trace's due time
// (so time a late submitter spends behind the lock counts).
int submit_at(fg_server *S, int set_id, long N, const void *B, void *C, long *req_id, Clock::time_point arrival) {
  if (!S || !B || !C || N <= 0) return set_error(FG_ERR_INVALID, "submit");
  std::lock_guard<std::mutex> lk(S->mu);
  if (set_id < 0 || set_id >= (int)S->sets.size()) return set_error(FG_ERR_INVALID, "weight set %d", set_id);
  const WeightSet &w = S->sets[set_id];
  size_t e = w.dtype == FG_DTYPE_F32 ? 4 : w.dtype == FG_DTYPE_BF16 ? 2 : 1;
  Request r{};
  r.id = (long)S->reqs.size();
  r.set = set_id;
  r.N = N;
  r.B = B;
  r.C = C;
  r.bytes = (size_t)w.K[0] * N * e + (size_t)w.M[w.num_gemm - 1] * N * 4;
  // A request that cannot ride the on-chip grouped launch runs in the level launch and
  // takes library scratch (two intermediates, split-K slabs): that is charged too.
  fg_chain_req x;
  memset(&x, 0, sizeof x);
  x.dtype = w.dtype;
  x.f32_mode = w.f32_mode;
  x.C = C;
  x.B = B;
  x.N = N;
  x.num_gemm = w.num_gemm;
  for (int g = 0; g < w.num_gemm; ++g) {
    x.A[g] = w.A[g];
    x.M[g] = w.M[g];
    x.K[g] = w.K[g];
  }
  const int fits = S->cfg.per_gemm_baseline ? 0 : fg_grouped_fits(&x);
  if (fits < 0) return fits;
  if (!fits)
    r.bytes += chain_scratch_bound(w.dtype, w.f32_mode, N, w.num_gemm, w.M, w.K, false);
  if (r.bytes + S->weight_bytes > S->cfg.hbm_budget_bytes)
    return set_error(FG_ERR_BUDGET, "request of %zu bytes can never fit the budget", r.bytes);
  r.arrival = arrival;
  r.status = FG_OK;
  S->reqs.push_back(r);
  S->pending.push_back(r.id);
  S->work_cv.notify_one();
  if (req_id) *req_id = r.id;
  return FG_OK;
}
}  // namespace

int fg_server_submit(fg_server *S, int set_id, long N, const void *B, void *C, long *req_id) {
  return submit_at(S, set_id, N, B, C, req_id, Clock::now());
}

int fg_server_drain(fg_server *S, double timeout_s) {
  if (!S) return set_error(FG_ERR_INVALID, "drain");
  std::unique_lock<std::mutex> lk(S->mu);
  bool ok = S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                           [&] { return S->completed == (long)S->reqs.size(); });
  if (S->first_error) return set_error(S->first_error, "server: %s", S->error_text.c_str());
  if (!ok) return set_error(FG_ERR_HIP, "drain timed out with %ld of %zu done", S->completed, S->reqs.size());
  return FG_OK;
}

int fg_server_cancel(fg_server *S, long req_id) {
  if (!S) return set_error(FG_ERR_INVALID, "cancel");
  std::lock_guard<std::mutex> lk(S->mu);
  if (req_id < 0 || req_id >= (long)S->reqs.size()) return set_error(FG_ERR_INVALID, "cancel: no request %ld", req_id);
  // only a request still waiting for admission holds nothing (no slot, no budget, no
  // launch): the scheduler takes requests off `pending` under this lock
  auto it = std::find(S->pending.begin(), S->pending.end(), req_id);
  if (it == S->pending.end())
    return set_error(FG_ERR_INVALID, "cancel: request %ld is already admitted, running or done", req_id);
  S->pending.erase(it);
  S->reqs[req_id].cancelled = true;
  S->completed += 1;  // drain counts it as settled
  S->cv.notify_all();
  return FG_OK;
}

int fg_server_run_trace(fg_server *S, int n, const int *set_ids, const long *Ns, const void *const *Bs,
                        void *const *Cs, const double *arrival_s, double timeout_s) {
  if (!S || n < 0) return set_error(FG_ERR_INVALID, "run_trace");
  auto start = Clock::now();
  for (int i = 0; i < n; ++i) {
    auto due = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
    while (Clock::now() < due) {
      if (due - Clock::now() > std::chrono::microseconds(200))
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    int rc = submit_at(S, set_ids[i], Ns[i], Bs[i], Cs[i], nullptr, due);
    if (rc) return rc;
  }
  return fg_server_drain(S, timeout_s);
}

// Latency (completion - arrival, seconds) of requests [first, first+n).
int fg_server_latencies(fg_server *S, long first, long n, double *out) {
  if (!S || !out || first < 0) return set_error(FG_ERR_INVALID, "latencies");
  std::lock_guard<std::mutex> lk(S->mu);
  if (first + n > (long)S->reqs.size()) return set_error(FG_ERR_INVALID, "latencies: only %zu requests", S->reqs.size());
  for (long i = 0; i < n; ++i) {
    const Request &r = S->reqs[first + i];
    out[i] = r.cancelled ? -2.0 : (r.done.time_since_epoch().count() == 0) ? -1.0 : seconds(r.arrival, r.done);
  }
  return FG_OK;
}

long fg_server_num_requests(fg_server *S) {
  if (!S) return 0;
  std::lock_guard<std::mutex> lk(S->mu);
  return (long)S->reqs.size();
}

void fg_server_destroy(fg_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop.store(true);
    S->work_cv.notify_all();
  }
  if (S->worker.joinable()) S->worker.join();
  for (auto s : S->streams) (void)hipStreamSynchronize(s);
  for (auto &b : S->inflight) (void)hipEventDestroy(b.ev);
  for (auto e : S->free_events) (void)hipEventDestroy(e);
  for (auto g : S->graphs) fg_grouped_graph_destroy(g);
  for (auto s : S->streams) {
    // the library's per-stream scratch and table ring die with the stream
    release_stream_scratch(s);
    release_stream_ring(s);
    (void)hipStreamDestroy(s);
  }
  for (auto &w : S->sets)
    for (int i = 0; i < w.num_gemm; ++i)
      if (w.packed[i]) (void)hipFree(w.packed[i]);
  delete S;
}

}  // extern "C"
