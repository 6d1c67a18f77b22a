// Serving engine: APSP and sub-graph requests on resident graphs, continuous batching
// under an HBM budget (declared in fap_hip.h, fap_server_*).
//
// Clients register preconditioned graphs once (their CSR, layout and float/int
// weights stay resident) and submit requests from any thread: a full APSP of a graph
// (the n x n result of solve()) or one sub-graph block (the s_p x n rows of
// solveSubGraph(id)).  A scheduler thread owns the GPU side:
//   * admission: waiting requests are taken in arrival order while the bytes they
//     need -- their result buffer plus their share of the batch scratch -- fit in the
//     budget next to everything already admitted and not yet released;
//   * continuous batching: all admitted requests go into one new batch (the persistent
//     work-queue launch of fap_hip_batch_*) on a free stream, so requests arriving while
//     others run start as soon as a stream and memory are free, and their work
//     overlaps on the device with the batches already running;
//   * completion: an event per batch; when it has passed, its requests are done (their
//     latency is submit -> observed completion), the batch scratch is freed and the
//     result stays resident until the client releases it.
// Requests on graphs of kLargeN vertices or more take a second lane: a worker thread
// runs them one after another through the stage-batched pipeline
// (fap_hip_solve_parts / _f32, whose kernels fill the GPU from a single graph) on a
// stream of its own, beside the batches.  The pipeline's workspaces belong to the
// library, so while a server runs no other thread may call solve() or
// fap_hip_solve_parts.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "capi_internal.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct GraphEntry {
  fap::fapGraph *g;
  fap_graph_dev dev;
  std::vector<fap_part_desc> parts;
  bool is_float;
  int64_t ld;
};

enum State { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  int graph = 0, sub_id = 0;
  int only_part = 0;  // > 0: parts tile the graph, only this part's rows are produced
  int state = kPending;
  double t_submit = 0, t_done = 0;
  void *d_out = nullptr;
  int64_t rows = 0, cols = 0, ld = 0, bytes = 0, scratch = 0;
  std::vector<fap_part_desc> parts;
};

struct Inflight {
  fap_batch *batch = nullptr;
  void *event = nullptr;
  int slot = 0;
  int64_t scratch = 0;
  std::vector<int64_t> tickets;
};

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

constexpr int kLargeN = 8192;  // graphs from this size take the pipeline lane

}  // namespace

struct fap_server {
  int64_t budget = 0;
  std::vector<void *> streams;
  std::vector<bool> busy;
  std::vector<GraphEntry> graphs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending;        // batch lane
  std::deque<int64_t> pending_large;  // pipeline lane
  void *pipe_stream = nullptr;
  std::thread pipe_worker;
  std::map<int64_t, Request> reqs;
  std::vector<Inflight> inflight;
  std::vector<double> latency_ms;
  int64_t next_ticket = 1, reserved = 0, peak_reserved = 0, submitted = 0, completed = 0,
          failed = 0, batches = 0;
  double entries = 0, t_first = 0, t_last = 0;
  bool stop = false;
  std::thread worker;
};

namespace {

// Scratch a request brings to a batch, an upper bound of what fap_hip_batch_create
// allocates for it (rows of n entries padded to 256 columns): each part's padded
// in-part matrix; when the parts tile the graph (a full APSP) the padded
// boundary-graph matrix and the boundary rows of the whole graph (in vertex and in
// sorted-position order), otherwise (a
// sub-graph) each part's SSSP boundary rows; and the work tables (merge units of at
// least 32 rows x 256 columns at 16 bytes each plus the per-XCD slot padding,
// boundary-row units of at least 16 rows x 256 columns, part and graph records).
int64_t scratch_estimate(const Request &r, bool is_float, int n) {
  const int elem = is_float ? 8 : 4;
  const int64_t ldr = align_up(n, 256);
  int64_t bytes = 0, tables = 4096, nb = 0, covered = 0, brow_units = 0;
  for (const fap_part_desc &p : r.parts) {
    const int64_t S = align_up(p.num, 64);
    bytes += align_up(S * S * elem, 256);
    nb += p.bdy;
    covered += p.num;
    const int64_t rbs = (p.num - p.bdy + 31) / 32;
    tables += (rbs * ((n + 255) / 256) + 8 * rbs) * 16 + 4 * p.part_id + 512;
    brow_units += (p.num + 255) / 256;
  }
  if (covered == n && nb > 0) {
    const int64_t SB = align_up(nb, 64);
    const int64_t rows = r.only_part > 0 ? std::max(r.parts[r.only_part - 1].bdy, 1) : nb;
    bytes += align_up(SB * SB * elem, 256) + 2 * align_up(rows * ldr * elem, 256);
    tables += brow_units * ((nb + 15) / 16) * 16 + 12 * nb;
  } else {
    for (const fap_part_desc &p : r.parts) bytes += align_up((int64_t)p.bdy * ldr * elem, 256);
    tables += 8 * nb;
  }
  return bytes + align_up(tables, 256) * 2 + (1 << 16);
}

void scheduler(fap_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    bool progress = false;
    // Completions.
    for (size_t i = 0; i < s->inflight.size();) {
      Inflight &f = s->inflight[i];
      int32_t done = 0;
      // A failed query means the batch's work cannot be trusted: wait for its stream to
      // drain (so nothing still writes the buffers), then fail its requests and free
      // their results instead of reporting them done.
      const bool failed = fap_hip_event_query(f.event, &done) != FAP_OK;
      if (failed) (void)fap_hip_stream_sync(s->streams[f.slot]);
      if (!done && !failed) {
        ++i;
        continue;
      }
      const double t = now_s();
      for (int64_t tk : f.tickets) {
        Request &r = s->reqs[tk];
        r.t_done = t;
        if (failed) {
          r.state = kFailed;
          ++s->failed;
          if (r.d_out) fap_hip_free(r.d_out);
          r.d_out = nullptr;
          s->reserved -= r.bytes;
          continue;
        }
        r.state = kDone;
        s->latency_ms.push_back((t - r.t_submit) * 1e3);
        s->entries += (double)r.rows * r.cols;
        ++s->completed;
      }
      s->t_last = t;
      fap_hip_batch_destroy(f.batch);
      fap_hip_event_destroy(f.event);
      s->reserved -= f.scratch;
      s->busy[f.slot] = false;
      s->inflight.erase(s->inflight.begin() + i);
      progress = true;
    }
    if (progress) s->cv_done.notify_all();
    if (s->stop && s->pending.empty() && s->inflight.empty()) break;

    // Admission into a new batch on a free stream.
    int slot = -1;
    for (size_t i = 0; i < s->busy.size(); ++i)
      if (!s->busy[i]) {
        slot = (int)i;
        break;
      }
    if (slot >= 0 && !s->pending.empty()) {
      std::vector<int64_t> take;
      int64_t need = 0;
      while (!s->pending.empty()) {
        Request &r = s->reqs[s->pending.front()];
        const int64_t want = r.bytes + r.scratch;
        if (want > s->budget) {  // larger than the whole budget: fail, do not block the queue
          r.state = kFailed;
          ++s->failed;
          s->pending.pop_front();
          s->cv_done.notify_all();
          continue;
        }
        if (s->reserved + need + want > s->budget) {
          if (s->stop && take.empty() && s->inflight.empty()) {  // shutting down: nothing frees
            r.state = kFailed;
            ++s->failed;
            s->pending.pop_front();
            s->cv_done.notify_all();
            continue;
          }
          break;  // wait for memory (arrival order kept)
        }
        need += want;
        take.push_back(s->pending.front());
        s->pending.pop_front();
      }
      if (!take.empty()) {
        s->reserved += need;
        s->peak_reserved = std::max(s->peak_reserved, s->reserved);
        s->busy[slot] = true;
        ++s->batches;
        std::vector<fap_batch_item> items;
        int64_t scratch = 0;
        bool ok = true;
        for (int64_t tk : take) {
          Request &r = s->reqs[tk];
          const GraphEntry &ge = s->graphs[r.graph];
          if (fap_hip_malloc(&r.d_out, std::max<int64_t>(r.bytes, 4)) != FAP_OK) ok = false;
          r.state = kRunning;
          scratch += r.scratch;
          fap_batch_item it{};
          it.g = ge.dev;
          it.parts = r.parts.data();
          it.nparts = (int32_t)r.parts.size();
          it.is_float = ge.is_float ? 1 : 0;
          it.d_out = r.d_out;
          it.ld_out = r.ld;
          it.flags = ge.g->boundaryBySssp() ? FAP_ITEM_BOUNDARY_SSSP : 0;
          it.only_part = r.only_part;
          items.push_back(it);
        }
        Inflight f;
        f.slot = slot;
        f.scratch = scratch;
        f.tickets = take;
        void *stream = s->streams[slot];
        lk.unlock();  // allocation, upload and launch without holding the lock
        // Floyd-Warshall on whole parts: the nested closure's host planning (a copy of
        // each graph and a sub-partition of each part, per batch) lengthened p50 from
        // 1.0 to 2.6 ms and held the saturated server at a third of its throughput.
        if (ok)
          ok = fap_hip_batch_create(&f.batch, items.data(), (int32_t)items.size(),
                                    FAP_BATCH_NO_NESTED) == FAP_OK;
        if (ok) ok = fap_hip_batch_run(f.batch, stream) == FAP_OK;
        if (ok) ok = fap_hip_event_create(&f.event) == FAP_OK;
        if (ok) ok = fap_hip_event_record(f.event, stream) == FAP_OK;
        lk.lock();
        if (ok) {
          s->inflight.push_back(f);
        } else {
          if (f.batch) fap_hip_batch_destroy(f.batch);
          if (f.event) fap_hip_event_destroy(f.event);
          for (int64_t tk : take) {
            Request &r = s->reqs[tk];
            r.state = kFailed;
            ++s->failed;
            if (r.d_out) fap_hip_free(r.d_out);
            r.d_out = nullptr;
            s->reserved -= r.bytes;
          }
          s->reserved -= scratch;
          s->busy[slot] = false;
          s->cv_done.notify_all();
        }
        continue;
      }
    }
    s->cv_work.wait_for(lk, std::chrono::microseconds(100));
  }
}

// The pipeline lane: large requests in arrival order, one at a time.
void pipeline_lane(fap_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    if (s->pending_large.empty()) {
      if (s->stop) break;
      s->cv_work.wait_for(lk, std::chrono::milliseconds(1));
      continue;
    }
    Request &r = s->reqs[s->pending_large.front()];
    if (r.bytes > s->budget) {
      r.state = kFailed;
      ++s->failed;
      s->pending_large.pop_front();
      s->cv_done.notify_all();
      continue;
    }
    if (s->reserved + r.bytes > s->budget) {
      if (s->stop) {  // shutting down: nothing will be released
        r.state = kFailed;
        ++s->failed;
        s->pending_large.pop_front();
        s->cv_done.notify_all();
        continue;
      }
      s->cv_work.wait_for(lk, std::chrono::milliseconds(1));  // wait for releases
      continue;
    }
    const int64_t tk = s->pending_large.front();
    s->pending_large.pop_front();
    s->reserved += r.bytes;
    s->peak_reserved = std::max(s->peak_reserved, s->reserved);
    r.state = kRunning;
    const GraphEntry &ge = s->graphs[r.graph];
    const fap_graph_dev dev = ge.dev;
    const bool is_float = ge.is_float;
    const bool by_sssp = (ge.g->solveModeFlags() & FAP_MODE_FLAG_SSSP) != 0;
    std::vector<fap_part_desc> parts = r.parts;
    const int64_t bytes = r.bytes, ld = r.ld;
    const int only_part = r.only_part;
    lk.unlock();
    void *d_out = nullptr;
    bool ok = fap_hip_malloc(&d_out, std::max<int64_t>(bytes, 4)) == FAP_OK;
    // int32 FW keeps the call asynchronous (the 16-bit form reads flags back).
    const int32_t mode = FAP_MODE_BATCHED | FAP_MODE_FLAG_INT32_FW |
                         (by_sssp ? FAP_MODE_FLAG_SSSP : 0);
    if (ok && only_part > 0) {  // one part's rows from the boundary graph
      fap_batch_item it{};
      it.g = dev;
      it.parts = parts.data();
      it.nparts = (int32_t)parts.size();
      it.is_float = is_float ? 1 : 0;
      it.d_out = d_out;
      it.ld_out = ld;
      it.only_part = only_part;
      ok = fap_hip_solve_batch(&it, 1, FAP_MODE_BATCHED, s->pipe_stream) == FAP_OK;
    } else if (ok) {
      ok = (is_float ? fap_hip_solve_parts_f32(&dev, parts.data(), (int32_t)parts.size(),
                                               static_cast<float *>(d_out), ld, mode,
                                               s->pipe_stream)
                     : fap_hip_solve_parts(&dev, parts.data(), (int32_t)parts.size(),
                                           static_cast<int32_t *>(d_out), ld, mode,
                                           s->pipe_stream)) == FAP_OK;
    }
    if (ok) ok = fap_hip_stream_sync(s->pipe_stream) == FAP_OK;
    const double t = now_s();
    lk.lock();
    Request &q = s->reqs[tk];
    q.d_out = d_out;
    if (ok) {
      q.state = kDone;
      q.t_done = t;
      s->latency_ms.push_back((t - q.t_submit) * 1e3);
      s->entries += (double)q.rows * q.cols;
      ++s->completed;
      s->t_last = t;
    } else {
      q.state = kFailed;
      ++s->failed;
      if (d_out) fap_hip_free(d_out);
      q.d_out = nullptr;
      s->reserved -= q.bytes;
    }
    s->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int fap_server_create(fap_server **out, int64_t budget_bytes, int32_t nstreams) {
  if (!out || budget_bytes <= 0 || nstreams < 1 || nstreams > 16) {
    fap_hip::set_error("fap_server_create: bad arguments");
    return FAP_ERR_INVALID;
  }
  *out = nullptr;
  int st = fap_hip_init(0);
  if (st != FAP_OK) return st;
  fap_server *s = new fap_server;
  s->budget = budget_bytes;
  for (int i = 0; i < nstreams && st == FAP_OK; ++i) {
    void *p = nullptr;
    st = fap_hip_stream_create(&p);
    if (st == FAP_OK) s->streams.push_back(p);
  }
  if (st != FAP_OK) {
    for (void *p : s->streams) fap_hip_stream_destroy(p);
    delete s;
    return st;
  }
  if (st == FAP_OK) st = fap_hip_stream_create(&s->pipe_stream);
  if (st != FAP_OK) {
    for (void *p : s->streams) fap_hip_stream_destroy(p);
    delete s;
    return st;
  }
  s->busy.assign(s->streams.size(), false);
  s->worker = std::thread(scheduler, s);
  s->pipe_worker = std::thread(pipeline_lane, s);
  *out = s;
  return FAP_OK;
}

int fap_server_destroy(fap_server *s) {
  if (!s) return FAP_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();       // drains pending and in-flight batches
  s->pipe_worker.join();  // and the pipeline lane
  for (auto &kv : s->reqs)
    if (kv.second.d_out) fap_hip_free(kv.second.d_out);
  for (void *p : s->streams) fap_hip_stream_destroy(p);
  fap_hip_stream_destroy(s->pipe_stream);
  delete s;
  return FAP_OK;
}

int fap_server_add_graph(fap_server *s, fap_graph *g, int32_t *index) {
  if (!s || !g || !g->g || !index) {
    fap_hip::set_error("fap_server_add_graph: null argument");
    return FAP_ERR_INVALID;
  }
  return guarded([&] {
    GraphEntry e;
    e.g = g->g;
    e.dev = g->g->deviceView();
    e.parts = g->g->partList();
    e.is_float = g->g->isFloat();
    e.ld = align_up(e.dev.n, 64);
    std::lock_guard<std::mutex> lk(s->mu);
    s->graphs.push_back(std::move(e));
    *index = (int32_t)s->graphs.size() - 1;
  });
}

int fap_server_submit(fap_server *s, int32_t graph, int32_t sub_id, int64_t *ticket) {
  if (!s || !ticket) {
    fap_hip::set_error("fap_server_submit: null argument");
    return FAP_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  if (graph < 0 || graph >= (int32_t)s->graphs.size() || s->stop) {
    fap_hip::set_error("fap_server_submit: unknown graph %d", graph);
    return FAP_ERR_INVALID;
  }
  const GraphEntry &ge = s->graphs[graph];
  const int K = (int)ge.parts.size();
  if (sub_id < 0 || sub_id > K) {
    fap_hip::set_error("fap_server_submit: sub graph %d out of range 0..%d", sub_id, K);
    return FAP_ERR_INVALID;
  }
  Request r;
  r.graph = graph;
  r.sub_id = sub_id;
  r.t_submit = now_s();
  r.cols = ge.dev.n;
  r.ld = ge.ld;
  if (sub_id == 0) {
    r.parts = ge.parts;  // out_row = offset: the n x n layout of solve()
    r.rows = ge.dev.n;
  } else if (ge.g->subGraphByBoundaryGraph()) {
    // Every part (in-part distances, boundary graph), rows of part sub_id only.
    r.parts = ge.parts;
    r.parts[sub_id - 1].out_row = 0;
    r.only_part = sub_id;
    r.rows = r.parts[sub_id - 1].num;
  } else {
    fap_part_desc p = ge.parts[sub_id - 1];
    p.out_row = 0;
    r.parts.push_back(p);
    r.rows = p.num;
  }
  r.bytes = r.rows * r.ld * 4;
  r.scratch = scratch_estimate(r, ge.is_float, ge.dev.n);
  const int64_t tk = s->next_ticket++;
  if (s->submitted == 0) s->t_first = r.t_submit;
  ++s->submitted;
  const bool large = ge.dev.n >= kLargeN;
  s->reqs[tk] = std::move(r);
  (large ? s->pending_large : s->pending).push_back(tk);
  *ticket = tk;
  lk.unlock();
  s->cv_work.notify_all();
  return FAP_OK;
}

int fap_server_wait(fap_server *s, int64_t ticket, double timeout_s, int32_t *state) {
  if (!s || !state) return FAP_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    fap_hip::set_error("fap_server_wait: unknown ticket %lld", (long long)ticket);
    return FAP_ERR_INVALID;
  }
  const auto deadline = std::chrono::steady_clock::now() +
                        std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                            std::chrono::duration<double>(std::max(timeout_s, 0.0)));
  s->cv_done.wait_until(lk, deadline, [&] { return it->second.state >= kDone; });
  *state = it->second.state;
  return FAP_OK;
}

int fap_server_result(fap_server *s, int64_t ticket, void **d_out, int64_t *rows, int64_t *cols,
                      int64_t *ld, int32_t *is_float) {
  if (!s) return FAP_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state != kDone) {
    fap_hip::set_error("fap_server_result: ticket %lld is not done", (long long)ticket);
    return FAP_ERR_STATE;
  }
  const Request &r = it->second;
  if (d_out) *d_out = r.d_out;
  if (rows) *rows = r.rows;
  if (cols) *cols = r.cols;
  if (ld) *ld = r.ld;
  if (is_float) *is_float = s->graphs[r.graph].is_float ? 1 : 0;
  return FAP_OK;
}

int fap_server_release(fap_server *s, int64_t ticket) {
  if (!s) return FAP_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state < kDone) {
    fap_hip::set_error("fap_server_release: ticket %lld is not finished", (long long)ticket);
    return FAP_ERR_STATE;
  }
  if (it->second.d_out) fap_hip_free(it->second.d_out);
  if (it->second.state == kDone) s->reserved -= it->second.bytes;
  s->reqs.erase(it);
  s->cv_work.notify_all();
  return FAP_OK;
}

int fap_server_get_stats(fap_server *s, fap_server_stats *out) {
  if (!s || !out) return FAP_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  fap_server_stats st{};
  st.submitted = s->submitted;
  st.completed = s->completed;
  st.failed = s->failed;
  st.batches = s->batches;
  st.reserved_bytes = s->reserved;
  st.peak_reserved_bytes = s->peak_reserved;
  st.entries = s->entries;
  st.span_s = s->t_last > s->t_first ? s->t_last - s->t_first : 0.0;
  if (!s->latency_ms.empty()) {
    std::vector<double> v = s->latency_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const size_t i = (size_t)std::min<double>(v.size() - 1, std::max(0.0, q * (v.size() - 1) + 0.5));
      return v[i];
    };
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    st.mean_ms = sum / v.size();
  }
  *out = st;
  return FAP_OK;
}

}  // extern "C"
