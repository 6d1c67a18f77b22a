// Host test double of the device entry points that the host layer (fap.cpp, server.cpp,
// capi_graph.cpp) calls, for the sanitizer builds of that layer: "device" memory is
// host memory, streams and events are host objects (an event completes a few hundred
// microseconds after it is recorded, so the server's completion polling really waits),
// and the solve entry points compute their rows with a host BFS / Dijkstra from the
// rows' sources.  It is linked only into the programs built by tests/sanitize/Makefile;
// the library itself has no host path.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <queue>
#include <thread>
#include <utility>
#include <vector>

#include "fap_hip.h"

namespace {

thread_local char g_error[512] = "";
std::mutex g_mem_mu;
int64_t g_in_use = 0, g_high = 0;
std::atomic<int64_t> g_live{0};

struct HostEvent {
  std::chrono::steady_clock::time_point ready{};
  bool recorded = false;
};

struct HostBatch {
  std::vector<fap_batch_item> items;
  std::vector<std::vector<fap_part_desc>> parts;
};

// Distances (double; +inf unreachable) and predecessors from one source.  Unit
// weights when both weight arrays are null.
void host_sssp(const fap_graph_dev &g, int src, std::vector<double> &d, std::vector<int> &pred) {
  const int n = g.n;
  d.assign(n, std::numeric_limits<double>::infinity());
  pred.assign(n, -1);
  d[src] = 0;
  pred[src] = src;
  typedef std::pair<double, int> P;
  std::priority_queue<P, std::vector<P>, std::greater<P>> q;
  q.push({0.0, src});
  while (!q.empty()) {
    const P x = q.top();
    q.pop();
    const int u = x.second;
    if (x.first > d[u]) continue;
    for (int e = g.d_row_offset[u]; e < g.d_row_offset[u + 1]; ++e) {
      const int v = g.d_col[e];
      const double w = g.d_wf ? (double)g.d_wf[e] : g.d_w ? (double)g.d_w[e] : 1.0;
      if (x.first + w < d[v]) {
        d[v] = x.first + w;
        pred[v] = u;
        q.push({d[v], v});
      }
    }
  }
}

// Rows out_row + i (i < num) of every part: the distances from vertex st2ed[offset + i]
// to every vertex, as int32 (FAP_INF unreachable), or as float32 (1e8 unreachable and
// for distances >= 1e8), with predecessors when path is not null.
int host_solve(const fap_graph_dev *g, const fap_part_desc *parts, int nparts, void *out,
               bool is_float, int32_t *path, int64_t ld) {
  if (!g || (!parts && nparts > 0) || !out || ld < g->n) return FAP_ERR_INVALID;
  std::vector<double> d;
  std::vector<int> pred;
  for (int p = 0; p < nparts; ++p) {
    const fap_part_desc &pd = parts[p];
    if (pd.offset < 0 || pd.num < 0 || (int64_t)pd.offset + pd.num > g->n) return FAP_ERR_INVALID;
    for (int i = 0; i < pd.num; ++i) {
      host_sssp(*g, g->d_st2ed[pd.offset + i], d, pred);
      const int64_t row = (pd.out_row + i) * ld;
      for (int v = 0; v < g->n; ++v) {
        if (is_float) {
          static_cast<float *>(out)[row + v] =
              std::isinf(d[v]) || d[v] >= 1e8 ? 1e8f : (float)d[v];
        } else {
          static_cast<int32_t *>(out)[row + v] =
              std::isinf(d[v]) || d[v] >= FAP_INF ? FAP_INF : (int32_t)d[v];
        }
        if (path) path[row + v] = pred[v];
      }
    }
  }
  return FAP_OK;
}

}  // namespace

namespace fap_hip {
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}
}  // namespace fap_hip

extern "C" {

int fap_hip_init(int) { return FAP_OK; }
const char *fap_hip_last_error(void) { return g_error; }

int fap_hip_malloc(void **p, int64_t bytes) {
  if (!p || bytes < 0) return FAP_ERR_INVALID;
  const int64_t sz = std::max<int64_t>(bytes, 1);
  char *q = static_cast<char *>(std::malloc(sz + 16));
  if (!q) return FAP_ERR_NOMEM;
  std::memcpy(q, &sz, sizeof(sz));
  std::memset(q + 16, 0x5a, sz);  // not zeroed, like device memory
  *p = q + 16;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_in_use += sz;
  g_high = std::max(g_high, g_in_use);
  ++g_live;
  return FAP_OK;
}

int fap_hip_free(void *p) {
  if (!p) return FAP_OK;
  char *q = static_cast<char *>(p) - 16;
  int64_t sz;
  std::memcpy(&sz, q, sizeof(sz));
  std::free(q);
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_in_use -= sz;
  --g_live;
  return FAP_OK;
}

int fap_hip_malloc_placed(void **p, int64_t rows, int64_t ld, int32_t, float *probe_ms) {
  if (probe_ms) *probe_ms = 0.0f;
  return fap_hip_malloc(p, rows * ld * 4);
}

int fap_hip_mem_stats(int64_t *in_use, int64_t *high_water, int64_t *budget) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  if (in_use) *in_use = g_in_use;
  if (high_water) *high_water = g_high;
  if (budget) *budget = 0;
  return FAP_OK;
}

int fap_hip_memcpy_h2d(void *dst, const void *src, int64_t bytes, void *) {
  if (bytes > 0) std::memcpy(dst, src, bytes);
  return FAP_OK;
}

int fap_hip_memcpy_d2h(void *dst, const void *src, int64_t bytes, void *) {
  if (bytes > 0) std::memcpy(dst, src, bytes);
  return FAP_OK;
}

int fap_hip_memcpy_d2d(void *dst, const void *src, int64_t bytes, void *) {
  if (bytes > 0) std::memcpy(dst, src, bytes);
  return FAP_OK;
}

int fap_hip_memcpy2d_d2h_i32(int32_t *dst, int64_t h_ld, const int32_t *src, int64_t d_ld,
                             int64_t rows, int64_t cols, void *) {
  if (rows < 0 || cols < 0 || h_ld < cols || d_ld < cols) return FAP_ERR_INVALID;
  for (int64_t r = 0; r < rows; ++r) std::memcpy(dst + r * h_ld, src + r * d_ld, cols * 4);
  return FAP_OK;
}

int fap_hip_stream_create(void **s) {
  *s = new int(0);
  return FAP_OK;
}
int fap_hip_stream_destroy(void *s) {
  delete static_cast<int *>(s);
  return FAP_OK;
}
int fap_hip_stream_sync(void *) { return FAP_OK; }

int fap_hip_event_create(void **ev) {
  *ev = new HostEvent;
  return FAP_OK;
}
int fap_hip_event_destroy(void *ev) {
  delete static_cast<HostEvent *>(ev);
  return FAP_OK;
}
int fap_hip_event_record(void *ev, void *) {
  HostEvent *e = static_cast<HostEvent *>(ev);
  e->ready = std::chrono::steady_clock::now() + std::chrono::microseconds(300);
  e->recorded = true;
  return FAP_OK;
}
int fap_hip_event_query(void *ev, int32_t *done) {
  const HostEvent *e = static_cast<const HostEvent *>(ev);
  *done = !e->recorded || std::chrono::steady_clock::now() >= e->ready;
  return FAP_OK;
}

int fap_hip_last_stage_times(fap_stage_times *t) {
  if (t) std::memset(t, 0, sizeof(*t));
  return FAP_OK;
}

int fap_hip_solve_parts(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                        int32_t *d_out, int64_t ld_out, int32_t, void *) {
  return host_solve(g, parts, nparts, d_out, false, nullptr, ld_out);
}

// The sweep pair: offsets as the device library computes them; the per-part block is
// solved from the graph directly (the prepared buffers are not read).
int fap_hip_prepare_parts(const fap_graph_dev *, const fap_part_desc *parts, int32_t nparts,
                          int32_t *, int64_t, int32_t *, int64_t *fw_off, int64_t *fw_entries,
                          void *) {
  if (!fw_off || !fw_entries || nparts < 0 || (nparts > 0 && !parts)) return FAP_ERR_INVALID;
  int64_t total = 0;
  for (int p = 0; p < nparts; ++p) {
    if (parts[p].num <= 0) {
      fw_off[p] = -1;
      continue;
    }
    const int64_t S = (parts[p].num + 63) / 64 * 64;
    fw_off[p] = total;
    total += S * S;
  }
  *fw_entries = total;
  return FAP_OK;
}

int fap_hip_merge_prepared(const fap_graph_dev *g, const fap_part_desc *part, const int32_t *,
                           const int32_t *, int64_t, int64_t, int32_t *d_out, int64_t ld_out,
                           void *) {
  return host_solve(g, part, 1, d_out, false, nullptr, ld_out);
}

int fap_hip_count_reached(const fap_graph_dev *, const fap_part_desc *, int32_t nparts,
                          const int32_t *, int64_t, int32_t *counts, void *) {
  for (int p = 0; p < nparts; ++p) counts[p] = 0;
  return FAP_OK;
}

int fap_hip_solve_parts_f32(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                            float *d_out, int64_t ld_out, int32_t, void *) {
  return host_solve(g, parts, nparts, d_out, true, nullptr, ld_out);
}

int fap_hip_solve_parts_path(const fap_graph_dev *g, const int32_t *, const int32_t *,
                             const int32_t *, const fap_part_desc *parts, int32_t nparts,
                             int32_t *d_out, int32_t *d_path, int64_t ld_out, void *) {
  if (!d_path) return FAP_ERR_INVALID;
  return host_solve(g, parts, nparts, d_out, false, d_path, ld_out);
}

int fap_hip_batch_create(fap_batch **b, const fap_batch_item *items, int32_t nitems, int32_t) {
  if (!b || (!items && nitems > 0) || nitems < 0) return FAP_ERR_INVALID;
  HostBatch *h = new HostBatch;
  for (int i = 0; i < nitems; ++i) {
    h->items.push_back(items[i]);
    h->parts.emplace_back(items[i].parts, items[i].parts + items[i].nparts);
  }
  *b = reinterpret_cast<fap_batch *>(h);
  return FAP_OK;
}

// One item: every part, or only the part it.only_part names (its out_row is the block's
// first row; the other parts write nothing).
static int host_item(const fap_batch_item &it, const fap_part_desc *parts, int nparts) {
  if (it.only_part > 0) {
    for (int k = 0; k < nparts; ++k)
      if (parts[k].part_id == it.only_part)
        return host_solve(&it.g, parts + k, 1, it.d_out, it.is_float != 0, nullptr, it.ld_out);
    return FAP_ERR_INVALID;
  }
  return host_solve(&it.g, parts, nparts, it.d_out, it.is_float != 0, nullptr, it.ld_out);
}

int fap_hip_batch_run(fap_batch *b, void *) {
  HostBatch *h = reinterpret_cast<HostBatch *>(b);
  for (size_t i = 0; i < h->items.size(); ++i) {
    const int s = host_item(h->items[i], h->parts[i].data(), (int)h->parts[i].size());
    if (s != FAP_OK) return s;
  }
  return FAP_OK;
}

int fap_hip_batch_destroy(fap_batch *b) {
  delete reinterpret_cast<HostBatch *>(b);
  return FAP_OK;
}

int fap_hip_solve_batch(const fap_batch_item *items, int32_t nitems, int32_t, void *) {
  for (int i = 0; i < nitems; ++i) {
    const int s = host_item(items[i], items[i].parts, items[i].nparts);
    if (s != FAP_OK) return s;
  }
  return FAP_OK;
}

// Allocations still live (for the leak check at the end of a program).
int64_t host_backend_live_allocations(void) { return g_live.load(); }

}  // extern "C"
