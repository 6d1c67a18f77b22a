// Concurrency check of the request server (fap_server_*): client threads submit full
// APSP and sub-graph requests on several resident graphs (integer and float weights,
// one graph large enough for the pipeline lane), wait, check every result row against
// a host BFS / Dijkstra, and release; a second server with a small budget makes
// admission wait for releases and fails a request larger than the budget.  Built with
// -fsanitize=thread and with -fsanitize=address,undefined by tests/sanitize/Makefile,
// linked with the host test double of the device entry points.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <queue>
#include <random>
#include <thread>
#include <utility>
#include <vector>

#include "fap_hip.h"

extern "C" int64_t host_backend_live_allocations(void);

namespace {

#define CHECK(cond, ...)                        \
  do {                                          \
    if (!(cond)) {                              \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);        \
      std::fprintf(stderr, "\n");               \
      std::exit(1);                             \
    }                                           \
  } while (0)

struct Host {
  fap_graph *g = nullptr;
  int n = 0, K = 0, is_float = 0;
  std::vector<int32_t> ro, col, w, st2ed, ed2st, num, off, bdy;
  std::vector<float> wf;
};

Host make_graph(int kind, int W, int H, int n_target, double pa, double pb, uint64_t seed,
                int K, bool float_weights) {
  Host h;
  fap_graph *g0 = nullptr;
  CHECK(fap_graph_generate(&g0, kind, W, H, n_target, pa, pb, seed, K) == FAP_OK, "generate: %s",
        fap_hip_last_error());
  int32_t n, k;
  int64_t m;
  CHECK(fap_graph_sizes(g0, &n, &m, &k) == FAP_OK, "sizes");
  h.ro.resize(n + 1);
  h.col.resize(m);
  h.w.resize(m);
  CHECK(fap_graph_get_csr(g0, h.ro.data(), h.col.data(), h.w.data()) == FAP_OK, "csr");
  if (float_weights) {  // symmetric float weights in [0.5, 10) from the arc's end points
    std::vector<float> wf(m);
    for (int u = 0; u < n; ++u)
      for (int e = h.ro[u]; e < h.ro[u + 1]; ++e) {
        const uint64_t a = std::min(u, h.col[e]), b = std::max(u, h.col[e]);
        wf[e] = 0.5f + (float)((a * 2654435761u + b * 40503u) % 9500) / 1000.0f;
      }
    fap_graph *g1 = nullptr;
    CHECK(fap_graph_create_from_csr(&g1, n, m, h.ro.data(), h.col.data(), wf.data(), 0, K) ==
              FAP_OK,
          "from csr: %s", fap_hip_last_error());
    fap_graph_destroy(g0);
    g0 = g1;
    h.wf = wf;
  }
  CHECK(fap_graph_precondition(g0) == FAP_OK, "precondition: %s", fap_hip_last_error());
  CHECK(fap_graph_sizes(g0, &n, &m, &k) == FAP_OK, "sizes");
  h.g = g0;
  h.n = n;
  h.K = k;
  CHECK(fap_graph_is_float(g0, &h.is_float) == FAP_OK, "is_float");
  CHECK(fap_graph_get_csr(g0, h.ro.data(), h.col.data(), h.w.data()) == FAP_OK, "csr");
  std::vector<int32_t> gid(n);
  h.st2ed.resize(n);
  h.ed2st.resize(n);
  h.num.resize(k + 1);
  h.off.resize(k + 1);
  h.bdy.resize(k + 1);
  CHECK(fap_graph_get_layout(g0, gid.data(), h.st2ed.data(), h.ed2st.data(), h.num.data(),
                             h.off.data(), h.bdy.data()) == FAP_OK,
        "layout");
  return h;
}

// Distances from src: int (FAP_INF unreachable) or float (1e8 unreachable).
std::vector<double> reference_row(const Host &h, int src) {
  std::vector<double> d(h.n, std::numeric_limits<double>::infinity());
  d[src] = 0;
  typedef std::pair<double, int> P;
  std::priority_queue<P, std::vector<P>, std::greater<P>> q;
  q.push({0, src});
  while (!q.empty()) {
    P x = q.top();
    q.pop();
    if (x.first > d[x.second]) continue;
    for (int e = h.ro[x.second]; e < h.ro[x.second + 1]; ++e) {
      const double w = h.is_float ? (double)h.wf[e] : (double)h.w[e];
      if (x.first + w < d[h.col[e]]) {
        d[h.col[e]] = x.first + w;
        q.push({d[h.col[e]], h.col[e]});
      }
    }
  }
  return d;
}

void check_rows(fap_server *s, int64_t t, const Host &h, int sub, int nrows_check,
                std::mt19937 &rng) {
  void *p = nullptr;
  int64_t rows, cols, ld;
  int32_t fl;
  CHECK(fap_server_result(s, t, &p, &rows, &cols, &ld, &fl) == FAP_OK, "result");
  const int first = sub > 0 ? h.off[sub] : 0;
  CHECK(rows == (sub > 0 ? h.num[sub] : h.n) && cols == h.n && fl == h.is_float,
        "shape %lld x %lld", (long long)rows, (long long)cols);
  for (int c = 0; c < nrows_check && rows > 0; ++c) {
    const int r = c == 0 ? 0 : (int)(rng() % rows);
    const std::vector<double> want = reference_row(h, h.st2ed[first + r]);
    for (int v = 0; v < h.n; ++v) {
      if (h.is_float) {
        const float got = static_cast<const float *>(p)[r * ld + v];
        const float exp = std::isinf(want[v]) ? 1e8f : (float)want[v];
        CHECK(got == exp, "ticket %lld row %d col %d: %g != %g", (long long)t, r, v, got, exp);
      } else {
        const int32_t got = static_cast<const int32_t *>(p)[r * ld + v];
        const int32_t exp = std::isinf(want[v]) ? FAP_INF : (int32_t)want[v];
        CHECK(got == exp, "ticket %lld row %d col %d: %d != %d", (long long)t, r, v, got, exp);
      }
    }
  }
}

void run_clients(fap_server *s, const std::vector<Host> &pool, const std::vector<int> &idx,
                 int threads, int per_thread) {
  std::vector<std::thread> ts;
  for (int c = 0; c < threads; ++c)
    ts.emplace_back([&, c] {
      std::mt19937 rng(100 + c);
      std::vector<std::pair<int64_t, std::pair<int, int>>> mine;
      for (int i = 0; i < per_thread; ++i) {
        const int gi = (int)(rng() % pool.size());
        // Full APSP only on the small graphs; sub-graphs anywhere.
        const int sub = (pool[gi].n <= 2500 && rng() % 3 == 0) ? 0 : 1 + (int)(rng() % pool[gi].K);
        int64_t t = 0;
        CHECK(fap_server_submit(s, idx[gi], sub, &t) == FAP_OK, "submit: %s", fap_hip_last_error());
        mine.push_back({t, {gi, sub}});
        if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
      }
      for (auto &x : mine) {
        int32_t st = -1;
        CHECK(fap_server_wait(s, x.first, 60.0, &st) == FAP_OK && st == FAP_REQ_DONE,
              "wait ticket %lld state %d", (long long)x.first, st);
        check_rows(s, x.first, pool[x.second.first], x.second.second, 2, rng);
        CHECK(fap_server_release(s, x.first) == FAP_OK, "release");
      }
    });
  for (auto &t : ts) t.join();
}

}  // namespace

int main() {
  std::vector<Host> pool;
  pool.push_back(make_graph(1, 0, 0, 1500, 0.55, 1e-3, 11, 6, false));   // road-like, unit
  pool.push_back(make_graph(1, 0, 0, 2200, 0.55, 1e-3, 12, 8, true));    // road-like, float
  pool.push_back(make_graph(2, 3, 50, 1200, 0, 0, 13, 5, false));        // kNN, int weights
  pool.push_back(make_graph(1, 0, 0, 9000, 0.55, 1e-3, 14, 24, false));  // pipeline lane
  {
    fap_server *s = nullptr;
    CHECK(fap_server_create(&s, (int64_t)8 << 30, 4) == FAP_OK, "create");
    std::vector<int> idx;
    for (const Host &h : pool) {
      int32_t i = -1;
      CHECK(fap_server_add_graph(s, h.g, &i) == FAP_OK, "add graph");
      idx.push_back(i);
    }
    run_clients(s, pool, idx, 4, 12);
    fap_server_stats st{};
    CHECK(fap_server_get_stats(s, &st) == FAP_OK, "stats");
    CHECK(st.submitted == 48 && st.completed == 48 && st.failed == 0 && st.reserved_bytes == 0,
          "stats: submitted %lld completed %lld failed %lld reserved %lld", (long long)st.submitted,
          (long long)st.completed, (long long)st.failed, (long long)st.reserved_bytes);
    std::printf("server: %lld requests in %lld batches, p50 %.2f ms, peak %.1f MB\n",
                (long long)st.completed, (long long)st.batches, st.p50_ms,
                st.peak_reserved_bytes / 1e6);
    CHECK(fap_server_destroy(s) == FAP_OK, "destroy");
  }
  {
    // A budget of about two full results of the 1500-vertex graph: admission waits for
    // releases; a full APSP of the 9000-vertex graph exceeds it and fails.
    const int64_t budget = (int64_t)2 * 1500 * 1536 * 4 + ((int64_t)8 << 20);
    fap_server *s = nullptr;
    CHECK(fap_server_create(&s, budget, 2) == FAP_OK, "create");
    std::vector<int> idx;
    for (const Host &h : pool) {
      int32_t i = -1;
      CHECK(fap_server_add_graph(s, h.g, &i) == FAP_OK, "add graph");
      idx.push_back(i);
    }
    int64_t big = 0;
    CHECK(fap_server_submit(s, idx[3], 0, &big) == FAP_OK, "submit big");
    std::vector<int64_t> ts(6);
    for (auto &t : ts) CHECK(fap_server_submit(s, idx[0], 0, &t) == FAP_OK, "submit");
    int32_t st = -1;
    CHECK(fap_server_wait(s, big, 60.0, &st) == FAP_OK && st == FAP_REQ_FAILED, "big: state %d", st);
    std::mt19937 rng(7);
    for (int64_t t : ts) {
      CHECK(fap_server_wait(s, t, 60.0, &st) == FAP_OK && st == FAP_REQ_DONE, "state %d", st);
      check_rows(s, t, pool[0], 0, 1, rng);
      CHECK(fap_server_release(s, t) == FAP_OK, "release");
    }
    fap_server_stats ss{};
    CHECK(fap_server_get_stats(s, &ss) == FAP_OK, "stats");
    CHECK(ss.peak_reserved_bytes <= budget && ss.failed == 1 && ss.completed == 6,
          "budget run: peak %lld failed %lld completed %lld", (long long)ss.peak_reserved_bytes,
          (long long)ss.failed, (long long)ss.completed);
    std::printf("budget: peak %.1f MB of %.1f MB\n", ss.peak_reserved_bytes / 1e6, budget / 1e6);
    CHECK(fap_server_destroy(s) == FAP_OK, "destroy");
  }
  for (Host &h : pool) fap_graph_destroy(h.g);
  CHECK(host_backend_live_allocations() == 0, "%lld device allocations leaked",
        (long long)host_backend_live_allocations());
  std::printf("server_load ok\n");
  return 0;
}
