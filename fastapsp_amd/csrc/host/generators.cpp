// Seeded synthetic graph generators (road grid, road-like percolated grid, k-NN).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <random>
#include <set>

#include "fap/fap.h"
#include "graph.h"

namespace fap {

namespace {

// Uniform double in [0, 1) from the top 53 bits (same on every platform).
inline double u01(std::mt19937_64 &rng) { return (double)(rng() >> 11) * 0x1.0p-53; }

struct Dsu {
  std::vector<int> p, sz;
  explicit Dsu(int n) : p(n), sz(n, 1) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) {
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
  }
  void unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (sz[a] < sz[b]) std::swap(a, b);
    p[b] = a;
    sz[a] += sz[b];
  }
};

// Marks the largest connected component (ties: the one holding the smallest id).
std::vector<char> largest_component(int n, const std::vector<Arc> &edges) {
  Dsu d(n);
  for (const Arc &e : edges) d.unite(e.a, e.b);
  int best = -1, best_sz = 0;
  for (int v = 0; v < n; ++v) {
    const int r = d.find(v);
    if (d.sz[r] > best_sz) {
      best_sz = d.sz[r];
      best = r;
    }
  }
  std::vector<char> keep(n, 0);
  for (int v = 0; v < n; ++v) keep[v] = d.find(v) == best;
  return keep;
}

// Keeps the marked vertices (relabelled in increasing id order) and the edges
// between them, in input order.
GenGraph induce(int n, const std::vector<Arc> &edges, const std::vector<char> &keep) {
  std::vector<int> id(n, -1);
  int cnt = 0;
  for (int v = 0; v < n; ++v)
    if (keep[v]) id[v] = cnt++;
  GenGraph g;
  g.n = cnt;
  for (const Arc &e : edges)
    if (keep[e.a] && keep[e.b]) g.arcs.push_back(Arc{id[e.a], id[e.b], e.w});
  return g;
}

std::vector<Arc> percolated_grid(int L, double p_keep, double chord_rate, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::vector<Arc> edges;
  for (int r = 0; r < L; ++r)
    for (int c = 0; c < L; ++c) {
      const int v = r * L + c;
      if (c + 1 < L && u01(rng) < p_keep) edges.push_back(Arc{v, v + 1, 1.0f});
      if (r + 1 < L && u01(rng) < p_keep) edges.push_back(Arc{v, v + L, 1.0f});
    }
  const int64_t N = (int64_t)L * L;
  const int64_t chords = std::llround(chord_rate * (double)N);
  std::set<std::pair<int, int>> seen;
  for (int64_t i = 0; i < chords; ++i) {
    int a = (int)(rng() % (uint64_t)N), b = (int)(rng() % (uint64_t)N);
    if (a == b) continue;
    if (a > b) std::swap(a, b);
    const int ra = a / L, ca = a % L, rb = b / L, cb = b % L;
    if (std::abs(ra - rb) + std::abs(ca - cb) == 1) continue;  // already a grid pair
    if (!seen.insert({a, b}).second) continue;
    edges.push_back(Arc{a, b, 1.0f});
  }
  return edges;
}

}  // namespace

GenGraph gen_road_grid(int W, int H, double p_drop, double p_diag, uint64_t seed) {
  if (W < 1 || H < 1) throw FapError(FAP_ERR_INVALID, "road grid: W and H must be >= 1");
  std::mt19937_64 rng(seed);
  std::vector<Arc> edges;
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) {
      const int v = r * W + c;
      if (c + 1 < W && u01(rng) >= p_drop) edges.push_back(Arc{v, v + 1, 1.0f});
      if (r + 1 < H && u01(rng) >= p_drop) edges.push_back(Arc{v, v + W, 1.0f});
      if (c + 1 < W && r + 1 < H && u01(rng) < p_diag) edges.push_back(Arc{v, v + W + 1, 1.0f});
    }
  const int n = W * H;
  GenGraph g = induce(n, edges, largest_component(n, edges));
  g.directed = false;
  g.weighted = false;
  return g;
}

GenGraph gen_road_like(int n_target, double p_keep, double chord_rate, uint64_t seed) {
  if (n_target < 1) throw FapError(FAP_ERR_INVALID, "road-like: n_target must be >= 1");
  int L = 339;
  std::vector<Arc> edges;
  std::vector<char> keep;
  for (;; L += 16) {
    edges = percolated_grid(L, p_keep, chord_rate, seed);
    keep = largest_component(L * L, edges);
    const int64_t cnt = std::count(keep.begin(), keep.end(), 1);
    if (cnt >= n_target) break;
    if (L > 8192) throw FapError(FAP_ERR_INVALID, "road-like: cannot reach n_target");
  }
  const int N = L * L;
  // Adjacency of the grid graph in insertion order, for the BFS trim.
  Csr adj = csr_from_arcs(N, edges, false);
  int start = -1;
  double best = 1e300;
  const double mid = (L - 1) / 2.0;
  for (int v = 0; v < N; ++v) {
    if (!keep[v]) continue;
    const double dr = v / L - mid, dc = v % L - mid;
    const double d2 = dr * dr + dc * dc;
    if (d2 < best) {
      best = d2;
      start = v;
    }
  }
  std::vector<char> take(N, 0);
  std::queue<int> q;
  q.push(start);
  take[start] = 1;
  int taken = 1;
  while (!q.empty() && taken < n_target) {
    const int u = q.front();
    q.pop();
    for (int e = adj.row_offset[u]; e < adj.row_offset[u + 1] && taken < n_target; ++e) {
      const int v = adj.col_val[e];
      if (!take[v]) {
        take[v] = 1;
        ++taken;
        q.push(v);
      }
    }
  }
  GenGraph g = induce(N, edges, take);
  g.directed = false;
  g.weighted = false;
  return g;
}

GenGraph gen_knn(int n, int k, int wmax, uint64_t seed) {
  if (n < 2 || k < 1 || k >= n || wmax < 1)
    throw FapError(FAP_ERR_INVALID, "k-NN: need n >= 2, 1 <= k < n, wmax >= 1");
  std::mt19937_64 rng(seed);
  std::vector<double> x(n), y(n);
  for (int i = 0; i < n; ++i) {
    x[i] = u01(rng);
    y[i] = u01(rng);
  }
  const int G = std::max(1, (int)std::sqrt(n / 2.0));
  std::vector<std::vector<int>> cell((size_t)G * G);
  auto cell_of = [&](double v) { return std::min(G - 1, (int)(v * G)); };
  for (int i = 0; i < n; ++i) cell[(size_t)cell_of(y[i]) * G + cell_of(x[i])].push_back(i);
  GenGraph g;
  g.n = n;
  g.directed = true;
  g.weighted = true;
  std::vector<std::pair<double, int>> cand;
  for (int i = 0; i < n; ++i) {
    const int cx = cell_of(x[i]), cy = cell_of(y[i]);
    // Grow the search ring until k candidates lie within the ring's guaranteed radius.
    for (int ring = 1;; ++ring) {
      cand.clear();
      for (int yy = std::max(0, cy - ring); yy <= std::min(G - 1, cy + ring); ++yy)
        for (int xx = std::max(0, cx - ring); xx <= std::min(G - 1, cx + ring); ++xx)
          for (int j : cell[(size_t)yy * G + xx]) {
            if (j == i) continue;
            const double dx = x[j] - x[i], dy = y[j] - y[i];
            cand.push_back({dx * dx + dy * dy, j});
          }
      const double radius = (double)ring / G;
      if ((int)cand.size() >= k) {
        std::partial_sort(cand.begin(), cand.begin() + k, cand.end());
        if (cand[k - 1].first <= radius * radius || ring >= G) break;
      } else if (ring >= G) {
        break;
      }
    }
    for (int t = 0; t < k && t < (int)cand.size(); ++t) g.arcs.push_back(Arc{i, cand[t].second, 0.0f});
  }
  for (Arc &e : g.arcs) e.w = (float)(1 + (int)(rng() % (uint64_t)wmax));
  return g;
}

}  // namespace fap
