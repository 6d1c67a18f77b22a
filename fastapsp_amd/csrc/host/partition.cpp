// k-way graph partitioner with METIS k-way semantics.
//
// The reference calls METIS_PartGraphKway with default options on the CSR
// (symmetrised for directed graphs, fap/utils/readIdFile.h:200-283) and stores
// 1-based part ids.  METIS is not available here, so this is an independent
// multilevel recursive-bisection partitioner:
//   coarsen   heavy-edge matching in random order until <= 96 vertices remain;
//   initial   greedy graph growing from several random seeds, each refined,
//             best cut kept;
//   refine    Fiduccia-Mattheyses passes (max-gain moves with rollback to the
//             best balanced prefix) at every level while uncoarsening;
//   k-way     recursive bisection with part-count-proportional targets.
// Distances never depend on the partition; only the boundary sizes (and so the
// min-plus work) do.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <random>

#include "fap/fap.h"
#include "graph.h"

namespace fap {

namespace {

struct WGraph {
  int n = 0;
  std::vector<int> xadj{0}, adj, ew, vw;
  int64_t total_vw = 0;
  int max_vw = 1;
};

WGraph symmetrise(const Csr &g) {
  const int n = g.n;
  std::vector<std::vector<int>> nb(n);
  for (int u = 0; u < n; ++u)
    for (int e = g.row_offset[u]; e < g.row_offset[u + 1]; ++e) {
      const int v = g.col_val[e];
      if (v == u) continue;
      nb[u].push_back(v);
      nb[v].push_back(u);
    }
  WGraph w;
  w.n = n;
  w.xadj.assign(n + 1, 0);
  w.vw.assign(n, 1);
  w.total_vw = n;
  for (int u = 0; u < n; ++u) {
    auto &l = nb[u];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    for (int v : l) {
      w.adj.push_back(v);
      w.ew.push_back(1);
    }
    w.xadj[u + 1] = (int)w.adj.size();
    std::vector<int>().swap(l);
  }
  return w;
}

WGraph coarsen(const WGraph &g, std::mt19937_64 &rng, std::vector<int> &cmap) {
  const int n = g.n;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::shuffle(order.begin(), order.end(), rng);
  const int cap = std::max<int64_t>(2, (int64_t)(1.5 * g.total_vw / std::max(n / 8, 1)));
  std::vector<int> match(n, -1);
  for (int v : order) {
    if (match[v] != -1) continue;
    int best = -1, bw = -1;
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adj[e];
      if (match[u] != -1 || g.vw[u] + g.vw[v] > cap) continue;
      if (g.ew[e] > bw) {
        bw = g.ew[e];
        best = u;
      }
    }
    match[v] = best >= 0 ? best : v;
    if (best >= 0) match[best] = v;
  }
  cmap.assign(n, -1);
  int nc = 0;
  for (int v = 0; v < n; ++v)
    if (cmap[v] == -1) {
      cmap[v] = nc;
      cmap[match[v]] = nc;
      ++nc;
    }
  WGraph c;
  c.n = nc;
  c.xadj.assign(nc + 1, 0);
  c.vw.assign(nc, 0);
  c.total_vw = g.total_vw;
  std::vector<int> members_first(nc, -1), members_second(nc, -1);
  for (int v = 0; v < n; ++v) {
    const int cv = cmap[v];
    c.vw[cv] += g.vw[v];
    if (members_first[cv] == -1) members_first[cv] = v;
    else members_second[cv] = v;
  }
  c.max_vw = *std::max_element(c.vw.begin(), c.vw.end());
  std::vector<int> slot(nc, -1);
  for (int cv = 0; cv < nc; ++cv) {
    const int start = (int)c.adj.size();
    for (int m = 0; m < 2; ++m) {
      const int v = m == 0 ? members_first[cv] : members_second[cv];
      if (v < 0) continue;
      for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
        const int cu = cmap[g.adj[e]];
        if (cu == cv) continue;
        if (slot[cu] >= start) {
          c.ew[slot[cu]] += g.ew[e];
        } else {
          slot[cu] = (int)c.adj.size();
          c.adj.push_back(cu);
          c.ew.push_back(g.ew[e]);
        }
      }
    }
    c.xadj[cv + 1] = (int)c.adj.size();
  }
  return c;
}

int64_t cut_of(const WGraph &g, const std::vector<char> &side) {
  int64_t cut = 0;
  for (int v = 0; v < g.n; ++v)
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
      if (side[v] != side[g.adj[e]]) cut += g.ew[e];
  return cut / 2;
}

// Fiduccia-Mattheyses refinement of a bisection; tw[s] = target weight of side s.
void fm_refine(const WGraph &g, std::vector<char> &side, const double tw[2], double eps,
               int passes) {
  const int n = g.n;
  if (n < 2) return;
  const double maxw[2] = {tw[0] * (1 + eps) + g.max_vw, tw[1] * (1 + eps) + g.max_vw};
  std::vector<int> gain(n);
  std::vector<char> locked(n);
  std::vector<int> moved;
  for (int pass = 0; pass < passes; ++pass) {
    int64_t w[2] = {0, 0};
    for (int v = 0; v < n; ++v) w[(int)side[v]] += g.vw[v];
    typedef std::pair<int, int> GV;  // (gain, vertex)
    std::priority_queue<GV> heap[2];
    for (int v = 0; v < n; ++v) {
      int ext = 0, in = 0;
      for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
        (side[g.adj[e]] != side[v] ? ext : in) += g.ew[e];
      gain[v] = ext - in;
      locked[v] = 0;
      if (ext > 0) heap[(int)side[v]].push({gain[v], v});
    }

    auto excess = [&](const int64_t *ww) {
      return std::max(0.0, ww[0] - maxw[0]) + std::max(0.0, ww[1] - maxw[1]);
    };
    int64_t cut = cut_of(g, side);
    int64_t best_cut = cut;
    double best_excess = excess(w);
    int best_len = 0;
    moved.clear();
    const int limit = std::max(64, n / 50);
    int since_best = 0;
    while (since_best < limit) {
      // Pick the side to move from: an overweight side first, else the best gain.
      int from = -1;
      for (int s = 0; s < 2; ++s)
        while (!heap[s].empty() && (locked[heap[s].top().second] ||
                                    heap[s].top().first != gain[heap[s].top().second] ||
                                    side[heap[s].top().second] != s))
          heap[s].pop();
      // An overweight side without boundary vertices (its part of the graph is
      // disconnected from the other side) gives up interior vertices instead.
      for (int s = 0; s < 2; ++s)
        if (w[s] > maxw[s] && heap[s].empty())
          for (int v = 0; v < n; ++v)
            if (side[v] == s && !locked[v]) heap[s].push({gain[v], v});
      if (w[0] > maxw[0] && !heap[0].empty()) from = 0;
      else if (w[1] > maxw[1] && !heap[1].empty()) from = 1;
      else {
        int best_gain = INT32_MIN;
        for (int s = 0; s < 2; ++s) {
          if (heap[s].empty()) continue;
          const int v = heap[s].top().second;
          if (w[1 - s] + g.vw[v] > maxw[1 - s]) continue;
          if (heap[s].top().first > best_gain) {
            best_gain = heap[s].top().first;
            from = s;
          }
        }
      }
      if (from < 0) break;
      const int v = heap[from].top().second;
      heap[from].pop();
      const int to = 1 - from;
      side[v] = (char)to;
      locked[v] = 1;
      w[from] -= g.vw[v];
      w[to] += g.vw[v];
      cut -= gain[v];
      moved.push_back(v);
      for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
        const int u = g.adj[e];
        gain[u] += (side[u] == to) ? -2 * g.ew[e] : 2 * g.ew[e];
        if (!locked[u]) heap[(int)side[u]].push({gain[u], u});
      }
      const double ex = excess(w);
      if (ex < best_excess || (ex == best_excess && cut < best_cut)) {
        best_excess = ex;
        best_cut = cut;
        best_len = (int)moved.size();
        since_best = 0;
      } else {
        ++since_best;
      }

    }
    for (int i = (int)moved.size() - 1; i >= best_len; --i) side[moved[i]] ^= 1;
    if (best_len == 0) break;
  }
}

std::vector<char> grow_bisection(const WGraph &g, double target0, std::mt19937_64 &rng) {
  const int n = g.n;
  std::vector<char> side(n, 1);
  std::vector<int> conn(n, 0);
  std::vector<char> in(n, 0);
  typedef std::pair<int, int> GV;
  std::priority_queue<GV> heap;
  int64_t w0 = 0;
  auto add = [&](int v) {
    side[v] = 0;
    in[v] = 1;
    w0 += g.vw[v];
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adj[e];
      if (in[u]) continue;
      conn[u] += g.ew[e];
      heap.push({conn[u], u});
    }
  };
  while (w0 < target0) {
    while (!heap.empty() && (in[heap.top().second] || heap.top().first != conn[heap.top().second]))
      heap.pop();
    int v;
    if (heap.empty()) {
      std::vector<int> rest;
      for (int u = 0; u < n; ++u)
        if (!in[u]) rest.push_back(u);
      if (rest.empty()) break;
      v = rest[rng() % rest.size()];
    } else {
      v = heap.top().second;
      heap.pop();
    }
    if (w0 + g.vw[v] > target0 * 1.02 + g.max_vw && w0 > 0) break;
    add(v);
  }
  return side;
}

std::vector<char> bisect(const WGraph &g0, double frac, std::mt19937_64 &rng) {
  std::vector<WGraph> levels;
  std::vector<std::vector<int>> maps;
  levels.push_back(g0);
  while (levels.back().n > 96) {
    std::vector<int> cmap;
    WGraph c = coarsen(levels.back(), rng, cmap);
    if (c.n > 0.92 * levels.back().n) break;
    maps.push_back(std::move(cmap));
    levels.push_back(std::move(c));
  }
  const double total = (double)g0.total_vw;
  const double tw[2] = {frac * total, (1 - frac) * total};
  const WGraph &cg = levels.back();
  std::vector<char> best;
  int64_t best_cut = INT64_MAX;
  for (int t = 0; t < 8; ++t) {
    std::vector<char> s = grow_bisection(cg, tw[0], rng);
    fm_refine(cg, s, tw, 0.01, 6);
    const int64_t c = cut_of(cg, s);
    if (c < best_cut) {
      best_cut = c;
      best = s;
    }
  }
  std::vector<char> side = best;
  for (int l = (int)levels.size() - 2; l >= 0; --l) {
    const WGraph &fg = levels[l];
    std::vector<char> fs(fg.n);
    for (int v = 0; v < fg.n; ++v) fs[v] = side[maps[l][v]];
    const double eps = l == 0 ? 0.003 : 0.01;
    fm_refine(fg, fs, tw, eps, l == 0 ? 8 : 4);
    side.swap(fs);
  }
  return side;
}

WGraph induced(const WGraph &g, const std::vector<int> &verts, std::vector<int> &local) {
  WGraph s;
  s.n = (int)verts.size();
  s.xadj.assign(s.n + 1, 0);
  s.vw.resize(s.n);
  for (int i = 0; i < s.n; ++i) local[verts[i]] = i;
  for (int i = 0; i < s.n; ++i) {
    const int v = verts[i];
    s.vw[i] = g.vw[v];
    s.total_vw += g.vw[v];
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adj[e];
      if (local[u] >= 0) {
        s.adj.push_back(local[u]);
        s.ew.push_back(g.ew[e]);
      }
    }
    s.xadj[i + 1] = (int)s.adj.size();
  }
  for (int v : verts) local[v] = -1;
  s.max_vw = s.n ? *std::max_element(s.vw.begin(), s.vw.end()) : 1;
  return s;
}

void recurse(const WGraph &g, const std::vector<int> &orig, int k, int first,
             std::vector<int> &graph_id, std::vector<int> &local, std::mt19937_64 &rng) {
  if (k == 1 || g.n == 0) {
    for (int v : orig) graph_id[v] = first;
    return;
  }
  const int k0 = k / 2;
  std::vector<char> side;
  if (g.n == 1) side.assign(1, 0);
  else side = bisect(g, (double)k0 / k, rng);
  std::vector<int> v0, v1;
  for (int v = 0; v < g.n; ++v) (side[v] == 0 ? v0 : v1).push_back(v);
  for (int half = 0; half < 2; ++half) {
    const std::vector<int> &vs = half == 0 ? v0 : v1;
    WGraph sg = induced(g, vs, local);
    std::vector<int> so(vs.size());
    for (size_t i = 0; i < vs.size(); ++i) so[i] = orig[vs[i]];
    recurse(sg, so, half == 0 ? k0 : k - k0, half == 0 ? first : first + k0, graph_id, local, rng);
  }
}

// Greedy k-way balancing and refinement on the original graph (unit vertex
// weights): boundary vertices of parts above max_w move to the adjacent part that
// loses the fewest cut edges and has room; then positive-gain moves that keep
// every part within [1, max_w].  Each candidate is re-validated when popped.
void kway_balance_refine(const WGraph &g, std::vector<int> &part, int K, double eps) {
  const int n = g.n;
  std::vector<int> pw(K + 1, 0);
  for (int v = 0; v < n; ++v) ++pw[part[v]];
  const int max_w = std::max((int)std::ceil((double)n / K), (int)std::floor((double)n / K * (1 + eps)));
  std::vector<int> conn(K + 1, 0);
  std::vector<int> touched;
  // Best destination of v (or -1) and its gain, given the balance rule.
  auto best_move = [&](int v, bool balancing, int *gain_out) {
    const int p = part[v];
    touched.clear();
    int own = 0;
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int q = part[g.adj[e]];
      if (q == p) {
        own += g.ew[e];
        continue;
      }
      if (conn[q] == 0) touched.push_back(q);
      conn[q] += g.ew[e];
    }
    // Balancing prefers a destination with room; otherwise any strictly lighter
    // neighbour part, so weight diffuses away from overfull parts over the passes.
    int best = -1, bg = INT32_MIN;
    bool best_fits = false;
    for (int q : touched) {
      const int gain = conn[q] - own;
      const bool fits = pw[q] + 1 <= max_w;
      bool ok;
      if (balancing) ok = fits || pw[q] + 1 < pw[p];
      else ok = fits && pw[p] > 1 && (gain > 0 || (gain == 0 && pw[q] + 1 < pw[p]));
      if (!ok) continue;
      if (balancing && fits != best_fits ? fits : gain > bg) {
        bg = gain;
        best = q;
        best_fits = fits;
      }
    }
    for (int q : touched) conn[q] = 0;
    *gain_out = bg;
    return best;
  };
  typedef std::pair<int, int> GV;
  for (int pass = 0; pass < 24; ++pass) {
    bool over = false;
    for (int q = 1; q <= K; ++q) over |= pw[q] > max_w;
    int moves = 0;
    if (over) {
      std::priority_queue<GV> heap;
      for (int v = 0; v < n; ++v) {
        if (pw[part[v]] <= max_w) continue;
        int gain;
        if (best_move(v, true, &gain) >= 0) heap.push({gain, v});
      }
      while (!heap.empty()) {
        const int v = heap.top().second;
        heap.pop();
        if (pw[part[v]] <= max_w) continue;
        int gain;
        const int q = best_move(v, true, &gain);
        if (q < 0) continue;
        if (!heap.empty() && gain < heap.top().first) {  // stale: re-queue with its gain now
          heap.push({gain, v});
          continue;
        }
        --pw[part[v]];
        ++pw[q];
        part[v] = q;
        ++moves;
        for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
          const int u = g.adj[e];
          if (pw[part[u]] > max_w && best_move(u, true, &gain) >= 0) heap.push({gain, u});
        }
      }
    }
    // Refinement: positive-gain moves within the balance bound.
    std::priority_queue<GV> heap;
    for (int v = 0; v < n; ++v) {
      int gain;
      if (best_move(v, false, &gain) >= 0) heap.push({gain, v});
    }
    while (!heap.empty()) {
      const int v = heap.top().second;
      heap.pop();
      int gain;
      const int q = best_move(v, false, &gain);
      if (q < 0) continue;
      if (!heap.empty() && gain < heap.top().first) {
        heap.push({gain, v});
        continue;
      }
      --pw[part[v]];
      ++pw[q];
      part[v] = q;
      ++moves;
      for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
        const int u = g.adj[e];
        if (best_move(u, false, &gain) >= 0) heap.push({gain, u});
      }
    }
    if (moves == 0) break;
  }
}

}  // namespace

std::vector<int> partition_kway(const Csr &g, int K, uint64_t seed, double *edge_cut) {
  if (K < 1 || K > g.n) throw FapError(FAP_ERR_INVALID, "partition: need 1 <= K <= n");
  std::mt19937_64 rng(seed);
  WGraph w = symmetrise(g);
  std::vector<int> graph_id(g.n, 1);
  std::vector<int> orig(g.n);
  std::iota(orig.begin(), orig.end(), 0);
  std::vector<int> local(g.n, -1);
  recurse(w, orig, K, 1, graph_id, local, rng);
  if (K > 1) kway_balance_refine(w, graph_id, K, 0.03);
  if (edge_cut) {
    int64_t cut = 0;
    for (int v = 0; v < w.n; ++v)
      for (int e = w.xadj[v]; e < w.xadj[v + 1]; ++e)
        if (graph_id[v] != graph_id[w.adj[e]]) cut += w.ew[e];
    *edge_cut = (double)(cut / 2);
  }
  return graph_id;
}

}  // namespace fap
