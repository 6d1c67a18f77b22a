// fapGraph::check_result -- deterministic verifier of a block of result rows
// (interface of fap/fap.h:155-159; the reference's fap::check_ans,
// fap/utils/checkResult.h:38-94, samples 20 sources x 10 columns with a time seed
// and checks distances only).
#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

#include "fap/fap.h"

namespace fap {

namespace {

// FP64 Dijkstra from src over the CSR (adj_size[v] arcs of v from row_offset[v]).
void dijkstra_f64(int src, int n, const int *adj_size, const int *row_offset, const int *col,
                  const float *w, std::vector<double> &d) {
  d.assign(n, std::numeric_limits<double>::infinity());
  d[src] = 0.0;
  typedef std::pair<double, int> P;
  std::priority_queue<P, std::vector<P>, std::greater<P>> q;
  q.push({0.0, src});
  while (!q.empty()) {
    const P x = q.top();
    q.pop();
    if (x.first > d[x.second]) continue;
    const int u = x.second;
    for (int e = row_offset[u]; e < row_offset[u] + adj_size[u]; ++e) {
      const double nd = x.first + (double)w[e];
      if (nd < d[col[e]]) {
        d[col[e]] = nd;
        q.push({nd, col[e]});
      }
    }
  }
}

bool agree(double got, double want) {
  return std::fabs(got - want) <= std::max(1e-5, 1e-6 * std::fabs(want));
}

}  // namespace

bool fapGraph::check_result(float *dist, int *path, int *source, int source_num, int vertexs,
                            int *adj_size, int *row_offset, int *col_val, float *weight,
                            int *graph_id) {
  if (!dist || !source || !adj_size || !row_offset || !col_val || !weight || vertexs <= 0 ||
      source_num <= 0) {
    std::printf("check_result: nothing to check\n");
    return false;
  }
  const int n = vertexs;
  const float unreached = MAXVALUE / 10;
  const int checks = std::min(source_num, 20);
  std::vector<double> d;
  for (int c = 0; c < checks; ++c) {
    // evenly spaced rows, first and last included
    const int r = checks == 1 ? 0 : (int)((int64_t)c * (source_num - 1) / (checks - 1));
    const int ver = source[r];
    if (ver < 0 || ver >= n) {
      std::printf("check_result: source %d of row %d out of range\n", ver, r);
      return false;
    }
    dijkstra_f64(ver, n, adj_size, row_offset, col_val, weight, d);
    const float *row = dist + (int64_t)r * n;
    const int *prow = path ? path + (int64_t)r * n : nullptr;
    for (int v = 0; v < n; ++v) {
      const bool want_unreached = !std::isfinite(d[v]) || d[v] > unreached;
      const bool got_unreached = row[v] > unreached;
      const bool ok_dist = want_unreached ? got_unreached : agree(row[v], d[v]);
      if (!ok_dist) {
        std::printf("check_result: the dist is wrong: source %d -> %d: %f, Dijkstra %f (%s)\n",
                    ver, v, row[v], d[v],
                    graph_id && graph_id[ver] == graph_id[v] ? "same sub graph"
                                                             : "different sub graphs");
        return false;
      }
      if (!prow) continue;
      const int u = prow[v];
      bool ok_path;
      if (v == ver) {
        ok_path = u == ver;
      } else if (want_unreached) {
        ok_path = u == -1;
      } else {
        ok_path = false;
        if (u >= 0 && u < n && std::isfinite(d[u]))
          for (int e = row_offset[u]; e < row_offset[u] + adj_size[u] && !ok_path; ++e)
            ok_path = col_val[e] == v && agree(d[u] + (double)weight[e], d[v]);
      }
      if (!ok_path) {
        std::printf("check_result: the path is wrong: source %d -> %d: predecessor %d\n", ver,
                    v, u);
        return false;
      }
    }
  }
  return true;
}

}  // namespace fap
