// Boundary detection and the sorted part layout
// (semantics of fap/utils/util_preprocess.h:32-106).
#include <algorithm>

#include "graph.h"

namespace fap {

void find_boundary(int K, const Csr &g, const std::vector<int> &graph_id,
                   std::vector<int> &is_boundary,
                   std::unordered_map<int, std::vector<int>> &block_boundary) {
  const int n = g.n;
  is_boundary.assign(n, 0);
  // Every vertex of every part 1..K is visited once (util_preprocess.h:37-51).
  for (int u = 0; u < n; ++u) {
    if (graph_id[u] < 1 || graph_id[u] > K) continue;
    for (int e = g.row_offset[u]; e < g.row_offset[u + 1]; ++e) {
      const int v = g.col_val[e];
      if (graph_id[v] != graph_id[u]) {
        is_boundary[u] = 1;
        is_boundary[v] = 1;
      }
    }
  }
  block_boundary.clear();
  for (int i = 0; i < n; ++i)
    if (is_boundary[i]) block_boundary[graph_id[i]].push_back(i);
}

namespace {
struct DenseNode {
  int v;
  int GraphId;
  int isBound;
};
// Part ascending; inside a part boundary vertices first (util_preprocess.h:66-71).
bool dense_less(const DenseNode &a, const DenseNode &b) {
  if (a.GraphId == b.GraphId) return a.isBound > b.isBound;
  return a.GraphId < b.GraphId;
}
}  // namespace

void sort_and_encode(int K, int n, const std::vector<int> &graph_id,
                     const std::vector<int> &is_boundary,
                     const std::unordered_map<int, std::vector<int>> &block_ver,
                     const std::unordered_map<int, std::vector<int>> &block_boundary,
                     std::vector<int> &blk_num, std::vector<int> &blk_off,
                     std::vector<int> &bdy_num, std::vector<int> &bdy_off,
                     std::vector<int> &st2ed, std::vector<int> &ed2st) {
  std::vector<DenseNode> ver(n);
  for (int i = 0; i < n; ++i) ver[i] = DenseNode{i, graph_id[i], is_boundary[i]};
  // The same (unstable) std::sort over the same input order as the reference, so
  // the order inside each (part, boundary) group is reproduced as well.
  std::vector<DenseNode> sorted(ver);
  std::sort(sorted.begin(), sorted.end(), dense_less);
  st2ed.assign(n, 0);
  ed2st.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    st2ed[i] = sorted[i].v;
    ed2st[sorted[i].v] = i;
  }
  blk_num.assign(K + 1, 0);
  blk_off.assign(K + 1, 0);
  bdy_num.assign(K + 1, 0);
  bdy_off.assign(K + 1, 0);
  int cnt = 0;
  for (int p = 1; p <= K; ++p) {
    auto it = block_ver.find(p);
    auto ib = block_boundary.find(p);
    const int nv = it == block_ver.end() ? 0 : (int)it->second.size();
    blk_off[p] = cnt;
    blk_num[p] = nv;
    bdy_off[p] = cnt;
    bdy_num[p] = ib == block_boundary.end() ? 0 : (int)ib->second.size();
    cnt += nv;
  }
}

}  // namespace fap
