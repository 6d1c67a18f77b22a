// C wrapper of fap::fapGraph for FFI callers (declared in fap_hip.h).
#include <cstring>
#include <new>

#include "fap/fap.h"
#include "graph.h"

#include "capi_internal.h"

extern "C" {

int fap_graph_create_from_file(fap_graph **out, const char *file, int directed, int weighted,
                               int32_t K, const char *partitioner) {
  if (!out || !file) {
    fap_hip::set_error("fap_graph_create_from_file: null argument");
    return FAP_ERR_INVALID;
  }
  *out = nullptr;
  return guarded([&] {
    std::chrono::time_point<std::chrono::high_resolution_clock> t;
    auto *h = new fap_graph;
    try {
      h->g = new fap::fapGraph(file, directed != 0, weighted != 0, K,
                               partitioner && *partitioner ? partitioner : "metis", true, t);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int fap_mtx_read(const char *file, int directed, int weighted, int32_t *n, int64_t *m,
                 int32_t *row_offset, int32_t *col, float *weight) {
  if (!file || !n || !m) {
    fap_hip::set_error("fap_mtx_read: null argument");
    return FAP_ERR_INVALID;
  }
  return guarded([&] {
    fap::Csr c = fap::read_mtx(file, directed != 0, weighted != 0);
    *n = c.n;
    *m = c.m();
    if (row_offset) memcpy(row_offset, c.row_offset.data(), 4 * c.row_offset.size());
    if (col) memcpy(col, c.col_val.data(), 4 * c.col_val.size());
    if (weight) memcpy(weight, c.weight.data(), 4 * c.weight.size());
  });
}

int fap_graph_create_from_csr(fap_graph **out, int32_t n, int64_t m, const int32_t *row_offset,
                              const int32_t *col_val, const float *weight, int directed,
                              int32_t K) {
  if (!out || !row_offset || (m > 0 && !col_val) || n < 1 || m < 0) {
    fap_hip::set_error("fap_graph_create_from_csr: bad argument");
    return FAP_ERR_INVALID;
  }
  *out = nullptr;
  return guarded([&] {
    std::vector<int32_t> ro(row_offset, row_offset + n + 1), cv(col_val, col_val + m);
    std::vector<float> w;
    if (weight) w.assign(weight, weight + m);
    auto *h = new fap_graph;
    try {
      h->g = new fap::fapGraph(n, m, std::move(ro), std::move(cv), std::move(w), K, directed != 0);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int fap_graph_generate(fap_graph **out, int32_t kind, int32_t W, int32_t H, int32_t n_target,
                       double p_a, double p_b, uint64_t seed, int32_t K) {
  if (!out) {
    fap_hip::set_error("fap_graph_generate: null argument");
    return FAP_ERR_INVALID;
  }
  *out = nullptr;
  return guarded([&] {
    fap::GenGraph gg;
    if (kind == 0) gg = fap::gen_road_grid(W, H, p_a, p_b, seed);
    else if (kind == 1) gg = fap::gen_road_like(n_target, p_a, p_b, seed);
    else if (kind == 2) gg = fap::gen_knn(n_target, W, H, seed);
    else throw fap::FapError(FAP_ERR_INVALID, "unknown generator kind");
    fap::Csr c = fap::csr_from_arcs(gg.n, gg.arcs, gg.directed);
    const int64_t m = c.m();
    std::vector<float> w;
    if (gg.weighted) w = std::move(c.weight);
    auto *h = new fap_graph;
    try {
      h->g = new fap::fapGraph(c.n, m, std::move(c.row_offset), std::move(c.col_val), std::move(w),
                               K, gg.directed);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int fap_graph_destroy(fap_graph *h) {
  if (!h) return FAP_OK;
  return guarded([&] {
    delete h->g;
    delete h;
  });
}

int fap_layout_for_partition(int32_t n, const int32_t *row_offset, const int32_t *col, int32_t K,
                             const int32_t *graph_id, int32_t *st2ed, int32_t *ed2st,
                             int32_t *blk_num, int32_t *blk_off, int32_t *bdy_num) {
  if (n < 1 || K < 1 || !row_offset || !graph_id || !st2ed || !ed2st || !blk_num || !blk_off ||
      !bdy_num || (row_offset[n] > 0 && !col)) {
    fap_hip::set_error("fap_layout_for_partition: bad argument");
    return FAP_ERR_INVALID;
  }
  return guarded([&] {
    fap::Csr g;
    g.n = n;
    g.row_offset.assign(row_offset, row_offset + n + 1);
    g.col_val.assign(col, col + row_offset[n]);
    std::vector<int> gid(graph_id, graph_id + n);
    for (int v : gid)
      if (v < 1 || v > K) throw fap::FapError(FAP_ERR_INVALID, "part id out of [1, K]");
    std::unordered_map<int, std::vector<int>> bv, bb;
    for (int i = 0; i < n; ++i) bv[gid[i]].push_back(i);
    std::vector<int> isb, num, off, bdy, boff, s2e, e2s;
    fap::find_boundary(K, g, gid, isb, bb);
    fap::sort_and_encode(K, n, gid, isb, bv, bb, num, off, bdy, boff, s2e, e2s);
    memcpy(st2ed, s2e.data(), 4 * (size_t)n);
    memcpy(ed2st, e2s.data(), 4 * (size_t)n);
    memcpy(blk_num, num.data(), 4 * (size_t)(K + 1));
    memcpy(blk_off, off.data(), 4 * (size_t)(K + 1));
    memcpy(bdy_num, bdy.data(), 4 * (size_t)(K + 1));
  });
}

#define FAP_HANDLE(h)                                          \
  if (!(h) || !(h)->g) {                                       \
    fap_hip::set_error("%s: null graph handle", __func__);     \
    return FAP_ERR_INVALID;                                    \
  }

int fap_graph_set_partition(fap_graph *h, const int32_t *graph_id) {
  FAP_HANDLE(h);
  if (!graph_id) {
    fap_hip::set_error("fap_graph_set_partition: null partition");
    return FAP_ERR_INVALID;
  }
  return guarded([&] {
    h->g->setPartition(std::vector<int>(graph_id, graph_id + h->g->num_vertexs));
  });
}

int fap_graph_precondition(fap_graph *h) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->preCondition(); });
}

int fap_graph_is_split(fap_graph *h, int *is_split) {
  FAP_HANDLE(h);
  return guarded([&] {
    if (is_split) *is_split = h->g->isSplit() ? 1 : 0;
  });
}

int fap_graph_set_mode(fap_graph *h, int32_t mode) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->setMode(mode); });
}

int fap_graph_set_subgraph_sweep(fap_graph *h, int32_t after_calls, int32_t *ready) {
  FAP_HANDLE(h);
  return guarded([&] {
    if (after_calls >= 0) h->g->setSubGraphSweep(after_calls);
    if (ready) *ready = h->g->subGraphSweepState();
  });
}

int fap_graph_solve(fap_graph *h) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->solve(); });
}

int fap_graph_solve_subgraph(fap_graph *h, int32_t id) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->solveSubGraph(id); });
}

int fap_graph_solve_ex(fap_graph *h, int32_t is_path_needed) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->solve(is_path_needed != 0); });
}

int fap_graph_solve_subgraph_ex(fap_graph *h, int32_t id, int32_t is_path_needed) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->solveSubGraph(id, is_path_needed != 0); });
}

int fap_graph_get_path_rows(fap_graph *h, int64_t row0, int64_t rows, int32_t *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getPathRowsI32(row0, rows, h_out); });
}

int fap_graph_get_subgraph_path_rows(fap_graph *h, int32_t id, int64_t row0, int64_t rows,
                                     int32_t *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getSubGraphPathRowsI32(id, row0, rows, h_out); });
}

int fap_graph_sizes(fap_graph *h, int32_t *n, int64_t *m, int32_t *K) {
  FAP_HANDLE(h);
  if (n) *n = h->g->num_vertexs;
  if (m) *m = h->g->num_edges;
  if (K) *K = h->g->numParts();
  return FAP_OK;
}

int fap_graph_get_csr(fap_graph *h, int32_t *row_offset, int32_t *col, int32_t *weight) {
  FAP_HANDLE(h);
  const fap::fapGraph &g = *h->g;
  if (row_offset) memcpy(row_offset, g.row_offset.data(), 4 * g.row_offset.size());
  if (col) memcpy(col, g.col_val.data(), 4 * g.col_val.size());
  if (weight)
    for (size_t i = 0; i < g.weight.size(); ++i) weight[i] = (int32_t)g.weight[i];
  return FAP_OK;
}

int fap_graph_get_weights_f32(fap_graph *h, float *weight) {
  FAP_HANDLE(h);
  const fap::fapGraph &g = *h->g;
  if (weight) memcpy(weight, g.weight.data(), 4 * g.weight.size());
  return FAP_OK;
}

int fap_graph_get_layout(fap_graph *h, int32_t *graph_id, int32_t *st2ed, int32_t *ed2st,
                         int32_t *blk_num, int32_t *blk_off, int32_t *bdy_num) {
  FAP_HANDLE(h);
  return guarded([&] {
    fap::fapGraph &g = *h->g;
    const std::vector<int> gid = g.getGraphId();
    if (graph_id) memcpy(graph_id, gid.data(), 4 * gid.size());
    if (g.sortedToVertex().empty() && (st2ed || ed2st || blk_num || blk_off || bdy_num))
      throw fap::FapError(FAP_ERR_STATE, "layout requested before preCondition()");
    if (st2ed) memcpy(st2ed, g.sortedToVertex().data(), 4 * g.sortedToVertex().size());
    if (ed2st) memcpy(ed2st, g.vertexToSorted().data(), 4 * g.vertexToSorted().size());
    if (blk_num) memcpy(blk_num, g.blockNum().data(), 4 * g.blockNum().size());
    if (blk_off) memcpy(blk_off, g.blockOffset().data(), 4 * g.blockOffset().size());
    if (bdy_num) memcpy(bdy_num, g.blockBoundary().data(), 4 * g.blockBoundary().size());
  });
}

int fap_graph_get_dist_rows(fap_graph *h, int64_t row0, int64_t rows, int32_t *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getDistanceRowsI32(row0, rows, h_out); });
}

int fap_graph_get_subgraph_dist(fap_graph *h, int32_t id, int32_t *h_out) {
  FAP_HANDLE(h);
  return guarded([&] {
    std::vector<int32_t> d = h->g->getSubGraphDistanceI32(id);
    memcpy(h_out, d.data(), 4 * d.size());
  });
}

int fap_graph_get_subgraph_rows(fap_graph *h, int32_t id, int64_t row0, int64_t rows,
                                int32_t *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getSubGraphDistanceRowsI32(id, row0, rows, h_out); });
}

int fap_graph_is_float(fap_graph *h, int32_t *is_float) {
  FAP_HANDLE(h);
  return guarded([&] { *is_float = h->g->isFloat() ? 1 : 0; });
}

int fap_graph_boundary_by_sssp(fap_graph *h, int32_t *by_sssp) {
  FAP_HANDLE(h);
  return guarded([&] { *by_sssp = h->g->boundaryBySssp() ? 1 : 0; });
}

int fap_graph_get_dist_rows_f32(fap_graph *h, int64_t row0, int64_t rows, float *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getDistanceRowsF32(row0, rows, h_out); });
}

int fap_graph_get_subgraph_rows_f32(fap_graph *h, int32_t id, int64_t row0, int64_t rows,
                                    float *h_out) {
  FAP_HANDLE(h);
  return guarded([&] { h->g->getSubGraphDistanceRowsF32(id, row0, rows, h_out); });
}

int fap_graph_device_dist(fap_graph *h, int32_t **d_dist, int64_t *ld) {
  FAP_HANDLE(h);
  if (d_dist) *d_dist = const_cast<int32_t *>(h->g->deviceDistance(ld));
  return FAP_OK;
}

int fap_graph_times(fap_graph *h, double *t_part, double *t_sssp, double *t_floyd,
                    double *t_minplus, double *t_data) {
  FAP_HANDLE(h);
  fap::fapGraph &g = *h->g;
  if (t_part) *t_part = g.getTimePartitioning();
  if (t_sssp) *t_sssp = g.getTimeSSSP();
  if (t_floyd) *t_floyd = g.getTimeFloyd();
  if (t_minplus) *t_minplus = g.getTimeMinPlus();
  if (t_data) *t_data = g.getTimeData();
  return FAP_OK;
}

int fap_graph_edge_cut(fap_graph *h, double *edge_cut) {
  FAP_HANDLE(h);
  if (edge_cut) *edge_cut = h->g->getEdgeCut();
  return FAP_OK;
}

}  // extern "C"

namespace {
// Batch items of a list of graphs; parts holds the part arrays the items point to.
void batch_items(fap_graph **graphs, int32_t count, std::vector<fap_batch_item> &items,
                 std::vector<std::vector<fap_part_desc>> &parts) {
  if (count < 0 || (count > 0 && !graphs))
    throw fap::FapError(FAP_ERR_INVALID, "graph list is null");
  items.resize(count);
  parts.resize(count);
  for (int32_t i = 0; i < count; ++i) {
    if (!graphs[i] || !graphs[i]->g) throw fap::FapError(FAP_ERR_INVALID, "null graph in list");
    fap::fapGraph &g = *graphs[i]->g;
    parts[i] = g.partList();
    fap_batch_item &it = items[i];
    it.g = g.deviceView();
    it.parts = parts[i].data();
    it.nparts = (int32_t)parts[i].size();
    it.is_float = g.isFloat() ? 1 : 0;
    it.d_out = g.resultBuffer(&it.ld_out);
    it.flags = g.boundaryBySssp() ? FAP_ITEM_BOUNDARY_SSSP : 0;
  }
}
}  // namespace

int fap_graph_batch_create_ex(fap_batch **b, fap_graph **graphs, int32_t count, int32_t flags,
                              const int32_t *item_flags) {
  if (!b) return FAP_ERR_INVALID;
  return guarded([&] {
    std::vector<fap_batch_item> items;
    std::vector<std::vector<fap_part_desc>> parts;
    batch_items(graphs, count, items, parts);
    if (item_flags)
      for (int32_t i = 0; i < count; ++i) items[i].flags |= item_flags[i];
    const int s = fap_hip_batch_create(b, items.data(), count, flags);
    if (s != FAP_OK) throw fap::FapError(s, fap_hip_last_error());
  });
}

int fap_graph_batch_create(fap_batch **b, fap_graph **graphs, int32_t count, int32_t flags) {
  return fap_graph_batch_create_ex(b, graphs, count, flags, nullptr);
}

int fap_graph_solve_batch(fap_graph **graphs, int32_t count, int32_t mode) {
  return guarded([&] {
    std::vector<fap_batch_item> items;
    std::vector<std::vector<fap_part_desc>> parts;
    batch_items(graphs, count, items, parts);
    int s = fap_hip_solve_batch(items.data(), count, mode, nullptr);
    if (s == FAP_OK) s = fap_hip_stream_sync(nullptr);
    if (s != FAP_OK) throw fap::FapError(s, fap_hip_last_error());
  });
}
