// Host-side graph utilities of libfap_hip: .mtx input, CSR construction,
// synthetic generators, k-way partitioning and the sorted part layout.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace fap {

struct Arc {
  int32_t a, b;  // 0-based endpoints
  float w;
};

// CSR in the reference's layout: adj_size[n], row_offset[n+1], col_val[m],
// weight[m] (fap/fap.h:89-94).
struct Csr {
  int32_t n = 0;
  std::vector<int32_t> adj_size, row_offset, col_val;
  std::vector<float> weight;
  int64_t m() const { return (int64_t)col_val.size(); }
};

// Builds the CSR the way readMatFile does (fap/utils/readMatFile.h:92-162):
// every input arc is pushed at the head of its source's list (undirected input
// adds a->b then b->a), so the arcs of one vertex come out in reverse input order.
Csr csr_from_arcs(int32_t n, const std::vector<Arc> &arcs, bool directed);

// Reads "<name>" from graph/<(un)weight>-<(un)directed>/<name>.mtx relative to the
// working directory (fap/utils/readMatFile.h:43-51), or a path ending in ".mtx" as
// given.  '%' lines are skipped, the header is "n n m", ids are 1-based,
// unweighted files are read as pairs with weight 1, weights are |w| as float.
// Throws FapError(FAP_ERR_IO) when the file is missing or malformed.
Csr read_mtx(const std::string &file, bool directed, bool weighted);
std::string mtx_path(const std::string &file, bool directed, bool weighted);

// Synthetic graphs (std::mt19937_64, identical on every machine).
//  road grid: W x H grid, each grid edge dropped with probability p_drop, each
//             down-right diagonal added with probability p_diag, largest component.
//  road-like: L x L grid (L grown from 339 in steps of 16 until the largest
//             component holds n_target vertices), each grid edge kept with
//             probability p_keep, round(chord_rate * L^2) random long chords,
//             largest component, trimmed to exactly n_target vertices in BFS order
//             from the vertex nearest the grid centre.
//  k-NN:      n points uniform in [0,1)^2, directed arcs to the k nearest
//             neighbours, integer weights uniform in [1, wmax].
// Vertex ids follow grid / point order; arcs are listed in generation order.
struct GenGraph {
  int32_t n = 0;
  std::vector<Arc> arcs;
  bool directed = false;
  bool weighted = false;
};
GenGraph gen_road_grid(int W, int H, double p_drop, double p_diag, uint64_t seed);
GenGraph gen_road_like(int n_target, double p_keep, double chord_rate, uint64_t seed);
GenGraph gen_knn(int n, int k, int wmax, uint64_t seed);

// k-way partition with METIS k-way semantics (fap/utils/readIdFile.h:200-283):
// directed graphs are symmetrised and self-loops dropped; returns 1-based part
// ids; parts hold n/K vertices within a 3% imbalance.  Multilevel recursive
// bisection (heavy-edge matching, greedy growing, Fiduccia-Mattheyses refinement).
std::vector<int> partition_kway(const Csr &g, int K, uint64_t seed, double *edge_cut);

// findBoundry (fap/utils/util_preprocess.h:32-58): both endpoints of every arc
// whose endpoints lie in different parts are boundary vertices.
void find_boundary(int K, const Csr &g, const std::vector<int> &graph_id,
                   std::vector<int> &is_boundary,
                   std::unordered_map<int, std::vector<int>> &block_boundary);

// sort_and_encode (fap/utils/util_preprocess.h:73-106): std::sort of
// (vertex, part, boundary) by part ascending then boundary first; st2ed[pos] =
// vertex, ed2st[vertex] = pos, per-part counts and offsets (index 0 unused).
void sort_and_encode(int K, int n, const std::vector<int> &graph_id,
                     const std::vector<int> &is_boundary,
                     const std::unordered_map<int, std::vector<int>> &block_ver,
                     const std::unordered_map<int, std::vector<int>> &block_boundary,
                     std::vector<int> &blk_num, std::vector<int> &blk_off,
                     std::vector<int> &bdy_num, std::vector<int> &bdy_off,
                     std::vector<int> &st2ed, std::vector<int> &ed2st);

}  // namespace fap
