// fap/fap.h -- drop-in C++ API of the MI355X Fast-APSP (class fap::fapGraph).
//
// Mirrors fap/fap.h:25-162 of TomerHaham/FastAPSP: the (file, directed, weighted,
// K[, partitioner, version, endTimeIo]) constructors, preCondition, isSplit,
// solve, solveSubGraph and the getters.  Differences a caller can see:
//  * distances are computed as int32 on the GPU (FAP_INF = INT32_MAX/2 for
//    unreachable); getAllDistance / getSubGraphDistance still return float with
//    the reference's MAXVALUE = 1e8 for unreachable, and *I32 variants return
//    the int32 values without conversion;
//  * the result of solve() stays resident in HBM (n x ld int32); host vectors are
//    produced only by the getters;
//  * errors throw fap::FapError (carrying the C-ABI status) instead of exit();
//  * solveSubGraph recomputes on every call (no process-wide "already processed"
//    skip, fap/fap.cpp:522-528);
//  * getDistanceFromOnePoint uses the sorted position of the vertex (the
//    reference indexes C_BlockVer_offset by vertex id, fap/fap.cpp:584-585);
//  * check_result is deterministic: it checks whole rows of evenly spaced sources
//    against an FP64 Dijkstra, and the predecessors when a path matrix is given
//    (the reference samples 20 sources x 10 columns with a time seed,
//    fap/utils/checkResult.h:38-94);
//  * getMemCons* report the library's device memory in use (GB) averaged over the
//    parts solved, from its own allocation accounting (the reference sums
//    cudaMemGetInfo readings per part, fap/fap.cpp:242-243,298-299,358);
//  * getDistanceP2P / getPathP2P have (u, v) forms; the argument-free forms the
//    reference declares but leaves unimplemented (fap/fap.cpp:669-675) throw.
// Predecessor (path) matrices: solve(true) / solveSubGraph(id, true) produce them
// for integer-weight graphs, equal to the reference CPU path's bit for bit (same
// tie-breaking); float-weight graphs throw FapError(FAP_ERR_STATE) for them, graphs
// with a zero-weight arc FapError(FAP_ERR_INVALID) (the reference's choice among
// equal distances then follows its heap order).
#pragma once

#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "fap_hip.h"

namespace fap {

extern const float MAXVALUE;  // 1e8, fap/utils/parameter.h:26

class FapError : public std::runtime_error {
 public:
  FapError(int status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

class fapGraph {
 protected:
  std::string input_graph;
  std::string partitioner;
  const int32_t max_vertexs_num_limit = 20000;  // fap/fap.h:35
  bool is_split = false;
  bool is_path_needed = false;
  bool version = true;
  bool preconditioned = false;
  int32_t mode = FAP_MODE_BATCHED;
  double time_partitioning = 0;
  double time_sssp = 0;
  double time_floyd = 0;
  double time_min_plus = 0;
  double time_data_transformation = 0;
  double edge_cut = 0;
  double mem_cons_sssp = 0, mem_cons_floyd = 0, mem_cons_min_plus = 0;  // GB x parts

  int32_t current_subgraph_id = 0;

  int32_t K = 0;
  std::vector<int> graph_id;
  std::unordered_map<int, std::vector<int>> BlockVer;
  std::unordered_map<int, std::vector<int>> BlockBoundary;
  std::vector<int> isBoundry;
  std::vector<int> st2ed;
  std::vector<int> ed2st;
  std::vector<int> C_BlockVer_num;
  std::vector<int> C_BlockVer_offset;
  std::vector<int> C_BlockBdy_num;
  std::vector<int> C_BlockBdy_offset;
  int32_t max_subgraph_vertexs = 0;

  // Device residency (owned).
  int32_t *d_row_offset = nullptr, *d_col = nullptr, *d_w = nullptr;
  int32_t *d_graph_id = nullptr, *d_st2ed = nullptr, *d_ed2st = nullptr;
  int32_t *d_comp = nullptr;      // component labels (symmetric graphs only)
  float *d_wf = nullptr;          // float32 weights (float-weight graphs)
  bool float_weights = false;     // some weight is not an integer below 2^24
  int32_t *d_dist = nullptr;      // n x ld_dist, result of solve()
  int32_t *d_sub_dist = nullptr;  // s_p x ld_dist, result of solveSubGraph()
  int32_t *d_path = nullptr;      // n x ld_dist predecessors of solve(true)
  int32_t *d_sub_path = nullptr;  // s_p x ld_dist predecessors of solveSubGraph(id, true)
  int32_t *d_rro = nullptr, *d_rcol = nullptr, *d_rw = nullptr;  // reverse CSR (paths)
  bool have_path = false, have_sub_path = false;
  int64_t ld_dist = 0;
  int64_t sub_rows = 0;

  void init_from_csr(bool directed, bool weighted, int32_t K, const std::string &partitioner);
  void upload();
  void release();
  void add_stage_times(int32_t parts);
  void alloc_result(int32_t **d, const char *what);
  // Parallel in-part arcs whose later arc is heavier: boundary rows by SSSP (the
  // reference's semantics for such graphs); set by preCondition.
  bool boundary_by_sssp = false;
  // Float weights and boundary sources that reach few vertices (sampled by BFS in
  // preCondition; e.g. directed k-NN graphs): one SSSP per boundary vertex is then
  // cheaper than closing the boundary graph, which full float solves do otherwise.
  bool sssp_cheap = false;
  int32_t solve_mode() const {
    return mode | (boundary_by_sssp || (float_weights && sssp_cheap) ? FAP_MODE_FLAG_SSSP : 0);
  }
  // solveSubGraph of a weighted graph (int weights or float) whose boundary sources
  // reach far, in the batched form: its boundary rows come from the boundary graph
  // (one label-correcting SSSP per boundary vertex took 0.15-0.65 s on weighted road
  // graphs of 31k-62k vertices).
  bool subgraph_by_boundary_graph() const {
    return (d_w || float_weights) && !boundary_by_sssp && !sssp_cheap &&
           (mode & 0xFF) == FAP_MODE_BATCHED && !(mode & FAP_MODE_FLAG_SSSP);
  }
  // solveSubGraph sweep: from the sweep_after-th distance-only call on, every part's
  // boundary rows and Floyd-Warshall matrix come from one batched pass (kept until the
  // layout changes) and each call only writes its block.  sweep_reached[p]: the columns
  // part p's boundary rows reach when its block takes the sparse product, -1 for the
  // dense product; sweep_sparse: some part takes the sparse product.
  int32_t sweep_after = 3;
  int32_t sweep_calls = 0;
  bool sweep_ready = false, sweep_failed = false, sweep_sparse = false;
  // Full-solve form of the sweep (weighted graphs whose blocks come from the boundary
  // graph): the blocks are rows of the resident solve() result.
  bool sweep_full = false;
  int32_t *d_sweep_fw = nullptr, *d_sweep_rows = nullptr;
  std::vector<int64_t> sweep_fw_off, sweep_row0, sweep_reached;
  bool prepare_sweep(const fap_graph_dev &g);
  void free_sweep();
  void solve_placed(const fap_graph_dev &g, const std::vector<fap_part_desc> &parts);
  void copy_rows(const int32_t *base, int64_t nrows, int64_t row0, int64_t rows, void *out) const;
  void require_int(const char *what) const;
  void upload_reverse();
  void run_path(const fap_part_desc *parts, int32_t nparts, int32_t *d_out, int32_t *d_p);

 public:
  int32_t num_vertexs = 0;
  int32_t num_edges = 0;
  std::vector<int32_t> adj_size;
  std::vector<int32_t> row_offset;
  std::vector<int32_t> col_val;
  std::vector<float> weight;
  bool directed = false;
  bool weighted = false;

  // (file, directed, weighted, K) -- README.md:44; partitioner "metis".
  fapGraph(std::string input_graph, bool directed, bool weighted, int32_t K);
  // fap/fap.h:100-101.  `version` is accepted and ignored: every solve runs on the GPU.
  fapGraph(std::string input_graph, bool directed, bool weighted, int32_t K,
           std::string partitioner, bool version,
           std::chrono::time_point<std::chrono::high_resolution_clock> &endTimeIo);
  // fap/fap.h:102-104 (the reference leaves it empty, fap/fap.cpp:94-98).
  fapGraph(int32_t num_vertexs, int64_t num_edges, std::vector<int32_t> row_offset,
           std::vector<int32_t> col_val, std::vector<float> weight, int32_t K,
           bool directed = false);
  ~fapGraph();
  fapGraph(const fapGraph &) = delete;
  fapGraph &operator=(const fapGraph &) = delete;

  // Use a caller's partition (1-based ids, size n) instead of the built-in one.
  void setPartition(const std::vector<int> &graph_id);
  void preCondition();
  bool isSplit();
  // Parallel arcs inside a part whose later arc is heavier (set by preCondition): the
  // boundary rows then come from SSSP, as in the reference.
  bool boundaryBySssp() const { return boundary_by_sssp; }
  // Mode flags of this graph's full solves (FAP_MODE_FLAG_SSSP when the boundary rows
  // come from SSSP: boundaryBySssp(), or float weights with a small sampled reach).
  int32_t solveModeFlags() const { return solve_mode() & ~0xFF; }
  // solveSubGraph blocks from the boundary graph (weighted graphs whose boundary
  // sources reach far) rather than one SSSP per boundary vertex.
  bool subGraphByBoundaryGraph() const { return subgraph_by_boundary_graph(); }
  std::vector<int> getGraphId();

  // Execution form of solve(): FAP_MODE_BATCHED (default), FAP_MODE_OVERLAP or
  // FAP_MODE_PLAIN (identical results).
  void setMode(int32_t mode);

  int32_t solve(bool is_path_needed = false);
  int32_t solveSubGraph(int32_t sub_graph_id, bool is_path_needed = false);
  // Distance-only solveSubGraph calls on integer-weight graphs: from the after_calls-th
  // call on (default 3; 0 = never), one batched pass prepares every part's boundary rows
  // and Floyd-Warshall matrix, held in device memory until the layout changes (about
  // 4 x (sum of boundary counts x n + sum of padded s_p^2) bytes, at most 32 GiB), and
  // each call then only writes its block.  Weighted graphs whose blocks come from the
  // boundary graph (subGraphByBoundaryGraph()) instead run one full solve() when its
  // n x n result fits in 32 GiB, and each block is a copy of its rows.  Results are
  // identical either way.
  void setSubGraphSweep(int32_t after_calls);
  // 0 = not prepared, 1 = prepared (dense product), 2 = prepared (sparse product),
  // 3 = blocks taken from a full solve() (weighted graphs whose blocks come from the
  // boundary graph, when the n x n result fits in 32 GiB).
  int32_t subGraphSweepState() const {
    return sweep_full ? 3 : sweep_ready ? (sweep_sparse ? 2 : 1) : 0;
  }

  std::vector<int32_t> getMapping();
  std::vector<float> getAllDistance();
  std::vector<int32_t> getAllDistanceI32();
  std::vector<int32_t> getAllPath();
  int32_t getCurrentSubGraphId();
  std::vector<int32_t> getSubGraphIndex(int32_t sub_graph_id);
  std::vector<float> getSubGraphDistance(int32_t sub_graph_id);
  std::vector<int32_t> getSubGraphDistanceI32(int32_t sub_graph_id);
  // Rows [row0, row0 + rows) of the last solveSubGraph block (int32) into out.
  void getSubGraphDistanceRowsI32(int32_t sub_graph_id, int64_t row0, int64_t rows,
                                  int32_t *out);
  std::vector<int32_t> getSubGraphPath(int32_t sub_graph_id);
  std::vector<float> getDistanceFromOnePoint(int32_t vertex_id);
  std::vector<int32_t> getPathFromOnePoint(int32_t vertex_id);
  // Rows [row0, row0 + rows) of the predecessors of solve(true) / of the last
  // solveSubGraph(id, true) block into out (n entries per row).
  void getPathRowsI32(int64_t row0, int64_t rows, int32_t *out);
  void getSubGraphPathRowsI32(int32_t sub_graph_id, int64_t row0, int64_t rows, int32_t *out);
  // Rows [row0, row0 + rows) of the solve() result, without a full copy: int32
  // (integer-weight graphs only) or float32 with MAXVALUE for unreachable (any graph).
  void getDistanceRowsI32(int64_t row0, int64_t rows, int32_t *out);
  void getDistanceRowsF32(int64_t row0, int64_t rows, float *out);
  void getSubGraphDistanceRowsF32(int32_t sub_graph_id, int64_t row0, int64_t rows, float *out);
  // True when the graph has a weight that is not an integer below 2^24: solves then
  // run the float path (FP64 arithmetic, float32 results) and the int32 getters throw.
  bool isFloat() const { return float_weights; }
  // Device pointer / row stride of the resident solve() result.
  const int32_t *deviceDistance(int64_t *ld) const;

  // For many-graph batches (fap_hip_batch_*): the device view (uploaded on first
  // use), the parts 1..K with out_row = offset, and the n x n result buffer of
  // solve() (int32, or float32 bits for float-weight graphs; allocated on first use).
  fap_graph_dev deviceView();
  std::vector<fap_part_desc> partList() const;
  void *resultBuffer(int64_t *ld);

  // fap/fap.h:136-138.  d(u, v) and the vertices of the chosen shortest path from u
  // to v (u first, v last; empty when v is unreachable), from the solve() result
  // (predecessors: solve(true)) or, when u lies in it, the last solved sub graph.
  float getDistanceP2P(int32_t u, int32_t v);
  std::vector<int32_t> getPathP2P(int32_t u, int32_t v);
  // The reference's argument-free forms name no pair: they throw FapError(FAP_ERR_STATE).
  float getDistanceP2P();
  std::vector<int32_t> getPathP2P();
  double getTimeSSSP();
  double getTimeFloyd();
  double getTimeMinPlus();
  double getTimePartitioning();
  double getTimeData();
  int getBoundaryNum();
  int getInnerNum();
  // fap/fap.h:143-145: device memory (GB) in use while the parts were solved, averaged
  // over K as the reference does.
  double getMemConsSSSP();
  double getMemConsFloyd();
  double getMemConsMinPlus();
  double getEdgeCut();

  // fap/fap.h:151: every entry of the solve() result, one line each, in the
  // reference's format (sorted row order, MAXVALUE for unreachable).
  void printResalts();

  // fap/fap.h:155-159: verifies a block of distance rows (row i = source source[i],
  // `vertexs` entries each) and, when path is not NULL, its predecessors, against an
  // FP64 Dijkstra over the given CSR.  Checks min(source_num, 20) evenly spaced rows
  // in full.  Distances must agree within max(1e-5, 1e-6 |d|); both above MAXVALUE/10
  // counts as unreachable.  A predecessor u of v must be an in-neighbour with
  // d(u) + w(u, v) = d(v) (same tolerance), the source itself on the diagonal and -1
  // where v is unreachable.  Prints the first mismatch; false when there is one or
  // when there is no row to check.  graph_id (may be NULL) only labels the message.
  bool check_result(float *dist, int *path, int *source, int source_num, int vertexs,
                    int *adj_size, int *row_offset, int *col_val, float *weight, int *graph_id);

  // Layout accessors (sorted order): C_BlockVer_num / offset, C_BlockBdy_num.
  const std::vector<int> &blockNum() const { return C_BlockVer_num; }
  const std::vector<int> &blockOffset() const { return C_BlockVer_offset; }
  const std::vector<int> &blockBoundary() const { return C_BlockBdy_num; }
  const std::vector<int> &sortedToVertex() const { return st2ed; }
  const std::vector<int> &vertexToSorted() const { return ed2st; }
  int32_t numParts() const { return K; }
};

}  // namespace fap
