// C entry points over the reference's own CPU kernels, compiled from the header
// sources where they lie (-I <reference root>), with neither WITH_CUDA nor
// WITH_HIP defined, so every seam takes its CPU branch.  Test infrastructure
// only: the product never links this.
#include "fap/kernel/batch_sssp.h"
#include "fap/kernel/floyd.h"
#include "fap/kernel/minplus.h"
#include "fap/utils/util_preprocess.h"
#include "fap/utils/readMatFile.h"

extern "C" {

void ref_dijkstra_path(int u, int n, int *adj_size, int *row_offset, int *col, float *w,
                       float *dist, int *path) {
  fap::dijkstra_path(u, n, adj_size, row_offset, col, w, dist, path);
}

void ref_batched_sssp_path(int *src, int nsrc, int n, int m, int *adj_size, int *row_offset,
                           int *col, float *w, float *dist, int *path) {
  fap::batched_sssp_path(src, nsrc, n, m, adj_size, row_offset, col, w, dist, path);
}

void ref_floyd_cpu_path(int num, float *mat, int *path) { fap::floyd_CPU_path(num, mat, path); }

void ref_min_plus_cpu_advanced(float *a, float *b, int *bpath, float *c, int *cpath, int M, int N,
                               int K) {
  fap::min_plus_path_CPU_advanced(a, b, bpath, c, cpath, M, N, K);
}

// findBoundry + sort_and_encode on a given 1-based partition.
void ref_layout(int K, int n, int *graph_id, int *adj_size, int *row_offset, int *col, float *w,
                int *is_boundary, int *blk_num, int *blk_off, int *bdy_num, int *bdy_off,
                int *st2ed, int *ed2st) {
  std::unordered_map<int, std::vector<int>> block_ver, block_bdy;
  std::unordered_map<int, int> isb;
  for (int i = 0; i < n; ++i) block_ver[graph_id[i]].push_back(i);
  fap::findBoundry(K, n, graph_id, adj_size, row_offset, col, w, block_ver, block_bdy, isb);
  fap::sort_and_encode(K, n, graph_id, isb, blk_num, blk_off, bdy_num, bdy_off, block_ver,
                       block_bdy, st2ed, ed2st);
  for (int i = 0; i < n; ++i) is_boundary[i] = isb.count(i) ? isb[i] : 0;
}

// readVerEdges + readMatFile: file is "<name>", opened as graph/<kind>/<name>.mtx
// relative to the working directory.  Returns n and m through the pointers; the
// arrays must hold n (+1) and m entries (call with NULL arrays first for sizes).
int ref_read_mtx(const char *file, int directed, int weighted, int *n_out, int *m_out,
                 int *adj_size, int *row_offset, int *col, float *w) {
  int n = 0, m = 0;
  fap::readVerEdges(n, m, file, directed != 0, weighted != 0);
  *n_out = n;
  *m_out = m;
  if (!adj_size) return 0;
  fap::readMatFile(n, m, adj_size, row_offset, col, w, file, directed != 0, weighted != 0);
  return 0;
}

}  // extern "C"
