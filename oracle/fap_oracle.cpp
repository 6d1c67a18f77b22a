// ORACLE -- test infrastructure only.  CPU restatement of the reference's CPU
// Fast-APSP path (TomerHaham/FastAPSP, no WITH_CUDA / WITH_HIP), in float with
// MAXVALUE = 1e8 exactly as the reference computes it, plus independent exact
// checkers (BFS / integer Dijkstra).  Only tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py may load this library; the product never does.
//
// Pins (see tests/test_oracle.py):
//   orc_dijkstra_path      == reference dijkstra_path (compiled _ref + golden
//                             vectors), BFS closed forms on grids and paths
//   orc_floyd_path         == reference floyd_CPU_path (_ref + golden), textbook
//                             brute force on tiny inputs
//   orc_minplus_advanced   == reference min_plus_path_CPU_advanced (_ref +
//                             golden), closed form min_k (alpha_i + k) + (beta_j + k)
//   orc_layout             == reference findBoundry + sort_and_encode (_ref + golden)
//   orc_submat_build / orc_mat1_build / orc_submat_decode: no compilable reference
//                             (util_Myalgorithm.h includes cuda_runtime.h); pinned
//                             through orc_solve == exact BFS / Dijkstra APSP (a
//                             wrong index in any of them breaks that equality) and
//                             permutation round trips.
//   orc_exact_rows_f64    FP64 Dijkstra: equals orc_exact_rows on integer weights and
//                             a numpy float64 Floyd-Warshall on small float graphs.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <utility>
#include <vector>

namespace {
const float kMaxValue = 1e8f;  // fap/utils/parameter.h:26
const int32_t kInf = 1073741823;
}  // namespace

extern "C" {

// fap/kernel/batch_sssp.h:79-114.  Binary heap keyed by (distance, vertex);
// strict '>' relaxation; path[v] = predecessor, path[u] = u.
void orc_dijkstra_path(int u, int n, const int *adj_size, const int *row_offset, const int *col,
                       const float *w, float *dist, int *path) {
  std::fill_n(dist, n, kMaxValue);
  std::fill_n(path, n, -1);
  dist[u] = 0;
  path[u] = u;
  std::vector<char> done(n, 0);
  typedef std::pair<float, int> PDI;
  std::priority_queue<PDI, std::vector<PDI>, std::greater<PDI>> q;
  q.push({0.0f, u});
  while (!q.empty()) {
    const PDI x = q.top();
    q.pop();
    const int v = x.second;
    const float d = x.first;
    if (done[v]) continue;
    done[v] = 1;
    const int off = row_offset[v];
    for (int i = 0; i < adj_size[v]; ++i) {
      const int nb = col[off + i];
      const float nd = d + w[off + i];
      if (dist[nb] > nd) {
        q.push({nd, nb});
        dist[nb] = nd;
        path[nb] = v;
      }
    }
  }
}

// fap/kernel/batch_sssp.h:232-252 (OpenMP over sources).
void orc_batched_sssp_path(const int *src, int nsrc, int n, const int *adj_size,
                           const int *row_offset, const int *col, const float *w, float *dist,
                           int *path) {
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nsrc; ++i)
    orc_dijkstra_path(src[i], n, adj_size, row_offset, col, w, dist + (int64_t)i * n,
                      path + (int64_t)i * n);
}

// fap/utils/util_Myalgorithm.h:44-84; subMat pre-filled by the caller with
// MAXVALUE / -1 (fap/fap.cpp:207-208).
void orc_submat_build(float *sub, int *sub_path, const float *sub_graph, const int *sub_graph_path,
                      const int *graph_id, int start, int s, int bdy, int n, const int *st2ed,
                      const int *ed2st, const int *adj_size, const int *row_offset,
                      const int *col, const float *w) {
  for (int i = 0; i < s; ++i) {
    const int ver = st2ed[i + start];
    const int off = row_offset[ver];
    for (int j = 0; j < adj_size[ver]; ++j) {
      const int nb = col[off + j];
      if (graph_id[ver] != graph_id[nb]) continue;
      const int idx = ed2st[nb] - start;
      if (idx >= 0 && idx < s) {
        sub[(int64_t)i * s + idx] = w[off + j];
        sub_path[(int64_t)i * s + idx] = ver;
      }
    }
  }
  for (int i = 0; i < bdy; ++i)
    for (int j = 0; j < s; ++j) {
      const int ver = st2ed[j + start];
      sub[(int64_t)i * s + j] = sub_graph[(int64_t)i * n + ver];
      sub_path[(int64_t)i * s + j] = sub_graph_path[(int64_t)i * n + ver];
    }
  for (int i = 0; i < s; ++i) {
    sub[(int64_t)i * s + i] = 0;
    sub_path[(int64_t)i * s + i] = st2ed[i + start];
  }
}

// fap/kernel/floyd.h:38-49.
void orc_floyd_path(int num, float *mat, int *path) {
  for (int k = 0; k < num; ++k)
    for (int i = 0; i < num; ++i) {
      for (int j = 0; j < num; ++j) {
        const float cand = mat[(int64_t)i * num + k] + mat[(int64_t)k * num + j];
        if (mat[(int64_t)i * num + j] > cand) {
          mat[(int64_t)i * num + j] = cand;
          path[(int64_t)i * num + j] = path[(int64_t)k * num + j];
        }
      }
    }
}

// fap/utils/util_Myalgorithm.h:278-285.
void orc_mat1_build(float *mat1, const float *sub, int inner, int bdy, int s) {
  for (int i = 0; i < inner; ++i)
    memcpy(mat1 + (int64_t)i * bdy, sub + (int64_t)(i + bdy) * s, sizeof(float) * bdy);
}

// fap/kernel/minplus.h:51-64 (overwrite semantics, init MAXVALUE, strict '>').
void orc_minplus_advanced(const float *a, const float *b, const int *bpath, float *c, int *cpath,
                          int M, int N, int K) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      float best = kMaxValue;
      int bp = -1;
      for (int k = 0; k < K; ++k) {
        const float cand = a[(int64_t)i * K + k] + b[(int64_t)k * N + j];
        if (best > cand) {
          best = cand;
          bp = bpath[(int64_t)k * N + j];
        }
      }
      c[(int64_t)i * N + j] = best;
      cpath[(int64_t)i * N + j] = bp;
    }
}

// fap/utils/util_Myalgorithm.h:236-247.
void orc_submat_decode(const float *sub, const int *sub_path, float *sub_graph,
                       int *sub_graph_path, int start, int rows, int cols, int n,
                       const int *st2ed) {
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      const int ver = st2ed[j + start];
      sub_graph[(int64_t)i * n + ver] = sub[(int64_t)i * cols + j];
      sub_graph_path[(int64_t)i * n + ver] = sub_path[(int64_t)i * cols + j];
    }
}

// fapGraph::run, CPU branch (fap/fap.cpp:195-208, 400-502): the s x n block of
// part `id` (1-based) into dist / path (row r = source st2ed[offset + r]).
void orc_run_part(int id, int n, const int *adj_size, const int *row_offset, const int *col,
                  const float *w, const int *graph_id, const int *st2ed, const int *ed2st,
                  const int *blk_num, const int *blk_off, const int *bdy_num, float *dist,
                  int *path) {
  const int bdy = bdy_num[id], s = blk_num[id], inner = s - bdy, start = blk_off[id];
  std::vector<float> mat1((size_t)inner * bdy, kMaxValue);
  std::vector<float> sub((size_t)s * s, kMaxValue);
  std::vector<int> sub_path((size_t)s * s, -1);
  orc_batched_sssp_path(st2ed + start, bdy, n, adj_size, row_offset, col, w, dist, path);
  orc_submat_build(sub.data(), sub_path.data(), dist, path, graph_id, start, s, bdy, n, st2ed,
                   ed2st, adj_size, row_offset, col, w);
  orc_floyd_path(s, sub.data(), sub_path.data());
  orc_mat1_build(mat1.data(), sub.data(), inner, bdy, s);
  const int64_t off = (int64_t)bdy * n;
  orc_minplus_advanced(mat1.data(), dist, path, dist + off, path + off, inner, n, bdy);
  orc_submat_decode(sub.data(), sub_path.data(), dist, path, start, s, s, n, st2ed);
}

// Parts [first, last] (1-based) in parallel, each written at rows blk_off[p]
// of dist (n x n) -- the per-part arithmetic is that of orc_run_part.
void orc_solve_parts_parallel(int first, int last, int n, const int *adj_size,
                              const int *row_offset, const int *col, const float *w,
                              const int *graph_id, const int *st2ed, const int *ed2st,
                              const int *blk_num, const int *blk_off, const int *bdy_num,
                              float *dist, int64_t row_base) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = first; p <= last; ++p) {
    std::vector<int> path((size_t)blk_num[p] * n);
    orc_run_part(p, n, adj_size, row_offset, col, w, graph_id, st2ed, ed2st, blk_num, blk_off,
                 bdy_num, dist + ((int64_t)blk_off[p] - row_base) * n, path.data());
  }
}

// The same, keeping the predecessor (path) block of every part beside the distances
// (path rows laid out like dist: row r = source st2ed[r], column = vertex).
void orc_solve_parts_path_parallel(int first, int last, int n, const int *adj_size,
                                   const int *row_offset, const int *col, const float *w,
                                   const int *graph_id, const int *st2ed, const int *ed2st,
                                   const int *blk_num, const int *blk_off, const int *bdy_num,
                                   float *dist, int *path, int64_t row_base) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = first; p <= last; ++p) {
    const int64_t off = ((int64_t)blk_off[p] - row_base) * n;
    orc_run_part(p, n, adj_size, row_offset, col, w, graph_id, st2ed, ed2st, blk_num, blk_off,
                 bdy_num, dist + off, path + off);
  }
}

// float result -> int32 with MAXVALUE -> INT32_MAX / 2 (exact for integers < 2^24).
void orc_to_i32(const float *in, int32_t *out, int64_t count) {
  for (int64_t i = 0; i < count; ++i) out[i] = in[i] >= kMaxValue ? kInf : (int32_t)in[i];
}

// Independent exact distances: BFS (w == NULL) or integer Dijkstra, one row per
// source, int32 with INT32_MAX / 2 for unreachable.  OpenMP over sources.
static void exact_row(int n, const int *row_offset, const int *col, const int *w, int src,
                      int32_t *d, std::vector<int> &q) {
  std::fill_n(d, n, kInf);
  d[src] = 0;
  if (!w) {
    q.resize(n);
    int head = 0, tail = 0;
    q[tail++] = src;
    while (head < tail) {
      const int u = q[head++];
      for (int e = row_offset[u]; e < row_offset[u + 1]; ++e)
        if (d[col[e]] == kInf) {
          d[col[e]] = d[u] + 1;
          q[tail++] = col[e];
        }
    }
  } else {
    typedef std::pair<int64_t, int> P;
    std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
    pq.push({0, src});
    while (!pq.empty()) {
      const P x = pq.top();
      pq.pop();
      if (x.first > d[x.second]) continue;
      for (int e = row_offset[x.second]; e < row_offset[x.second + 1]; ++e) {
        const int64_t nd = x.first + w[e];
        if (nd < d[col[e]]) {
          d[col[e]] = (int32_t)nd;
          pq.push({nd, col[e]});
        }
      }
    }
  }
}

void orc_exact_rows(int n, const int *row_offset, const int *col, const int *w, const int *src,
                    int nsrc, int32_t *out) {
#pragma omp parallel
  {
    std::vector<int> q;
#pragma omp for schedule(dynamic)
    for (int i = 0; i < nsrc; ++i) exact_row(n, row_offset, col, w, src[i], out + (int64_t)i * n, q);
  }
}

// FP64 oracle for float32 weights: Dijkstra in double (w == NULL: unit weights),
// one row per source, +inf for unreachable.  The target of the float path's
// relative-1e-6 parity.  OpenMP over sources.
void orc_exact_rows_f64(int n, const int *row_offset, const int *col, const float *w,
                        const int *src, int nsrc, double *out) {
  const double inf = std::numeric_limits<double>::infinity();
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nsrc; ++i) {
    double *d = out + (int64_t)i * n;
    std::fill_n(d, n, inf);
    d[src[i]] = 0.0;
    typedef std::pair<double, int> P;
    std::priority_queue<P, std::vector<P>, std::greater<P>> q;
    q.push({0.0, src[i]});
    while (!q.empty()) {
      const P x = q.top();
      q.pop();
      if (x.first > d[x.second]) continue;
      for (int e = row_offset[x.second]; e < row_offset[x.second + 1]; ++e) {
        const double nd = x.first + (w ? (double)w[e] : 1.0);
        if (nd < d[col[e]]) {
          d[col[e]] = nd;
          q.push({nd, col[e]});
        }
      }
    }
  }
}

// Streaming checkers for full-size results (nothing n x n is materialised): each
// computes its expected rows per thread and compares them with the rows a caller
// copied out of the device, returning the number of differing rows / entries.
//
// Rows got[r * ld_got .. + n) against exact distances from src[r] (BFS when w is
// NULL, else integer Dijkstra, as orc_exact_rows).  *first_bad = the first
// differing row (or -1).
int64_t orc_compare_exact_rows(int n, const int *row_offset, const int *col, const int *w,
                               const int *src, int64_t nsrc, const int32_t *got, int64_t ld_got,
                               int64_t *first_bad) {
  int64_t bad = 0, first = -1;
#pragma omp parallel
  {
    std::vector<int32_t> row(n);
    std::vector<int> q;
    int64_t my_bad = 0, my_first = -1;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < nsrc; ++i) {
      exact_row(n, row_offset, col, w, src[i], row.data(), q);
      if (std::memcmp(row.data(), got + i * ld_got, sizeof(int32_t) * n) != 0) {
        ++my_bad;
        if (my_first < 0 || i < my_first) my_first = i;
      }
    }
#pragma omp critical
    {
      bad += my_bad;
      if (my_first >= 0 && (first < 0 || my_first < first)) first = my_first;
    }
  }
  if (first_bad) *first_bad = first;
  return bad;
}

// Parts [first, last] (1-based) computed by orc_run_part (the reference CPU
// path's arithmetic, in float with MAXVALUE) and compared, after the float ->
// int32 mapping of orc_to_i32, with got: part p's row r is got row
// blk_off[p] + r - row_base (stride ld_got).  Returns the number of differing
// entries; *first_bad = the first differing got row (or -1).
int64_t orc_compare_parts(int first, int last, int n, const int *adj_size,
                          const int *row_offset, const int *col, const float *w,
                          const int *graph_id, const int *st2ed, const int *ed2st,
                          const int *blk_num, const int *blk_off, const int *bdy_num,
                          const int32_t *got, int64_t ld_got, int64_t row_base,
                          int64_t *first_bad) {
  int64_t bad = 0, firstrow = -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = first; p <= last; ++p) {
    const int s = blk_num[p];
    std::vector<float> dist((size_t)s * n);
    std::vector<int> path((size_t)s * n);
    orc_run_part(p, n, adj_size, row_offset, col, w, graph_id, st2ed, ed2st, blk_num, blk_off,
                 bdy_num, dist.data(), path.data());
    int64_t my_bad = 0, my_first = -1;
    for (int r = 0; r < s; ++r) {
      const int64_t gr = (int64_t)blk_off[p] + r - row_base;
      const int32_t *g = got + gr * ld_got;
      const float *d = dist.data() + (size_t)r * n;
      for (int j = 0; j < n; ++j) {
        const int32_t want = d[j] >= kMaxValue ? kInf : (int32_t)d[j];
        if (g[j] != want) {
          ++my_bad;
          if (my_first < 0) my_first = gr;
        }
      }
    }
#pragma omp critical
    {
      bad += my_bad;
      if (my_first >= 0 && (firstrow < 0 || my_first < firstrow)) firstrow = my_first;
    }
  }
  if (first_bad) *first_bad = firstrow;
  return bad;
}

// Textbook Floyd-Warshall on int32 (k-i-j), for kernel-level checks.
void orc_floyd_i32(int s, int32_t *mat, int64_t ld) {
  for (int k = 0; k < s; ++k)
    for (int i = 0; i < s; ++i) {
      const int32_t dik = mat[(int64_t)i * ld + k];
      for (int j = 0; j < s; ++j) {
        const int32_t c = dik + mat[(int64_t)k * ld + j];
        if (mat[(int64_t)i * ld + j] > c) mat[(int64_t)i * ld + j] = c;
      }
    }
}

// Overwrite min-plus on int32, init INT32_MAX / 2, strict '>'.
void orc_minplus_i32(const int32_t *a, int64_t lda, const int32_t *b, int64_t ldb, int32_t *c,
                     int64_t ldc, int M, int N, int K) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      int32_t best = kInf;
      for (int k = 0; k < K; ++k) {
        const int32_t cand = a[(int64_t)i * lda + k] + b[(int64_t)k * ldb + j];
        if (best > cand) best = cand;
      }
      c[(int64_t)i * ldc + j] = best;
    }
}

}  // extern "C"
