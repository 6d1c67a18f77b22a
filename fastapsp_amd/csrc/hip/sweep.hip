// One part's block from prepared parts: the per-call half of a solveSubGraph sweep.
//
// The reference's solveSubGraph(id) (fap/fap.cpp:195-503, run() for one part) computes
// the part's boundary SSSP, sub-matrix build, Floyd-Warshall, min-plus of the inner
// rows and decode on every call.  A caller that walks many parts (MoreTaskMpiExample.cpp:
// 91-96 solves every task of its rank in turn) pays the SSSP and Floyd-Warshall latency
// once per part, although neither depends on anything but the graph.  The sweep splits
// the call: fap_hip_prepare_parts (pipeline.hip) runs the SSSP, build and Floyd-Warshall
// of every part in one batched pass into caller-owned buffers, and
// fap_hip_merge_prepared writes one part's s_p x n block from them:
//   rows [0, b)   the part's boundary rows (a copy of the prepared SSSP rows) with their
//                 own columns taken from the Floyd-Warshall matrix (the reference's
//                 MysubMatDecode_path, fap.cpp:492-496);
//   rows [b, s)   min_k FW[b + i][k] + B[k][v] over the boundary rows (the reference's
//                 min_plus_path_advanced, fap.cpp:437-485), own columns from FW.
// Dense form: the streaming / tiled merge of the full solve.  Sparse form, for parts
// whose boundary rows reach few columns (directed k-NN graphs: a boundary source reaches
// a few hundred of 262,144 vertices; fap_hip_count_reached counts them per part once):
// a slot map names, per column, the part's own index, its index in the list of reached
// columns, or nothing; the product runs over the reached columns into a compact matrix;
// then one coalesced pass writes every row of the block (own columns from FW, boundary
// rows from the prepared rows, reached columns from the compact matrix, FAP_INF
// elsewhere).  Both forms give the same block bit for bit: the merge starts each entry
// at FAP_INF and takes minima of sums of values in [0, FAP_INF], so a column no boundary
// row reaches is FAP_INF exactly; an own column's FW value is never above the merge value
// (the closed matrix holds the boundary rows' own columns, so FW[i][v] <= FW[i][k] +
// B[k][v]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"

namespace fap_hip {

int launch_merge(const MergeDesc *d_descs, int nd, int total_tiles, int N, bool vec,
                 const int *d_graph_id, const int *d_ed2st, hipStream_t stream);
int merge_tiles(int M, int N);
int merge_stream_units(int M, int N, const StreamShape &sh);
int merge_stream_row_blocks(int M, const StreamShape &sh);
int merge_kind(int K);
int launch_merge_stream(const MergeDesc *d_descs, int nd, int total_units, int N, int kp,
                        const int *d_graph_id, const int *d_ed2st, const int *d_max_b, int lim16,
                        const StreamShape &sh, hipStream_t stream);
int launch_decode(const DecodeDesc *d_descs, int nd, int64_t total, const int *st2ed,
                  hipStream_t stream);

namespace {

constexpr int kSpRows = 16;   // rows per work unit of the sparse product
constexpr int kSpK = 16;      // boundary rows per staged chunk
constexpr int kSpWaves = 4;   // waves per workgroup (each takes its own units)

Workspace g_ws_sweep;  // descriptors, slot map, reached-column list, compact product

// Per column v of the block: slot[v] = the own index ed2st[v] - offset for a vertex of
// part `part_id`, num + j when v is the j-th column some boundary row reaches
// (B[k][v] < FAP_INF for some k < K; listed in cols in no particular order), else -1.
// *count is zero on entry.
__global__ __launch_bounds__(256) void slot_cols_kernel(const int *__restrict__ B, int64_t ldb,
                                                        int K, int n,
                                                        const int *__restrict__ graph_id,
                                                        const int *__restrict__ ed2st, int part_id,
                                                        int offset, int num,
                                                        int *__restrict__ slot,
                                                        int *__restrict__ cols,
                                                        int *__restrict__ count) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  bool active = false;
  int own = -1;
  if (v < n) {
    if (graph_id[v] == part_id) {
      own = ed2st[v] - offset;
    } else {
      const FAP_GLOBAL int *b = gptr(B) + v;
      for (int k = 0; k < K && !active; ++k) active = b[(int64_t)k * ldb] < kInf;
    }
  }
  const uint64_t mask = __ballot(active);
  int j = -1;
  if (mask != 0) {
    const int lane = threadIdx.x & 63, first = __builtin_ctzll(mask);
    int base = 0;
    if (lane == first) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, first);
    j = base + __popcll(mask & ((1ull << lane) - 1));
    if (active) cols[j] = v;
  }
  if (v < n) slot[v] = own >= 0 ? own : (active ? num + j : -1);
}

// Reached columns per part: counts[pi] += columns v outside part parts[pi] that one of its
// boundary rows (rows row0 .. row0 + bdy - 1 of B) reaches.  Grid: (ceil(n / 256), parts).
__global__ __launch_bounds__(256) void count_reached_kernel(const int *__restrict__ B,
                                                            int64_t ldb, int n,
                                                            const int *__restrict__ graph_id,
                                                            const int4 *__restrict__ parts,
                                                            int *__restrict__ counts) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const int4 q = parts[blockIdx.y];  // (part id, first row, boundary rows, -)
  bool active = false;
  if (v < n && graph_id[v] != q.x) {
    const FAP_GLOBAL int *b = gptr(B) + (int64_t)q.y * ldb + v;
    for (int k = 0; k < q.z && !active; ++k) active = b[(int64_t)k * ldb] < kInf;
  }
  const uint64_t mask = __ballot(active);
  if (mask != 0 && (threadIdx.x & 63) == __builtin_ctzll(mask))
    atomicAdd(counts + blockIdx.y, __popcll(mask));
}

// Q[num + j][b + i] = min_k A[i][k] + B[k][cols[j]] for inner rows i < M and listed
// columns j < min(*count, nq) (Q: the block's special columns, one row of ldq entries per
// column).  A work unit is 64 listed columns (one per lane) x kSpRows rows; the waves of
// the grid take units in turn until the list is done.  A rows are wave-uniform (scalar
// loads); A holds at least align16(K) readable entries per row (rows of the padded
// Floyd-Warshall matrix, whose stride is a multiple of 64 and at least K), and B entries
// past K read as FAP_INF.
__global__ __launch_bounds__(64 * kSpWaves) void sparse_merge_kernel(
    const int *__restrict__ A, int64_t lda, const int *__restrict__ B, int64_t ldb,
    int *__restrict__ Q, int64_t ldq, int num, int nq, int M, int K,
    const int *__restrict__ cols, const int *__restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int ncols = min(*count, nq);
  const int b0 = num - M;
  const int col_blocks = (ncols + 63) / 64;
  const int row_blocks = (M + kSpRows - 1) / kSpRows;
  const int64_t units = (int64_t)col_blocks * row_blocks;
  const int64_t waves = (int64_t)gridDim.x * kSpWaves;
  const FAP_CONST int *abase = uniform_cptr(A);
  for (int64_t u = (int64_t)blockIdx.x * kSpWaves + (threadIdx.x >> 6); u < units; u += waves) {
    const int cb = (int)(u / row_blocks), rb = (int)(u - (int64_t)cb * row_blocks);
    const int j = cb * 64 + lane;
    const int v = j < ncols ? cols[j] : -1;
    const int r0 = rb * kSpRows;
    int acc[kSpRows];
#pragma unroll
    for (int r = 0; r < kSpRows; ++r) acc[r] = kInf;
    for (int k0 = 0; k0 < K; k0 += kSpK) {
      int b[kSpK];
#pragma unroll
      for (int kk = 0; kk < kSpK; ++kk)
        b[kk] = (v >= 0 && k0 + kk < K) ? B[(int64_t)(k0 + kk) * ldb + v] : kInf;
#pragma unroll
      for (int r = 0; r < kSpRows; ++r) {
        if (r0 + r >= M) break;  // wave-uniform
        const FAP_CONST int *arow = abase + (int64_t)(r0 + r) * lda + k0;
        int a[kSpK];
#pragma unroll
        for (int kk = 0; kk < kSpK; ++kk) a[kk] = arow[kk];
#pragma unroll
        for (int kk = 0; kk < kSpK; kk += 2)
          acc[r] = min(acc[r], min(a[kk] + b[kk], a[kk + 1] + b[kk + 1]));
      }
    }
    if (v >= 0) {
      FAP_GLOBAL int *q = gptr(Q) + (int64_t)(num + j) * ldq + b0 + r0;
#pragma unroll
      for (int r = 0; r < kSpRows; ++r)
        if (r0 + r < M) q[r] = acc[r];
    }
  }
}

typedef int v4i __attribute__((ext_vector_type(4)));

// Q[c][i] = FW[i][c] for i, c < num (the block's own columns, one Q row per column),
// through a 64 x 65 LDS tile.  Grid: (ceil(num / 64), ceil(num / 64)), 256 threads.
__global__ __launch_bounds__(256) void own_cols_kernel(const int *__restrict__ fw, int64_t ldf,
                                                       int num, int *__restrict__ Q,
                                                       int64_t ldq) {
  __shared__ int t[64][65];
  const int c0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4)
    t[r][tx] = (i0 + r < num && c0 + tx < num) ? fw[(int64_t)(i0 + r) * ldf + c0 + tx] : kInf;
  __syncthreads();
  for (int r = ty; r < 64; r += 4)
    if (c0 + r < num && i0 + tx < num) Q[(int64_t)(c0 + r) * ldq + i0 + tx] = t[tx][r];
}

constexpr int kGrRows = 16;   // rows of the block per workgroup of the row pass
constexpr int kGrChunks = 1;  // 1024-column chunks per workgroup (4 columns per thread each)

// Rows blockIdx.y * kGrRows .. (+ kGrRows) of the block over kGrChunks chunks of 1024
// columns.  Column v with slot sv: an own column (sv < num) or, in inner rows, a reached
// column takes Q[sv][i]; other columns take the prepared row in boundary rows (i < b) and
// FAP_INF in inner rows.  A thread first loads everything its 4 x kGrChunks columns need
// for its kGrRows rows (the slot map, then two 16-byte loads of Q per special column),
// then writes one 16-byte non-temporal store per row and chunk: no store waits on a load
// made for its own row (with one gather per row and chunk in the loop, the pass wrote the
// 1.1 GB block of a 262,144-column k-NN part at 4.1 TB/s on MI355X; in this form 5.4 TB/s
// with 8 rows x 2048 columns per workgroup, 5.7 with 16 x 1024, 5.6 with 32 x 1024).
// ldq: a multiple of
// kGrRows at least num, so Q reads stay inside the row of the column.  vec: n, ld_out and
// ldb multiples of 4 and 16-byte aligned rows.
__global__ __launch_bounds__(256) void gather_rows_kernel(
    int *__restrict__ out, int64_t ld_out, int num, int b, int n, const int *__restrict__ B,
    int64_t ldb, const int *__restrict__ Q, int64_t ldq, int nq,
    const int *__restrict__ slot, int vec) {
  const int i0 = blockIdx.y * kGrRows;
  const int base = blockIdx.x * (1024 * kGrChunks) + threadIdx.x * 4;
  int s[kGrChunks][4];
#pragma unroll
  for (int c = 0; c < kGrChunks; ++c) {
    const int v0 = base + c * 1024;
    if (vec && v0 + 3 < n) {
      const v4i x = *(const FAP_GLOBAL v4i *)(gptr(slot) + v0);
      s[c][0] = x.x, s[c][1] = x.y, s[c][2] = x.z, s[c][3] = x.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) s[c][e] = v0 + e < n ? slot[v0 + e] : -1;
    }
  }
  int qv[kGrChunks][4][kGrRows];
#pragma unroll
  for (int c = 0; c < kGrChunks; ++c)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int sv = s[c][e];
      if (sv >= 0 && sv < num + nq) {
        const FAP_GLOBAL v4i *q = (const FAP_GLOBAL v4i *)(gptr(Q) + (int64_t)sv * ldq + i0);
#pragma unroll
        for (int h = 0; h < kGrRows / 4; ++h) {
          const v4i x = q[h];
          qv[c][e][4 * h] = x.x, qv[c][e][4 * h + 1] = x.y;
          qv[c][e][4 * h + 2] = x.z, qv[c][e][4 * h + 3] = x.w;
        }
      } else {
#pragma unroll
        for (int r = 0; r < kGrRows; ++r) qv[c][e][r] = kInf;
      }
    }
#pragma unroll
  for (int r = 0; r < kGrRows; ++r) {
    const int i = i0 + r;
    if (i >= num) break;  // uniform
    const FAP_GLOBAL int *brow = gptr(B) + (int64_t)i * ldb;
    FAP_GLOBAL int *orow = gptr(out) + (int64_t)i * ld_out;
#pragma unroll
    for (int c = 0; c < kGrChunks; ++c) {
      const int v0 = base + c * 1024;
      if (v0 >= n) break;
      const bool full = vec && v0 + 3 < n;
      int val[4];
      if (i < b) {  // boundary row: the prepared row except at own columns
        int bv[4];
        if (full) {
          const v4i y = *(const FAP_GLOBAL v4i *)(brow + v0);
          bv[0] = y.x, bv[1] = y.y, bv[2] = y.z, bv[3] = y.w;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) bv[e] = v0 + e < n ? brow[v0 + e] : kInf;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) val[e] = (s[c][e] >= 0 && s[c][e] < num) ? qv[c][e][r] : bv[e];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) val[e] = qv[c][e][r];
      }
      if (full) {
        const v4i y = {val[0], val[1], val[2], val[3]};
        __builtin_nontemporal_store(y, (FAP_GLOBAL v4i *)(orow + v0));
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (v0 + e < n) orow[v0 + e] = val[e];
      }
    }
  }
}

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_count_reached(const fap_graph_dev *g, const fap_part_desc *parts,
                                     int32_t nparts, const int32_t *d_rows, int64_t ld_rows,
                                     int32_t *counts, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && g->n > 0 && g->d_graph_id && nparts >= 0 && (nparts == 0 || (parts && counts)),
              "fap_hip_count_reached: null argument");
  if (nparts == 0) return FAP_OK;
  int64_t rows = 0;
  std::vector<int4> tab(nparts);
  for (int p = 0; p < nparts; ++p) {
    const fap_part_desc &q = parts[p];
    FAP_REQUIRE(q.bdy >= 0 && q.out_row >= 0, "fap_hip_count_reached: bad part %d", q.part_id);
    tab[p] = make_int4(q.part_id, (int)q.out_row, q.bdy, 0);
    rows = std::max<int64_t>(rows, q.out_row + q.bdy);
  }
  FAP_REQUIRE(rows == 0 || (d_rows && ld_rows >= g->n && rows < (1LL << 31)),
              "fap_hip_count_reached: bad boundary-row buffer");
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  const int64_t o_cnt = 16LL * nparts;
  FAP_TRY(g_ws_sweep.ensure(o_cnt + 4LL * nparts));
  char *meta = static_cast<char *>(g_ws_sweep.ptr);
  FAP_CHECK_HIP(hipMemcpyAsync(meta, tab.data(), o_cnt, hipMemcpyHostToDevice, st));
  int *d_counts = reinterpret_cast<int *>(meta + o_cnt);
  FAP_TRY(fap_hip_fill_i32(d_counts, 0, nparts, st));
  for (int p0 = 0; p0 < nparts; p0 += 65535) {  // parts on the grid's y dimension
    const int cnt = std::min(65535, nparts - p0);
    hipLaunchKernelGGL(count_reached_kernel, dim3((g->n + 255) / 256, cnt), dim3(256), 0, st,
                       d_rows, ld_rows, g->n, g->d_graph_id,
                       reinterpret_cast<const int4 *>(meta) + p0, d_counts + p0);
    FAP_CHECK_LAUNCH();
  }
  FAP_CHECK_HIP(hipMemcpyAsync(counts, d_counts, 4LL * nparts, hipMemcpyDeviceToHost, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return ws_fence_record(st);
}

extern "C" int fap_hip_merge_prepared(const fap_graph_dev *g, const fap_part_desc *part,
                                      const int32_t *d_fw, const int32_t *d_rows, int64_t ld_rows,
                                      int64_t reached, int32_t *d_out, int64_t ld_out,
                                      void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && part && d_out && g->n > 0 && g->d_graph_id && g->d_st2ed && g->d_ed2st,
              "fap_hip_merge_prepared: null argument");
  const fap_part_desc q = *part;
  const int n = g->n;
  FAP_REQUIRE(q.part_id >= 1 && q.num >= 0 && q.bdy >= 0 && q.bdy <= q.num && q.offset >= 0 &&
                  (int64_t)q.offset + q.num <= n && q.out_row >= 0 && ld_out >= n,
              "fap_hip_merge_prepared: bad part %d", q.part_id);
  FAP_REQUIRE(q.num == 0 || d_fw, "fap_hip_merge_prepared: no Floyd-Warshall matrix");
  FAP_REQUIRE(q.bdy == 0 || (d_rows && ld_rows >= n), "fap_hip_merge_prepared: no boundary rows");
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  FAP_TRY(timing_start(st));
  if (q.num == 0) {
    FAP_TRY(timing_finish(st));
    return ws_fence_record(st);
  }
  FAP_REQUIRE(reached < 0 || (int64_t)(q.num - q.bdy) * reached < (1LL << 40),
              "fap_hip_merge_prepared: reached-column count out of range");
  const bool sparse = reached >= 0;
  const int S = (q.num + 63) / 64 * 64;
  const int b = q.bdy, M = q.num - q.bdy;
  int32_t *out = d_out + q.out_row * ld_out;
  int32_t *C = out + (int64_t)b * ld_out;

  MergeDesc m{};
  m.A = d_fw + (int64_t)b * S;
  m.lda = S;
  m.B = d_rows;
  m.ldb = ld_rows;
  m.C = C;
  m.ldc = ld_out;
  m.M = M;
  m.K = b;
  m.part_id = q.part_id;
  m.offset = q.offset;
  m.fw = d_fw;
  m.ldf = S;
  m.fw_row0 = b;
  const int kind = merge_kind(b);
  const StreamShape shape = stream_shape();
  int units = 0;
  if (M > 0 && !sparse) {
    m.row_blocks = kind > 0 ? merge_stream_row_blocks(M, shape) : (M + 63) / 64;
    units = kind > 0 ? merge_stream_units(M, n, shape) : merge_tiles(M, n);
  }
  // Dense form: the boundary rows' own columns (the merge epilogue writes the inner
  // rows').
  const DecodeDesc dd{d_fw, S, out, ld_out, b, q.num, q.offset, 0};

  // Workspace: descriptors, count, slot map (n), reached columns (n), compact product.
  const int64_t o_md = 0, o_dd = 256, o_cnt = 512, o_slot = 768;
  const int64_t o_cols = o_slot + (4LL * n + 255) / 256 * 256;
  const int64_t o_q = o_cols + (4LL * n + 255) / 256 * 256;
  // Q: one row of ldq entries per special column (own columns, then reached columns).
  const int nq = sparse ? (int)std::min<int64_t>(reached, n) : 0;
  const int64_t ldq = (q.num + kGrRows - 1) / kGrRows * kGrRows;
  FAP_TRY(g_ws_sweep.ensure(o_q + (sparse ? 4LL * (q.num + nq) * ldq : 0)));
  char *meta = static_cast<char *>(g_ws_sweep.ptr);
  if (!sparse) {  // the dense kernels read descriptors; the sparse ones take arguments
    char host[512] = {};
    std::memcpy(host + o_md, &m, sizeof m);
    std::memcpy(host + o_dd, &dd, sizeof dd);
    FAP_CHECK_HIP(hipMemcpyAsync(meta, host, sizeof host, hipMemcpyHostToDevice, st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));  // the descriptors come from the stack
  }
  const MergeDesc *d_md = reinterpret_cast<const MergeDesc *>(meta + o_md);
  const DecodeDesc *d_dd = reinterpret_cast<const DecodeDesc *>(meta + o_dd);
  int *d_cnt = reinterpret_cast<int *>(meta + o_cnt);
  int *d_slot = reinterpret_cast<int *>(meta + o_slot);
  int *d_cols = reinterpret_cast<int *>(meta + o_cols);
  int *d_q = reinterpret_cast<int *>(meta + o_q);

  hipEvent_t e0, e1, e2, e3;
  FAP_TRY(timing_record(st, &e0));
  if (sparse) {  // slot map and reached columns
    FAP_TRY(fap_hip_fill_i32(d_cnt, 0, 1, st));
    hipLaunchKernelGGL(slot_cols_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_rows,
                       ld_rows, b, n, g->d_graph_id, g->d_ed2st, q.part_id, q.offset, q.num,
                       d_slot, d_cols, d_cnt);
    FAP_CHECK_LAUNCH();
  } else if (b > 0) {  // boundary rows: the prepared SSSP rows
    FAP_CHECK_HIP(hipMemcpy2DAsync(out, ld_out * 4, d_rows, ld_rows * 4, (size_t)n * 4, b,
                                   hipMemcpyDeviceToDevice, st));
  }
  FAP_TRY(timing_record(st, &e1));
  timing_add_span(0, e0, e1);
  if (sparse) {
    if (M > 0 && b > 0 && reached > 0) {
      const int grid = std::max(1, device_cus() * 4);
      hipLaunchKernelGGL(sparse_merge_kernel, dim3(grid), dim3(64 * kSpWaves), 0, st, m.A, m.lda,
                         d_rows, ld_rows, d_q, ldq, q.num, nq, M, b, d_cols, d_cnt);
      FAP_CHECK_LAUNCH();
    }
    const unsigned tiles = (unsigned)((q.num + 63) / 64);
    hipLaunchKernelGGL(own_cols_kernel, dim3(tiles, tiles), dim3(256), 0, st, d_fw, (int64_t)S,
                       q.num, d_q, ldq);
    FAP_CHECK_LAUNCH();
  } else if (M > 0) {
    if (kind == 0)
      FAP_TRY(launch_merge(d_md, 1, units, n, (ld_out % 4 == 0) && ((uintptr_t)d_out % 16 == 0),
                           g->d_graph_id, g->d_ed2st, st));
    else
      FAP_TRY(launch_merge_stream(d_md, 1, units, n, kind, g->d_graph_id, g->d_ed2st, nullptr, 0,
                                  shape, st));
  }
  FAP_TRY(timing_record(st, &e2));
  timing_add_span(3, e1, e2);
  if (sparse) {  // every row of the block in one coalesced pass
    const int vec = n % 4 == 0 && ld_out % 4 == 0 && ld_rows % 4 == 0 &&
                    (uintptr_t)out % 16 == 0 && (uintptr_t)d_rows % 16 == 0;
    hipLaunchKernelGGL(gather_rows_kernel,
                       dim3((n + 1024 * kGrChunks - 1) / (1024 * kGrChunks),
                            (q.num + kGrRows - 1) / kGrRows),
                       dim3(256), 0, st, out,
                       ld_out, q.num, b, n, b > 0 ? d_rows : out, ld_rows, d_q, ldq, nq, d_slot,
                       vec);
    FAP_CHECK_LAUNCH();
  } else {
    FAP_TRY(launch_decode(d_dd, 1, (int64_t)b * q.num, g->d_st2ed, st));
  }
  FAP_TRY(timing_record(st, &e3));
  timing_add_span(4, e2, e3);
  FAP_TRY(timing_finish(st));
  return ws_fence_record(st);
}
