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
// Dense form: the streaming / tiled merge of the full solve.  Sparse form, for graphs
// whose boundary rows are mostly unreachable (directed k-NN graphs: a boundary source
// reaches a few hundred of 262,144 vertices): the inner rows are filled with FAP_INF,
// the columns some boundary row reaches are listed on the device, the product runs over
// those columns only, and the own columns of every row are decoded from FW.  Both forms
// give the same block bit for bit: the merge starts each entry at FAP_INF and takes
// minima of sums of values in [0, FAP_INF], so a column no boundary row reaches is
// FAP_INF exactly; an own column's FW value is never above the merge value (the closed
// matrix holds the boundary rows' own columns, so FW[i][v] <= FW[i][k] + B[k][v]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>

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

Workspace g_ws_sweep;  // descriptors + active-column list and count

// Columns v outside part `part_id` that some boundary row reaches (B[k][v] < FAP_INF for
// some k < K), appended to cols in no particular order; *count is zero on entry.
__global__ __launch_bounds__(256) void active_cols_kernel(const int *__restrict__ B, int64_t ldb,
                                                          int K, int n,
                                                          const int *__restrict__ graph_id,
                                                          int part_id, int *__restrict__ cols,
                                                          int *__restrict__ count) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  bool active = false;
  if (v < n && graph_id[v] != part_id) {
    const FAP_GLOBAL int *b = gptr(B) + v;
    for (int k = 0; k < K && !active; ++k) active = b[(int64_t)k * ldb] < kInf;
  }
  const uint64_t mask = __ballot(active);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == __builtin_ctzll(mask)) base = atomicAdd(count, __popcll(mask));
  base = __shfl(base, __builtin_ctzll(mask));
  if (active) cols[base + __popcll(mask & ((1ull << lane) - 1))] = v;
}

// C[i][cols[j]] = min_k A[i][k] + B[k][cols[j]] for i < M and j < *count.  A work unit is
// 64 listed columns (one per lane) x kSpRows rows; the waves of the grid take units in
// turn until the list is done.  A rows are wave-uniform (scalar loads); A holds at least
// align16(K) readable entries per row (rows of the padded Floyd-Warshall matrix, whose
// stride is a multiple of 64 and at least K), and B entries past K read as FAP_INF.
__global__ __launch_bounds__(64 * kSpWaves) void sparse_merge_kernel(
    const int *__restrict__ A, int64_t lda, const int *__restrict__ B, int64_t ldb,
    int *__restrict__ C, int64_t ldc, int M, int K, const int *__restrict__ cols,
    const int *__restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int ncols = *count;
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
      FAP_GLOBAL int *c = gptr(C) + v;
#pragma unroll
      for (int r = 0; r < kSpRows; ++r)
        if (r0 + r < M) c[(int64_t)(r0 + r) * ldc] = acc[r];
    }
  }
}

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_merge_prepared(const fap_graph_dev *g, const fap_part_desc *part,
                                      const int32_t *d_fw, const int32_t *d_rows, int64_t ld_rows,
                                      int32_t sparse, int32_t *d_out, int64_t ld_out,
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
  // Sparse form: every row's own columns; dense form: the boundary rows' (the merge
  // epilogue writes the inner rows' own columns).
  const int dec_rows = sparse ? q.num : b;
  const DecodeDesc dd{d_fw, S, out, ld_out, dec_rows, q.num, q.offset, 0};

  const int64_t o_md = 0, o_dd = 256, o_cnt = 512, o_cols = 768;
  FAP_TRY(g_ws_sweep.ensure(o_cols + 4LL * n));
  char *meta = static_cast<char *>(g_ws_sweep.ptr);
  char host[512] = {};
  std::memcpy(host + o_md, &m, sizeof m);
  std::memcpy(host + o_dd, &dd, sizeof dd);
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host, sizeof host, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));  // the descriptors come from the stack
  const MergeDesc *d_md = reinterpret_cast<const MergeDesc *>(meta + o_md);
  const DecodeDesc *d_dd = reinterpret_cast<const DecodeDesc *>(meta + o_dd);
  int *d_cnt = reinterpret_cast<int *>(meta + o_cnt);
  int *d_cols = reinterpret_cast<int *>(meta + o_cols);

  hipEvent_t e0, e1, e2, e3;
  FAP_TRY(timing_record(st, &e0));
  if (b > 0)  // boundary rows: the prepared SSSP rows
    FAP_CHECK_HIP(hipMemcpy2DAsync(out, ld_out * 4, d_rows, ld_rows * 4, (size_t)n * 4, b,
                                   hipMemcpyDeviceToDevice, st));
  FAP_TRY(timing_record(st, &e1));
  timing_add_span(0, e0, e1);
  if (M > 0 && sparse) {
    FAP_TRY(fap_hip_fill_i32(d_cnt, 0, 1, st));
    FAP_TRY(fap_hip_fill_i32(C, kInf, (int64_t)M * ld_out, st));
    if (b > 0) {
      hipLaunchKernelGGL(active_cols_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_rows,
                         ld_rows, b, n, g->d_graph_id, q.part_id, d_cols, d_cnt);
      FAP_CHECK_LAUNCH();
      const int grid = std::max(1, device_cus() * 4);
      hipLaunchKernelGGL(sparse_merge_kernel, dim3(grid), dim3(64 * kSpWaves), 0, st, m.A, m.lda,
                         d_rows, ld_rows, C, ld_out, M, b, d_cols, d_cnt);
      FAP_CHECK_LAUNCH();
    }
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
  FAP_TRY(launch_decode(d_dd, 1, (int64_t)dec_rows * q.num, g->d_st2ed, st));
  FAP_TRY(timing_record(st, &e3));
  timing_add_span(4, e2, e3);
  FAP_TRY(timing_finish(st));
  return ws_fence_record(st);
}
