// Min-plus merge: inner rows of every part against its boundary rows, with the
// decode of the part's own columns fused into the epilogue.
//
// Replaces min_plus_path_CPU_advanced (fap/kernel/minplus.h:51-64; overwrite
// semantics, init MAXVALUE, strict '>') plus MyMat1Build
// (fap/utils/util_Myalgorithm.h:278-285: the A operand is read in place as a
// strided view of the Floyd-Warshall matrix) and MysubMatDecode_path (:236-247,
// fused: columns of the part's own vertices take the Floyd-Warshall distance).
// The reference's GPU twin is MatrixMulCUDA6_path<32,32,32,4,4>
// (fap/kernel/cuda/minplus.cu:40-189) with padded copies in and out (:268-359).
//
// gfx950 mapping:
//  * workgroup = 256 threads (4 wave64), output tile 64 rows x 256 columns,
//    8 x 8 accumulators per thread (column groups tx*4 and 128 + tx*4 so each
//    16-byte LDS read and each 16-byte global store of a wave is one contiguous
//    run);
//  * A^T (K x 64) and B (K x 256) staged through LDS in chunks of 32 k-steps;
//  * VALU v_add_u32 + v_min3_i32 (1.5 instructions per min-plus step);
//  * one launch covers every product of the batch through a work table
//    (MergeDesc.tile_begin) and tiles that share a B panel are mapped to the same
//    XCD back to back, so each 32 KB B panel comes from HBM once per XCD L2;
//  * output rows are written once with non-temporal 16-byte stores.
#include "fap_internal.h"
#include "fw_desc.h"
#include "mp_tile.h"

namespace fap_hip {

constexpr int kMpRows = 64;
constexpr int kMpCols = 256;
constexpr int kMpKc = 32;

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_desc(const MergeDesc *__restrict__ d, int nd, int t) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].tile_begin <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// blockIdx -> logical tile so that consecutive logical tiles run on one XCD
// (workgroups are dispatched round-robin over the 8 XCDs).  gridDim.x is a
// multiple of 8.
__device__ __forceinline__ int xcd_logical_tile() {
  const int per_xcd = gridDim.x >> 3;
  return (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
}

template <bool kVec>
__global__ __launch_bounds__(256) void merge_kernel(const MergeDesc *__restrict__ descs, int nd,
                                                    int total_tiles, int N, int ncb,
                                                    const int *__restrict__ graph_id,
                                                    const int *__restrict__ ed2st) {
  const int t = xcd_logical_tile();
  if (t >= total_tiles) return;
  const MergeDesc d = descs[find_desc(descs, nd, t)];
  const int local = t - d.tile_begin;
  const int cb = local / d.row_blocks;
  const int rb = local - cb * d.row_blocks;
  const int row0 = rb * kMpRows;
  const int col0 = cb * kMpCols;

  __shared__ int As[kMpKc][kMpRows];
  __shared__ int Bs[kMpKc][kMpCols];

  const int tid = threadIdx.x;
  const int tx = tid & 31;  // columns col0 + tx*4 + {0..3} and col0 + 128 + tx*4 + {0..3}
  const int ty = tid >> 5;  // rows row0 + ty*8 + {0..7}

  int acc[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[r][c] = kInf;

  for (int k0 = 0; k0 < d.K; k0 += kMpKc) {
    if (k0 > 0) __syncthreads();
    // A chunk: As[kk][i] = A[row0 + i][k0 + kk]
    {
      const int i = tid & 63;
      const int kq = (tid >> 6) * 8;
      const bool row_ok = row0 + i < d.M;
      const int *arow = d.A + (int64_t)(row0 + i) * d.lda + k0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int kk = kq + q;
        As[kk][i] = (row_ok && k0 + kk < d.K) ? arow[kk] : kInf;
      }
    }
    // B chunk: Bs[kk][c] = B[k0 + kk][col0 + c]
    {
      const int c4 = (tid & 63) * 4;
      const int j = col0 + c4;
#pragma unroll
      for (int q = 0; q < kMpKc / 4; ++q) {
        const int kk = (tid >> 6) + 4 * q;
        const bool k_ok = k0 + kk < d.K;
        const int *brow = d.B + (int64_t)(k0 + kk) * d.ldb;
        int4 v;
        if (kVec && k_ok && j + 3 < N) {
          v = *reinterpret_cast<const int4 *>(brow + j);
        } else {
          v.x = (k_ok && j + 0 < N) ? brow[j + 0] : kInf;
          v.y = (k_ok && j + 1 < N) ? brow[j + 1] : kInf;
          v.z = (k_ok && j + 2 < N) ? brow[j + 2] : kInf;
          v.w = (k_ok && j + 3 < N) ? brow[j + 3] : kInf;
        }
        *reinterpret_cast<int4 *>(&Bs[kk][c4]) = v;
      }
    }
    __syncthreads();
    const int kc = min(kMpKc, d.K - k0);
    if (kc == kMpKc)
      mp_accumulate<8, 8, kMpKc>(acc, &As[0][0], kMpRows, &Bs[0][0], kMpCols, ty * 8, 4, tx * 4,
                                 128);
    else
      mp_accumulate_n<8, 8>(acc, &As[0][0], kMpRows, &Bs[0][0], kMpCols, ty * 8, 4, tx * 4, 128,
                            (kc + 1) & ~1);
  }

  // Epilogue: fused decode of the part's own columns, then one store per row.
  int cols[8];
  bool col_ok[8];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      cols[4 * g + e] = col0 + g * 128 + tx * 4 + e;
      col_ok[4 * g + e] = cols[4 * g + e] < N;
    }
  int local_col[8];
  bool any_own = false;
  if (d.part_id > 0) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      local_col[c] = -1;
      if (col_ok[c] && graph_id[cols[c]] == d.part_id) {
        local_col[c] = ed2st[cols[c]] - d.offset;
        any_own = true;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int i = row0 + ty * 8 + r;
    if (i >= d.M) break;
    if (any_own) {
      const int *fwrow = d.fw + (int64_t)(d.fw_row0 + i) * d.ldf;
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (local_col[c] >= 0) acc[r][c] = fwrow[local_col[c]];
    }
    int *crow = d.C + (int64_t)i * d.ldc;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int j = cols[4 * g];
      if (kVec && j + 3 < N) {
        v4i v = {acc[r][4 * g + 0], acc[r][4 * g + 1], acc[r][4 * g + 2], acc[r][4 * g + 3]};
        __builtin_nontemporal_store(v, reinterpret_cast<v4i *>(crow + j));
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (j + e < N) crow[j + e] = acc[r][4 * g + e];
      }
    }
  }
}

int launch_merge(const MergeDesc *d_descs, int nd, int total_tiles, int N, bool vec,
                 const int *d_graph_id, const int *d_ed2st, hipStream_t stream) {
  if (nd <= 0 || total_tiles <= 0 || N <= 0) return FAP_OK;
  const int ncb = (N + kMpCols - 1) / kMpCols;
  const int grid = (total_tiles + 7) / 8 * 8;
  if (vec)
    hipLaunchKernelGGL(merge_kernel<true>, dim3(grid), dim3(256), 0, stream, d_descs, nd,
                       total_tiles, N, ncb, d_graph_id, d_ed2st);
  else
    hipLaunchKernelGGL(merge_kernel<false>, dim3(grid), dim3(256), 0, stream, d_descs, nd,
                       total_tiles, N, ncb, d_graph_id, d_ed2st);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int merge_tiles(int M, int N) {
  if (M <= 0 || N <= 0) return 0;
  return ((M + kMpRows - 1) / kMpRows) * ((N + kMpCols - 1) / kMpCols);
}

}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_minplus(const int32_t *d_A, int64_t lda, const int32_t *d_B, int64_t ldb,
                               int32_t *d_C, int64_t ldc, int32_t M, int32_t N, int32_t K,
                               void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(M >= 0 && N >= 0 && K >= 0, "fap_hip_minplus: negative size");
  if (M == 0 || N == 0) return FAP_OK;
  FAP_REQUIRE(d_C && ldc >= N, "fap_hip_minplus: bad C");
  FAP_REQUIRE(K == 0 || (d_A && d_B && lda >= K && ldb >= N), "fap_hip_minplus: bad A/B");
  const int64_t tiles64 = (int64_t)((M + kMpRows - 1) / kMpRows) * ((N + kMpCols - 1) / kMpCols);
  FAP_REQUIRE(tiles64 < (1LL << 30), "fap_hip_minplus: product too large for one launch");
  hipStream_t st = resolve(stream);
  MergeDesc h{};
  h.A = d_A;
  h.lda = lda;
  h.B = d_B;
  h.ldb = ldb;
  h.C = d_C;
  h.ldc = ldc;
  h.M = M;
  h.K = K;
  h.row_blocks = (M + kMpRows - 1) / kMpRows;
  h.tile_begin = 0;
  h.part_id = 0;
  static Workspace ws;
  FAP_TRY(ws.ensure(sizeof(MergeDesc)));
  FAP_CHECK_HIP(hipMemcpyAsync(ws.ptr, &h, sizeof(h), hipMemcpyHostToDevice, st));
  const bool vec = (ldb % 4 == 0) && (ldc % 4 == 0) && ((uintptr_t)d_B % 16 == 0) &&
                   ((uintptr_t)d_C % 16 == 0);
  FAP_TRY(launch_merge(static_cast<MergeDesc *>(ws.ptr), 1, (int)tiles64, N, vec, nullptr,
                       nullptr, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}
