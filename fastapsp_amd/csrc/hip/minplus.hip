// Min-plus merge: inner rows of every part against its boundary rows, with the
// decode of the part's own columns fused into the epilogue.
//
// Replaces min_plus_path_CPU_advanced (fap/kernel/minplus.h:51-64; overwrite
// semantics, init MAXVALUE, strict '>') plus MyMat1Build
// (fap/utils/util_Myalgorithm.h:278-285: the A operand is read in place as a
// strided view of the Floyd-Warshall matrix) and MysubMatDecode_path (:236-247,
// fused: columns of the part's own vertices take the minimum of the product and
// the Floyd-Warshall distance -- the distance itself when the FW matrix holds the
// boundary rows, which no product entry undercuts).
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
#include <algorithm>
#include <atomic>
#include <vector>
#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"
#include "mp_tile.h"
#include "pk16.h"

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
      // Descriptor pointers are generic: gptr keeps every access a global one (a flat
      // access would also hold up the LDS counter of the staging writes).
      const FAP_GLOBAL int *arow = gptr(d.A + (int64_t)(row0 + i) * d.lda + k0);
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
        const FAP_GLOBAL int *brow = gptr(d.B + (int64_t)(k0 + kk) * d.ldb);
        v4i v;
        if (kVec && k_ok && j + 3 < N) {
          v = *reinterpret_cast<const FAP_GLOBAL v4i *>(brow + j);
        } else {
          v.x = (k_ok && j + 0 < N) ? brow[j + 0] : kInf;
          v.y = (k_ok && j + 1 < N) ? brow[j + 1] : kInf;
          v.z = (k_ok && j + 2 < N) ? brow[j + 2] : kInf;
          v.w = (k_ok && j + 3 < N) ? brow[j + 3] : kInf;
        }
        *reinterpret_cast<v4i *>(&Bs[kk][c4]) = v;
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
      const FAP_GLOBAL int *fwrow = gptr(d.fw + (int64_t)(d.fw_row0 + i) * d.ldf);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (local_col[c] >= 0) acc[r][c] = min(acc[r][c], fwrow[local_col[c]]);
    }
    FAP_GLOBAL int *crow = gptr(d.C + (int64_t)i * d.ldc);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int j = cols[4 * g];
      if (kVec && j + 3 < N) {
        v4i v = {acc[r][4 * g + 0], acc[r][4 * g + 1], acc[r][4 * g + 2], acc[r][4 * g + 3]};
        __builtin_nontemporal_store(v, reinterpret_cast<FAP_GLOBAL v4i *>(crow + j));
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (j + e < N) crow[j + e] = acc[r][4 * g + e];
      }
    }
  }
}

// Streaming form for small K (the boundary count of a road-graph part is ~10):
// the product is a write stream of M x N int32, so one wave keeps its B panel
// (KP x 256 columns, KP = K rounded up to a multiple of 2, FAP_INF padding) in
// registers, reads each A row wave-uniformly (scalar loads) and writes one
// contiguous 1 KB row segment per row.  Lane l holds columns cb * 256 + l + 64 e
// (e < 4), so each of its four dword stores covers 256 contiguous bytes: measured
// in one process on MI355X, into the same 52.5 GB result buffer, against 4 adjacent
// columns per lane with one dwordx4 store: 10.69 ms against 10.82.  Work unit: one
// 256-column block x kStRows rows of one product; one launch per KP group.  Host requirements:
// K <= KP and every A row holds KP readable entries (rows of the padded
// Floyd-Warshall matrices, stride >= 64).  Entries A[i][k] with K <= k < KP are
// distances in [0, FAP_INF] and meet B = FAP_INF, so they never win a minimum.
constexpr int kStCols = 256;

// Work-unit shape of the streaming merge (StreamShape, fap_internal.h): rows per
// unit, unit order inside a product (0: the row blocks of one column block back to
// back, 1: the column blocks of one row block back to back) and a cap on the resident
// waves per CU (0: none; enforced through the LDS each one-wave workgroup reserves).
// Default: 128 rows in row-block order (a unit's B panel is reused from L2 by the next
// row blocks of its column block).  Measured on MI355X (the 52.5 GB result of the
// 114,599-vertex graph, K-groups on three streams, one result buffer per process):
// 9.11 ms against 9.25 for column-block order; 64 rows 9.45, 256 rows 9.41, 512 rows
// 9.88; a cap of 24 resident waves per CU 9.11.  In a result buffer in a slowly
// written range of device memory the order was reversed (10.73 against 10.59 ms).
// Set with fap_hip_set_merge_shape; a solve reads it once, when it plans its merge.
std::atomic<int64_t> g_stream_shape{128};

StreamShape stream_shape() {
  const int64_t v = g_stream_shape.load(std::memory_order_relaxed);
  return StreamShape{(int)(v & 0xFFFF), (int)((v >> 16) & 0xFF), (int)((v >> 24) & 0xFF)};
}

// One row per loop iteration, so a wave keeps few row stores in flight: on MI355X
// the merge stage of a 52.5 GB result took 11.8 ms with 8 rows in flight per wave
// and 9.6 ms with one.

// The A values come from scalar loads (SGPRs), so they never contend with the VGPRs
// of stores still in flight.
template <int KP>
__device__ __forceinline__ void stream_row(int (&acc)[4], const int (&a)[KP],
                                           const int (&b)[KP][4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = kInf;
#pragma unroll
  for (int k = 0; k < KP; k += 2)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[e] = min3_i32(acc[e], a[k] + b[k][e], a[k + 1] + b[k + 1][e]);
}

// The same row on packed 16-bit distances (pk16.h): the B panel holds column pairs
// min(B, kCap16), each A value is capped and broadcast to both halves, so a step on
// two entries is one v_pk_add_u16 and half a v_pk_minimum3_f16 (0.75 VALU
// instructions per step and entry against 1.5).  Used only when every finite sum
// A + B lies below kCap16 (the caller's bound), so a result at the cap means
// unreachable and becomes FAP_INF.
template <int KP>
__device__ __forceinline__ void stream_row16(int (&acc)[4], const int (&ai)[KP],
                                             const unsigned (&bp)[KP][2]) {
  unsigned a[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) a[k] = min((unsigned)ai[k], (unsigned)kCap16);
  unsigned s0[KP], s1[KP];  // all sums first: a minimum right after its sum waits a cycle
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    s0[k] = pk_add_lo(a[k], bp[k][0]);
    s1[k] = pk_add_lo(a[k], bp[k][1]);
  }
  unsigned m0 = kPk16Top, m1 = kPk16Top;
#pragma unroll
  for (int k = 0; k < KP; k += 2) {
    m0 = pk_min3(m0, s0[k], s0[k + 1]);
    m1 = pk_min3(m1, s1[k], s1[k + 1]);
  }
  const unsigned v[4] = {m0 & 0xFFFFu, m0 >> 16, m1 & 0xFFFFu, m1 >> 16};
#pragma unroll
  for (int e = 0; e < 4; ++e) acc[e] = v[e] >= (unsigned)kCap16 ? kInf : (int)v[e];
}

// Columns j + 64 e.  Non-temporal: on MI355X the merge measured 1 ms slower with
// plain stores.
__device__ __forceinline__ void stream_store(FAP_GLOBAL int *crow, int j, int N,
                                             const int (&acc)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (j + 64 * e < N) __builtin_nontemporal_store(acc[e], crow + j + 64 * e);
}

template <int KP, bool kOwn>
__device__ __forceinline__ void stream_own(int (&acc)[4], const MergeDesc &d, int i,
                                           const int (&own)[4]) {
  if (kOwn) {
    const FAP_GLOBAL int *f = gptr(d.fw + (int64_t)(d.fw_row0 + i) * d.ldf);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (own[e] >= 0) acc[e] = min(acc[e], f[own[e]]);
  }
}

// A rows are prefetched one row ahead (scalar loads of row i + 1 in flight while
// row i is computed and stored); the prefetch index is clamped to the last row, so it
// never reads past the product.
template <int KP>
__device__ __forceinline__ void load_arow(int (&a)[KP], const FAP_CONST int *arow) {
#pragma unroll
  for (int k = 0; k < KP; ++k) a[k] = arow[k];
}

template <int KP, bool kOwn>
__device__ __forceinline__ void stream_rows(const MergeDesc &d, int row0, int row1, int j, int N,
                                            const int (&b)[KP][4], const int (&own)[4]) {
  FAP_GLOBAL int *crow = gptr(d.C + (int64_t)row0 * d.ldc);
  const FAP_CONST int *abase = uniform_cptr(d.A);
  int an[KP];
  load_arow<KP>(an, abase + (int64_t)row0 * d.lda);
  for (int i = row0; i < row1; ++i, crow += d.ldc) {
    int a[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) a[k] = an[k];
    load_arow<KP>(an, abase + (int64_t)min(i + 1, row1 - 1) * d.lda);
    int acc[4];
    stream_row<KP>(acc, a, b);
    stream_own<KP, kOwn>(acc, d, i, own);
    stream_store(crow, j, N, acc);
  }
}

template <int KP, bool kOwn>
__device__ __forceinline__ void stream_rows16(const MergeDesc &d, int row0, int row1, int j,
                                              int N, const unsigned (&bp)[KP][2],
                                              const int (&own)[4]) {
  FAP_GLOBAL int *crow = gptr(d.C + (int64_t)row0 * d.ldc);
  const FAP_CONST int *abase = uniform_cptr(d.A);
  int an[KP];
  load_arow<KP>(an, abase + (int64_t)row0 * d.lda);
  for (int i = row0; i < row1; ++i, crow += d.ldc) {
    int a[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) a[k] = an[k];
    load_arow<KP>(an, abase + (int64_t)min(i + 1, row1 - 1) * d.lda);
    int acc[4];
    stream_row16<KP>(acc, a, bp);
    stream_own<KP, kOwn>(acc, d, i, own);
    stream_store(crow, j, N, acc);
  }
}

// max_b / lim16: when max_b is set and *max_b <= lim16 (every finite B below
// lim16 + 1, every finite A at most kCap16 - 1 - lim16, both guaranteed by the
// caller), the rows run on packed 16-bit distances.
template <int KP>
__global__ __launch_bounds__(64) void merge_stream_kernel(const MergeDesc *__restrict__ descs,
                                                          int nd, int total_units, int N,
                                                          const int *__restrict__ graph_id,
                                                          const int *__restrict__ ed2st,
                                                          const int *__restrict__ max_b,
                                                          int lim16, int rpu, int col_major) {
  const int t = xcd_logical_tile();
  if (t >= total_units) return;
  const MergeDesc d = descs[find_desc(descs, nd, t)];
  const int local = t - d.tile_begin;
  int cb, rb;
  if (col_major) {
    const int ncb = (N + kStCols - 1) / kStCols;
    rb = local / ncb;
    cb = local - rb * ncb;
  } else {
    cb = local / d.row_blocks;
    rb = local - cb * d.row_blocks;
  }
  const int row0 = rb * rpu;
  const int row1 = min(row0 + rpu, d.M);
  const int j = cb * kStCols + threadIdx.x;  // columns j + 64 e

  int b[KP][4];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    if (k < d.K) {
      const FAP_GLOBAL int *brow = gptr(d.B + (int64_t)k * d.ldb);
#pragma unroll
      for (int e = 0; e < 4; ++e) b[k][e] = (j + 64 * e < N) ? brow[j + 64 * e] : kInf;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) b[k][e] = kInf;
    }
  }
  int own[4];
  bool any_own = false;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    own[e] = -1;
    if (d.part_id > 0 && j + 64 * e < N && graph_id[j + 64 * e] == d.part_id) {
      own[e] = ed2st[j + 64 * e] - d.offset;
      any_own = true;
    }
  }
  const bool packed = max_b != nullptr && *max_b <= lim16;  // uniform
  // Retire the B-panel and column loads here, so the row loop never waits on them.
  __builtin_amdgcn_s_waitcnt(0);
  const bool none_own = __builtin_amdgcn_ballot_w64(any_own) == 0;
  if (packed) {
    unsigned bp[KP][2];
#pragma unroll
    for (int k = 0; k < KP; ++k)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        bp[k][h] = (unsigned)min(b[k][2 * h], kCap16) |
                   ((unsigned)min(b[k][2 * h + 1], kCap16) << 16);
    if (none_own)
      stream_rows16<KP, false>(d, row0, row1, j, N, bp, own);
    else
      stream_rows16<KP, true>(d, row0, row1, j, N, bp, own);
  } else if (none_own) {
    stream_rows<KP, false>(d, row0, row1, j, N, b, own);
  } else {
    stream_rows<KP, true>(d, row0, row1, j, N, b, own);
  }
}

// Write probe with the streaming merge's store pattern: units of `rows` rows x 256
// columns, the column blocks of one row block back to back, the merge's XCD remap, one
// wave per unit writing one 1 KB row segment (four 256-byte runs of non-temporal dword
// stores) per row.  Only every `stride`-th row block is written.  Its duration on a
// buffer tells how fast the merge will write that buffer: on MI355X the same merge
// measured 9.0 ms in some 52.5 GB ranges of device memory and 10.6 ms in others.
__global__ __launch_bounds__(64) void write_probe_kernel(int *__restrict__ C, int64_t ldc,
                                                         int M, int N, int rows, int stride,
                                                         int total_units, int value) {
  const int t = xcd_logical_tile();
  if (t >= total_units) return;
  const int ncb = (N + kStCols - 1) / kStCols;
  const int rbi = t / ncb;
  const int cb = t - rbi * ncb;
  const int64_t row0 = (int64_t)rbi * stride * rows;
  const int row1 = (int)min<int64_t>(row0 + rows, M);
  const int j = cb * kStCols + threadIdx.x;
  FAP_GLOBAL int *crow = gptr(C + row0 * ldc);
  for (int i = (int)row0; i < row1; ++i, crow += ldc)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (j + 64 * e < N) __builtin_nontemporal_store(value + i, crow + j + 64 * e);
}

// Row blocks the placement probe of fap_hip_malloc_placed writes: one in 8 (on MI355X
// about 1 ms for a 52.5 GB buffer, and the same ordering of buffers as the full merge:
// 9.0 against 10.6 ms when its probe read 0.93-0.97 against 1.17-1.20 ms).
constexpr int kPlacementSample = 8;

int launch_write_probe(int *d, int64_t ld, int M, int N, int rows, int stride, int value,
                       hipStream_t stream) {
  if (M <= 0 || N <= 0) return FAP_OK;
  const int64_t rbs = ((int64_t)M + rows - 1) / rows;
  const int64_t units = (rbs + stride - 1) / stride * ((N + kStCols - 1) / kStCols);
  FAP_REQUIRE(units < (1LL << 30), "write probe: too many units");
  const int grid = (int)((units + 7) / 8 * 8);
  hipLaunchKernelGGL(write_probe_kernel, dim3(grid), dim3(64), 0, stream, d, ld, M, N, rows,
                     stride, (int)units, value);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int merge_stream_units(int M, int N, const StreamShape &sh) {
  if (M <= 0 || N <= 0) return 0;
  const int r = sh.rows;
  return ((M + r - 1) / r) * ((N + kStCols - 1) / kStCols);
}
int merge_stream_row_blocks(int M, const StreamShape &sh) { return (M + sh.rows - 1) / sh.rows; }
// K rounded up to a multiple of 2 (the pairs of v_min3 / v_pk_minimum3): against a
// multiple of 4, the merge stage of the 52.5 GB result measured 10.39 ms against 10.61
// in one process and buffer (K of the headline parts: 2..23, mean 10).
int merge_stream_kpad(int K) { return K <= 0 ? 2 : (K + 1) / 2 * 2; }

// kp: padded K, a multiple of 2 in [2, 32].
int launch_merge_stream(const MergeDesc *d_descs, int nd, int total_units, int N, int kp,
                        const int *d_graph_id, const int *d_ed2st, const int *d_max_b, int lim16,
                        const StreamShape &sh, hipStream_t stream) {
  if (nd <= 0 || total_units <= 0 || N <= 0) return FAP_OK;
  const int grid = (total_units + 7) / 8 * 8;
  // Resident-wave cap: each one-wave workgroup reserves 1/waves_per_cu of the CU's
  // 160 KB LDS (the kernel itself uses none).
  const size_t lds = sh.waves_per_cu > 0 ? (size_t)(163840 / sh.waves_per_cu) / 1024 * 1024 : 0;
#define FAP_STREAM_CASE(KP)                                                                   \
  case KP:                                                                                    \
    hipLaunchKernelGGL((merge_stream_kernel<KP>), dim3(grid), dim3(64), lds, stream, d_descs,  \
                       nd, total_units, N, d_graph_id, d_ed2st, d_max_b, lim16, sh.rows,      \
                       sh.col_major);                                                          \
    break;
  switch (kp) {
    FAP_STREAM_CASE(2)
    FAP_STREAM_CASE(4)
    FAP_STREAM_CASE(6)
    FAP_STREAM_CASE(8)
    FAP_STREAM_CASE(10)
    FAP_STREAM_CASE(12)
    FAP_STREAM_CASE(14)
    FAP_STREAM_CASE(16)
    FAP_STREAM_CASE(18)
    FAP_STREAM_CASE(20)
    FAP_STREAM_CASE(22)
    FAP_STREAM_CASE(24)
    FAP_STREAM_CASE(26)
    FAP_STREAM_CASE(28)
    FAP_STREAM_CASE(30)
    FAP_STREAM_CASE(32)
    default:
      set_error("launch_merge_stream: unsupported padded K %d", kp);
      return FAP_ERR_INVALID;
  }
#undef FAP_STREAM_CASE
  FAP_CHECK_LAUNCH();
  return FAP_OK;
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

extern "C" int fap_hip_set_merge_shape(int32_t rows, int32_t unit_order, int32_t waves_per_cu) {
  if (rows < 8 || rows > 1024 || (rows & 7) || unit_order < 0 || unit_order > 1 ||
      waves_per_cu < 0 || (waves_per_cu > 0 && waves_per_cu < 4) || waves_per_cu > 32) {
    set_error("fap_hip_set_merge_shape: rows %d (8..1024, multiple of 8), order %d (0/1), "
              "waves per CU %d (0 or 4..32)", rows, unit_order, waves_per_cu);
    return FAP_ERR_INVALID;
  }
  g_stream_shape.store(rows | ((int64_t)unit_order << 16) | ((int64_t)waves_per_cu << 24),
                       std::memory_order_relaxed);
  return FAP_OK;
}

extern "C" int fap_hip_get_merge_shape(int32_t *rows, int32_t *unit_order, int32_t *waves_per_cu) {
  const StreamShape sh = stream_shape();
  if (rows) *rows = sh.rows;
  if (unit_order) *unit_order = sh.col_major;
  if (waves_per_cu) *waves_per_cu = sh.waves_per_cu;
  return FAP_OK;
}

extern "C" int fap_hip_write_probe(int32_t *d, int64_t ld, int32_t rows, int32_t cols,
                                   int32_t sample_stride, void *stream, float *ms) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(d && rows >= 0 && cols >= 0 && ld >= cols && sample_stride >= 1 && ms,
              "fap_hip_write_probe: invalid arguments");
  const hipStream_t st = resolve(stream);
  hipEvent_t e0, e1;
  FAP_CHECK_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("fap_hip_write_probe: event creation failed");
    return FAP_ERR_HIP;
  }
  int s = hipEventRecord(e0, st) == hipSuccess ? FAP_OK : FAP_ERR_HIP;
  if (s == FAP_OK) s = launch_write_probe(d, ld, rows, cols, stream_shape().rows, sample_stride, 0, st);
  if (s == FAP_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                      hipEventElapsedTime(ms, e0, e1) != hipSuccess)) {
    set_error("fap_hip_write_probe: timing failed");
    s = FAP_ERR_HIP;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return s;
}

extern "C" int fap_hip_malloc_placed(void **d_ptr, int64_t rows, int64_t ld, int32_t candidates,
                                     float *probe_ms) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(d_ptr && rows > 0 && ld > 0 && rows < (1LL << 31) && ld < (1LL << 31) &&
                  candidates >= 1 && candidates <= 8,
              "fap_hip_malloc_placed: invalid arguments");
  *d_ptr = nullptr;
  const int64_t bytes = rows * ld * 4;
  const hipStream_t st = resolve(nullptr);
  std::vector<std::pair<void *, float>> got;
  auto release = [&](void *keep) {
    for (auto &x : got)
      if (x.first != keep) (void)dev_free(x.first);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (candidates > 1) {
    FAP_CHECK_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      set_error("fap_hip_malloc_placed: event creation failed");
      return FAP_ERR_HIP;
    }
  }
  int status = FAP_OK;
  for (int c = 0; c < candidates && status == FAP_OK; ++c) {
    void *p = nullptr;
    const int s = dev_alloc(&p, bytes);
    if (s != FAP_OK) {
      if (c == 0) status = s;  // not even one buffer: the caller's error
      else clear_error();      // no room for another candidate: choose among those held
      break;
    }
    float best = 0.0f;
    if (candidates > 1) {
      best = 1e30f;
      for (int rep = 0; rep < 2 && status == FAP_OK; ++rep) {
        float ms = 0.0f;
        if (hipEventRecord(e0, st) != hipSuccess) status = FAP_ERR_HIP;
        if (status == FAP_OK)
          status = launch_write_probe(static_cast<int *>(p), ld, (int)rows, (int)ld,
                                      stream_shape().rows, kPlacementSample, 0, st);
        if (status == FAP_OK &&
            (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
             hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
          set_error("fap_hip_malloc_placed: probe timing failed");
          status = FAP_ERR_HIP;
        }
        best = std::min(best, ms);
      }
    }
    got.push_back({p, best});
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (status != FAP_OK) {
    release(nullptr);
    return status;
  }
  size_t pick = 0;
  for (size_t k = 1; k < got.size(); ++k)
    if (got[k].second < got[pick].second) pick = k;
  release(got[pick].first);
  *d_ptr = got[pick].first;
  if (probe_ms) *probe_ms = got[pick].second;
  return FAP_OK;
}

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
