// Float-weight path of the part pipeline, computed in FP64, results in float32.
//
// The reference computes everything in float (dijkstra_path, floyd_CPU_path and
// min_plus_path_CPU_advanced on float matrices, fap/kernel/batch_sssp.h:79-114,
// floyd.h:38-49, minplus.h:51-64).  Here every distance is carried in double: a
// float32 weight is an integer multiple of its ulp, so as long as a path sum spans
// fewer than 53 significant bits (weights in [0.5, 10) and sums below 2^29, for
// instance) every sum is exact and the results equal an FP64 Dijkstra bit for bit;
// only the final conversion to float32 rounds (relative error <= 2^-24).  In every
// other case the FP64 rounding is far below float's.  Unreachable entries, and
// distances of 1e8 or more, are written as MAXVALUE = 1e8 (fap/parameter.h:26),
// which is what the reference's float path reports for them.
//
// Kernels (each the FP64 twin of an int32 kernel of this library):
//   sssp64_kernel   one wave per boundary source, label-correcting relaxation with
//                   atomicMin on the bit patterns (non-negative doubles and +inf
//                   order as unsigned integers), frontier queues in LDS;
//   build64_kernel  one workgroup per padded row of a part's matrix;
//   fw64_phase*     blocked three-phase Floyd-Warshall on 64 x 64 tiles;
//   merge64_kernel  LDS-tiled min-plus of the inner rows against the boundary rows
//                   with the decode of the part's own columns fused, float output;
//   bdy64_kernel    the boundary rows: SSSP row, own columns from the FW matrix.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"

namespace fap_hip {

int launch_gather_sources(const int *offsets, const int *src_begin, const int64_t *out_rows,
                          int nparts, int total, const int *st2ed, int64_t ld, int *src,
                          int64_t *rows, hipStream_t stream);

namespace {

typedef unsigned long long u64;
constexpr double kInfD = __builtin_inf();
constexpr float kMaxValue = 1e8f;
constexpr int kT = 64;
constexpr int kBuildChunk64 = 8192;  // LDS row buffer: 64 KB

__device__ __forceinline__ u64 dbits(double x) { return (u64)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(u64 x) { return __longlong_as_double((long long)x); }
__device__ __forceinline__ float to_out(double d) { return d < 1e8 ? (float)d : kMaxValue; }

__device__ __forceinline__ double load_agent(const double *p) {
  return bitsd(__hip_atomic_load(reinterpret_cast<const u64 *>(p), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- SSSP
constexpr int kSWaves = 4;
constexpr int kSQ = 1024;

// Relaxes every out-arc of u with its current label.
__device__ __forceinline__ void relax64(const int *__restrict__ row_off, const int *__restrict__ col,
                                        const float *__restrict__ w, double *row, int u,
                                        int *q_next, int *q_count) {
  const double du = load_agent(row + u);
  if (!(du < kInfD)) return;
  const int e1 = row_off[u + 1];
  for (int e = row_off[u]; e < e1; ++e) {
    const int v = col[e];
    const double nd = du + (w ? (double)w[e] : 1.0);
    if (nd < row[v]) {
      const u64 old = atomicMin(reinterpret_cast<u64 *>(row + v), dbits(nd));
      if (nd < bitsd(old)) {
        const int pos = atomicAdd(q_count, 1);
        if (pos < kSQ) q_next[pos] = v;
      }
    }
  }
}

// One wave per source; row i of `rows` (stride ld) belongs to source src[i].
__global__ __launch_bounds__(64 * kSWaves) void sssp64_kernel(
    const int *__restrict__ row_off, const int *__restrict__ col, const float *__restrict__ w,
    int n, const int *__restrict__ src, int nsrc, double *__restrict__ rows, int64_t ld) {
  __shared__ int queue[kSWaves][2][kSQ];
  __shared__ int qcount[kSWaves][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sidx = blockIdx.x * kSWaves + wave;
  if (sidx >= nsrc) return;  // whole wave leaves; no workgroup barrier below
  const int s = src[sidx];
  double *row = rows + (int64_t)sidx * ld;
  for (int v = lane; v < n; v += 64) row[v] = (v == s) ? 0.0 : kInfD;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // fill visible before the atomics
  if (lane == 0) {
    queue[wave][0][0] = s;
    qcount[wave][0] = 1;
    qcount[wave][1] = 0;
  }
  wave_sync();
  int cur = 0, qsize = 1;
  bool scan = false;
  while (true) {
    const int nxt = cur ^ 1;
    int *q_next = queue[wave][nxt];
    int *q_cnt = &qcount[wave][nxt];
    if (!scan) {
      for (int i = lane; i < qsize; i += 64) relax64(row_off, col, w, row, queue[wave][cur][i], q_next, q_cnt);
    } else {
      for (int u = lane; u < n; u += 64) relax64(row_off, col, w, row, u, q_next, q_cnt);
    }
    wave_sync();
    const int cnt = __hip_atomic_load(q_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync();
    if (cnt == 0) break;
    scan = cnt > kSQ;  // overflowed queue: next round relaxes every vertex
    qsize = scan ? 0 : cnt;
    if (lane == 0) qcount[wave][cur] = 0;
    wave_sync();
    cur = nxt;
  }
}

// ---------------------------------------------------------------- build
struct Build64Desc {
  double *fw;
  int S, s, bdy, offset, part_id, row_begin;
  const double *rows;  // the part's SSSP rows (row i = boundary vertex i)
  int64_t ld_rows;
};

template <typename D>
__device__ __forceinline__ int find_begin(const D *__restrict__ d, int nd, int64_t key,
                                          int64_t (*get)(const D &)) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (get(d[mid]) <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ int64_t build_begin(const Build64Desc &d) { return d.row_begin; }

// Row i of part p: boundary rows from the SSSP rows gathered at st2ed, inner rows
// FAP_INF then the part's own out-arcs in CSR order (a later arc overwrites an
// earlier one), diagonal 0, padding +inf (MysubMatBuild_path,
// fap/utils/util_Myalgorithm.h:44-84).
__global__ __launch_bounds__(256) void build64_kernel(const Build64Desc *__restrict__ descs, int nd,
                                                      int total_rows, const int *__restrict__ row_off,
                                                      const int *__restrict__ col,
                                                      const float *__restrict__ w,
                                                      const int *__restrict__ graph_id,
                                                      const int *__restrict__ st2ed,
                                                      const int *__restrict__ ed2st) {
  const int grow = blockIdx.x;
  if (grow >= total_rows) return;
  const Build64Desc d = descs[find_begin(descs, nd, grow, build_begin)];
  const int i = grow - d.row_begin;
  __shared__ double rowbuf[kBuildChunk64];
  const int tid = threadIdx.x;
  const bool from_rows = i < d.s && i < d.bdy;
  for (int c0 = 0; c0 < d.S; c0 += kBuildChunk64) {  // rows in chunks of LDS size
    const int cw = min(kBuildChunk64, d.S - c0);
    if (c0 > 0) __syncthreads();
    if (from_rows) {
      const double *src = d.rows + (int64_t)i * d.ld_rows;
      for (int j = tid; j < cw; j += 256)
        rowbuf[j] = (c0 + j < d.s) ? src[st2ed[d.offset + c0 + j]] : kInfD;
    } else {
      for (int j = tid; j < cw; j += 256) rowbuf[j] = kInfD;
    }
    __syncthreads();
    if (tid == 0 && i < d.s) {
      if (!from_rows) {
        const int ver = st2ed[d.offset + i];
        for (int e = row_off[ver]; e < row_off[ver + 1]; ++e) {
          const int nb = col[e];
          if (graph_id[nb] != d.part_id) continue;
          const int idx = ed2st[nb] - d.offset - c0;
          if (idx >= 0 && idx < cw && c0 + idx < d.s) rowbuf[idx] = w ? (double)w[e] : 1.0;
        }
      }
      if (i >= c0 && i < c0 + cw) rowbuf[i - c0] = 0.0;
    }
    __syncthreads();
    double *dst = d.fw + (int64_t)i * d.S + c0;
    for (int j = tid; j < cw; j += 256) dst[j] = rowbuf[j];
  }
}

// ---------------------------------------------------------------- Floyd-Warshall
// Tiles are addressed in the global address space (the descriptor pointer is generic;
// flat accesses would also hold up the LDS counter of the staging writes).
__device__ __forceinline__ FAP_GLOBAL double *tile64(const Fw64Desc &d, int ti, int tj) {
  const int64_t S = (int64_t)d.nb * kT;
  return gptr(d.mat + (int64_t)ti * kT * S + tj * kT);
}

// Phase 1: the diagonal tile closed by 256 threads.  Wave w holds rows w, w + 4, ...,
// w + 60 in registers, lane j their column j.  At step k, column k (the entries
// D[i][k] of a wave's rows) comes from lane k by readlane and row k from LDS, where the
// wave owning row k published it after step k - 1 (two buffers, one barrier per step).
// Row k and column k do not change during step k (D[k][k] = 0, or the padded row and
// column are +inf throughout), so this is the k-i-j loop.  An earlier form with 16
// consecutive entries per thread and a 64-double LDS row stride hit every row of a wave
// on the same banks.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__global__ __launch_bounds__(256) void fw64_phase1_kernel(const Fw64Desc *__restrict__ descs, int r) {
  const Fw64Desc d = descs[blockIdx.x];
  if (r >= d.nb) return;
  const int64_t S = (int64_t)d.nb * kT;
  __shared__ double Rk[2][kT];
  FAP_GLOBAL double *g = tile64(d, r, r);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = g[(wave + 4 * q) * S + lane];
  if (wave == 0) Rk[0][lane] = v[0];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kT; ++k) {
    const double rk = Rk[k & 1][lane];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = fmin(v[q], readlane_f64(v[q], k) + rk);
    if (k + 1 < kT && wave == ((k + 1) & 3)) Rk[(k + 1) & 1][lane] = v[(k + 1) >> 2];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) g[(wave + 4 * q) * S + lane] = v[q];
}

// Phase 3 keeps the LDS-tiled form: one 256-thread workgroup per 64 x 64 tile, thread t
// owns rows (t / 16) * 4 + {0..3} and columns (t % 16) + 16 * {0..3} (4 x 4
// accumulators), k in LDS chunks of 16: As[k][i] (transposed A, four consecutive rows
// read as two 16-byte broadcasts), Bs[k][j] (16 consecutive lanes read 16 consecutive
// doubles: conflict-free).  Its A tile differs per tile row, so scalar loads of A (the
// phase-2 form below) miss the scalar cache: measured 2x slower here.
constexpr int kKc64 = 16;
// As rows padded by 16 bytes: the staging writes (16 lanes on 16 consecutive k of one
// row) land in 16 distinct bank pairs instead of one.
constexpr int kLdA64 = kT + 2;

__device__ __forceinline__ void tile_product64(double (&acc)[4][4], const FAP_GLOBAL double *a,
                                               const FAP_GLOBAL double *b, int64_t S,
                                               double (*As)[kLdA64], double (*Bs)[kT], int tr,
                                               int tc) {
  const int t = threadIdx.x;
  for (int k0 = 0; k0 < kT; k0 += kKc64) {
    if (k0 > 0) __syncthreads();
    // 16 x 64 entries of each operand, four per thread: A row-major (i = e / 16),
    // B row-major (k = e / 64).
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = t + 256 * q;
      const int ai = e >> 4, ak = e & 15;
      As[ak][ai] = a[ai * S + k0 + ak];
      const int bk = e >> 6, bj = e & 63;
      Bs[bk][bj] = b[(k0 + bk) * S + bj];
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kKc64; ++k) {
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = As[k][tr * 4 + q];
        bv[q] = Bs[k][tc + 16 * q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = fmin(acc[q][c], av[q] + bv[c]);
    }
  }
}

__device__ __forceinline__ void load_tile64(double (&acc)[4][4], const FAP_GLOBAL double *t,
                                            int64_t S, int tr, int tc) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[q][c] = t[(tr * 4 + q) * S + tc + 16 * c];
}

__device__ __forceinline__ void store_tile64(const double (&acc)[4][4], FAP_GLOBAL double *t,
                                             int64_t S, int tr, int tc) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) t[(tr * 4 + q) * S + tc + 16 * c] = acc[q][c];
}

// Phase 2: C = min(C, A (x) B) on one 64 x 64 tile per 512-thread workgroup.
// Wave w owns rows 8w .. 8w + 7 of C, lane l column l (8 accumulators).  B (64 x 64)
// is staged in LDS and read as Bs[k][lane] (one conflict-free 8-byte read per k for
// 16 VALU instructions); the A entries of a wave's rows are wave-uniform and come
// from scalar loads (SGPR operands of v_add_f64), two k ahead of their use.  The
// earlier form (4 x 4 accumulators, A and B both from LDS) spent as many LDS cycles
// as VALU cycles per k.  Tiles of one part run back to back on one XCD (its L2 keeps
// the panels).
constexpr int kFwRows = 8;    // rows of C per wave
constexpr int kFwWaves = 8;   // waves per workgroup (kFwRows * kFwWaves = 64)
constexpr int kFwLdB = kT;    // Bs row stride (doubles)

// acc[q] = min(acc[q], min_k A[q][k] + Bs[k][lane]) over the 64 k of the tile, where
// `a` points at row 0 of the wave's A rows (stride S, wave-uniform).
__device__ __forceinline__ void tile_rows64(double (&acc)[kFwRows], const FAP_CONST double *a,
                                           int64_t S, const double (*Bs)[kFwLdB], int lane) {
  double an[kFwRows][2];
#pragma unroll
  for (int q = 0; q < kFwRows; ++q) {
    an[q][0] = a[q * S + 0];
    an[q][1] = a[q * S + 1];
  }
#pragma unroll 2
  for (int k = 0; k < kT; k += 2) {
    double ac[kFwRows][2];
#pragma unroll
    for (int q = 0; q < kFwRows; ++q) {
      ac[q][0] = an[q][0];
      ac[q][1] = an[q][1];
    }
    // Next pair (the last pair again at the end: an in-bounds, unused load).
    const int kn = k + 2 < kT ? k + 2 : k;
#pragma unroll
    for (int q = 0; q < kFwRows; ++q) {
      an[q][0] = a[q * S + kn];
      an[q][1] = a[q * S + kn + 1];
    }
    const double b0 = Bs[k][lane], b1 = Bs[k + 1][lane];
#pragma unroll
    for (int q = 0; q < kFwRows; ++q) acc[q] = fmin(acc[q], ac[q][0] + b0);
#pragma unroll
    for (int q = 0; q < kFwRows; ++q) acc[q] = fmin(acc[q], ac[q][1] + b1);
  }
}

// Stages the 64 x 64 tile t (row stride S) into Bs: 512 threads, 8 doubles each.
__device__ __forceinline__ void stage_tile64(double (*Bs)[kFwLdB], const FAP_GLOBAL double *t,
                                             int64_t S) {
  const int e = threadIdx.x, row = e >> 3, c0 = (e & 7) * 8;
  typedef double d2 __attribute__((ext_vector_type(2)));
  const FAP_GLOBAL d2 *src = reinterpret_cast<const FAP_GLOBAL d2 *>(t + row * S + c0);
  d2 *dst = reinterpret_cast<d2 *>(&Bs[row][c0]);
  d2 v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = src[q];
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[q] = v[q];
}

__device__ __forceinline__ int skip64(int o, int p) { return o < p ? o : o + 1; }

// One launch of phase 2 or 3 over every descriptor: grid (tiles per descriptor,
// descriptors), remapped so each XCD runs a contiguous range of (descriptor, tile).
//   phase 2, row panel T(r, j):    A = pivot tile (r, r), B = T(r, j) itself;
//   phase 2, column panel T(i, r): A = T(i, r) itself, B = pivot tile (r, r);
//   phase 3, T(i, j):              A = T(i, r), B = T(r, j).
// A tile that is its own operand is safe: as B it is read from the LDS copy, and as A
// each wave reads only the rows it writes itself, before it writes them.
template <int kPhase>
__global__ __launch_bounds__(64 * kFwWaves) void fw64_tile_kernel(const Fw64Desc *__restrict__ descs,
                                                                  int r) {
  const unsigned L = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int o = (int)(L % gridDim.x), di = (int)(L / gridDim.x);
  const Fw64Desc d = descs[di];
  const int others = d.nb - 1;
  if (r >= d.nb || o >= (kPhase == 2 ? 2 * others : others * others)) return;
  const int64_t S = (int64_t)d.nb * kT;
  int ci, cj, ai, aj, bi, bj;
  if (kPhase == 2) {
    if (o < others) {  // row panel
      ci = r, cj = skip64(o, r), ai = r, aj = r, bi = ci, bj = cj;
    } else {           // column panel
      ci = skip64(o - others, r), cj = r, ai = ci, aj = cj, bi = r, bj = r;
    }
  } else {
    ci = skip64(o / others, r), cj = skip64(o % others, r), ai = ci, aj = r, bi = r, bj = cj;
  }
  __shared__ __attribute__((aligned(16))) double Bs[kT][kFwLdB];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  FAP_GLOBAL double *c = tile64(d, ci, cj) + (int64_t)(wave * kFwRows) * S + lane;
  stage_tile64(Bs, tile64(d, bi, bj), S);
  double acc[kFwRows];
#pragma unroll
  for (int q = 0; q < kFwRows; ++q) acc[q] = c[q * S];
  __syncthreads();
  const FAP_CONST double *a =
      uniform_cptr(d.mat + ((int64_t)ai * kT + wave * kFwRows) * S + (int64_t)aj * kT);
  tile_rows64(acc, a, S, Bs, lane);
#pragma unroll
  for (int q = 0; q < kFwRows; ++q) c[q * S] = acc[q];
}


// Phase 3 with the LDS-tiled product; tiles of one part back to back on one XCD.
__global__ __launch_bounds__(256) void fw64_phase3_kernel(const Fw64Desc *__restrict__ descs, int r) {
  const unsigned L = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int o = (int)(L % gridDim.x), di = (int)(L / gridDim.x);
  const Fw64Desc d = descs[di];
  const int others = d.nb - 1;
  if (r >= d.nb || o >= others * others) return;
  const int i = skip64(o / others, r), j = skip64(o % others, r);
  const int64_t S = (int64_t)d.nb * kT;
  __shared__ __attribute__((aligned(16))) double As[kKc64][kLdA64];
  __shared__ __attribute__((aligned(16))) double Bs[kKc64][kT];
  FAP_GLOBAL double *t = tile64(d, i, j);
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4];
  load_tile64(acc, t, S, tr, tc);
  tile_product64(acc, tile64(d, i, r), tile64(d, r, j), S, As, Bs, tr, tc);
  store_tile64(acc, t, S, tr, tc);
}

}  // namespace

namespace {

// ---- 32 x 32 tiles: the Floyd-Warshall tile of the batch's FP64 parts when the
// autotuner picks it (parts padded to multiples of 32 instead of 64: a part of 33-64
// vertices does 1/8 of the padded work).  Same three phases on smaller tiles.
constexpr int kT32 = 32;

__device__ __forceinline__ FAP_GLOBAL double *tile32(const Fw64Desc &d, int ti, int tj) {
  const int64_t S = (int64_t)d.nb * kT32;
  return gptr(d.mat + (int64_t)ti * kT32 * S + tj * kT32);
}

// Phase 1: 128 threads; wave w holds rows w, w + 2, ..., w + 30 in registers, lane j
// column j & 31 (lanes 32..63 mirror lanes 0..31 and store nothing).  Column k comes
// from lane k by readlane, row k from LDS (published by its wave after step k - 1).
__global__ __launch_bounds__(128) void fw32_phase1_kernel(const Fw64Desc *__restrict__ descs,
                                                          int r) {
  const Fw64Desc d = descs[blockIdx.x];
  if (r >= d.nb) return;
  const int64_t S = (int64_t)d.nb * kT32;
  __shared__ double Rk[2][kT32];
  FAP_GLOBAL double *g = tile32(d, r, r);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = lane & 31;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = g[(wave + 2 * q) * S + col];
  if (wave == 0 && lane < kT32) Rk[0][lane] = v[0];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kT32; ++k) {
    const double rk = Rk[k & 1][col];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = fmin(v[q], readlane_f64(v[q], k) + rk);
    if (k + 1 < kT32 && wave == ((k + 1) & 1) && lane < kT32) Rk[(k + 1) & 1][lane] = v[(k + 1) >> 1];
    __syncthreads();
  }
  if (lane < kT32) {
#pragma unroll
    for (int q = 0; q < 16; ++q) g[(wave + 2 * q) * S + col] = v[q];
  }
}

// C = min(C, A (x) B) on 32 x 32 tiles, one wave: thread t owns rows (t / 8) * 4 +
// {0..3} and columns (t % 8) + 8 * {0..3}; k in LDS chunks of 16.  A tile that is its own
// operand is read (staged) before any thread stores.
__device__ __forceinline__ void tile_product32(double (&acc)[4][4], const FAP_GLOBAL double *a,
                                               const FAP_GLOBAL double *b, int64_t S,
                                               double (*As)[kT32 + 2], double (*Bs)[kT32],
                                               int tr, int tc) {
  const int t = threadIdx.x;
  for (int k0 = 0; k0 < kT32; k0 += 16) {
    if (k0 > 0) __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int e = t + 64 * q;
      const int ai = e >> 4, ak = e & 15;
      As[ak][ai] = a[ai * S + k0 + ak];
      const int bk = e >> 5, bj = e & 31;
      Bs[bk][bj] = b[(k0 + bk) * S + bj];
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      double av[4], bv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        av[q] = As[k][tr * 4 + q];
        bv[q] = Bs[k][tc + 8 * q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = fmin(acc[q][c], av[q] + bv[c]);
    }
  }
}

// Phase 2 (kPhase 2, grid x = 2 * (nb - 1)) and phase 3 (kPhase 3, grid x = (nb - 1)^2);
// the tiles of one part back to back on one XCD.
template <int kPhase>
__global__ __launch_bounds__(64) void fw32_tile_kernel(const Fw64Desc *__restrict__ descs, int r) {
  const unsigned L = xcd_contiguous(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const int o = (int)(L % gridDim.x), di = (int)(L / gridDim.x);
  const Fw64Desc d = descs[di];
  const int others = d.nb - 1;
  if (r >= d.nb || o >= (kPhase == 2 ? 2 * others : others * others)) return;
  const int64_t S = (int64_t)d.nb * kT32;
  int ci, cj, ai, aj, bi, bj;
  if (kPhase == 2) {
    if (o < others) {  // row panel: P* (x) T(r, j)
      ci = r, cj = skip64(o, r), ai = r, aj = r, bi = ci, bj = cj;
    } else {           // column panel: T(i, r) (x) P*
      ci = skip64(o - others, r), cj = r, ai = ci, aj = cj, bi = r, bj = r;
    }
  } else {
    ci = skip64(o / others, r), cj = skip64(o % others, r), ai = ci, aj = r, bi = r, bj = cj;
  }
  __shared__ __attribute__((aligned(16))) double As[16][kT32 + 2];
  __shared__ __attribute__((aligned(16))) double Bs[16][kT32];
  FAP_GLOBAL double *c = tile32(d, ci, cj);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  double acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[q][e] = c[(tr * 4 + q) * S + tc + 8 * e];
  tile_product32(acc, tile32(d, ai, aj), tile32(d, bi, bj), S, As, Bs, tr, tc);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) c[(tr * 4 + q) * S + tc + 8 * e] = acc[q][e];
}

}  // namespace

// The same rounds on 32 x 32 tiles (descriptors' nb counts 32-tiles).
int launch_fw64_t32(const Fw64Desc *d_descs, int nd, int max_nb, hipStream_t st) {
  for (int r = 0; r < max_nb; ++r) {
    hipLaunchKernelGGL(fw32_phase1_kernel, dim3(nd), dim3(128), 0, st, d_descs, r);
    FAP_CHECK_LAUNCH();
    const int others = max_nb - 1;
    if (others > 0) {
      hipLaunchKernelGGL(fw32_tile_kernel<2>, dim3(2 * others, nd), dim3(64), 0, st, d_descs, r);
      FAP_CHECK_LAUNCH();
      hipLaunchKernelGGL(fw32_tile_kernel<3>, dim3(others * others, nd), dim3(64), 0, st, d_descs,
                         r);
      FAP_CHECK_LAUNCH();
    }
  }
  return FAP_OK;
}

// Phase 2 reads the row tile it writes only through its own workgroup and the
// pivot tile, which phase 2 does not write; phase 3 reads panel tiles only.
int launch_fw64(const Fw64Desc *d_descs, int nd, int max_nb, hipStream_t st) {
  for (int r = 0; r < max_nb; ++r) {
    hipLaunchKernelGGL(fw64_phase1_kernel, dim3(nd), dim3(256), 0, st, d_descs, r);
    FAP_CHECK_LAUNCH();
    const int others = max_nb - 1;
    if (others > 0) {
      hipLaunchKernelGGL(fw64_tile_kernel<2>, dim3(2 * others, nd), dim3(64 * kFwWaves), 0, st,
                         d_descs, r);
      FAP_CHECK_LAUNCH();
      hipLaunchKernelGGL(fw64_phase3_kernel, dim3(others * others, nd), dim3(256), 0, st, d_descs,
                         r);
      FAP_CHECK_LAUNCH();
    }
  }
  return FAP_OK;
}

namespace {

// ---------------------------------------------------------------- merge
// C[i][j] = min_k A[i][k] + B[k][j] (overwrite, +inf when K == 0) for the inner rows
// of one part, with the part's own columns taken from its FW matrix; float output.
struct Merge64Desc {
  const double *A;  // inner x boundary block of the FW matrix (stride lda)
  int64_t lda;
  const double *B;  // boundary SSSP rows (stride ldb)
  int64_t ldb;
  float *C;
  int64_t ldc;
  int M, K, row_blocks, tile_begin, part_id, offset;
  const double *fw;  // FW matrix; row fw_row0 + i holds C's row i
  int64_t ldf;
  int fw_row0;
};

constexpr int kMRows = 64, kMCols = 256, kMKc = 16;

__device__ int64_t merge_begin(const Merge64Desc &d) { return d.tile_begin; }

template <bool kVec>
__global__ __launch_bounds__(256) void merge64_kernel(const Merge64Desc *__restrict__ descs, int nd,
                                                      int total_tiles, int N,
                                                      const int *__restrict__ graph_id,
                                                      const int *__restrict__ ed2st) {
  const int t = blockIdx.x;
  if (t >= total_tiles) return;
  const Merge64Desc d = descs[find_begin(descs, nd, t, merge_begin)];
  const int local = t - d.tile_begin;
  const int cb = local / d.row_blocks, rb = local - cb * d.row_blocks;
  const int row0 = rb * kMRows, col0 = cb * kMCols;
  __shared__ double As[kMKc][kMRows];
  __shared__ double Bs[kMKc][kMCols];
  const int tid = threadIdx.x;
  const int tx = tid & 31;  // columns col0 + tx*4 + {0..3}, col0 + 128 + tx*4 + {0..3}
  const int ty = tid >> 5;  // rows row0 + ty*8 + {0..7}
  double acc[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[r][c] = kInfD;
  for (int k0 = 0; k0 < d.K; k0 += kMKc) {
    if (k0 > 0) __syncthreads();
    {  // As[kk][i] = A[row0 + i][k0 + kk]: 64 x 16, four per thread
      const int i = tid & 63, kq = (tid >> 6) * 4;
      const bool row_ok = row0 + i < d.M;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = kq + q;
        As[kk][i] = (row_ok && k0 + kk < d.K) ? gptr(d.A)[(int64_t)(row0 + i) * d.lda + k0 + kk] : kInfD;
      }
    }
    {  // Bs[kk][c] = B[k0 + kk][col0 + c]: 16 x 256, sixteen per thread
#pragma unroll
      for (int q = 0; q < kMKc; ++q) {
        const int c = tid, j = col0 + c;
        Bs[q][c] = (k0 + q < d.K && j < N) ? gptr(d.B)[(int64_t)(k0 + q) * d.ldb + j] : kInfD;
      }
    }
    __syncthreads();
    const int kc = min(kMKc, d.K - k0);
    for (int k = 0; k < kc; ++k) {
      double av[8], bv[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) av[r] = As[k][ty * 8 + r];
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int e = 0; e < 4; ++e) bv[4 * g + e] = Bs[k][g * 128 + tx * 4 + e];
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = fmin(acc[r][c], av[r] + bv[c]);
    }
  }
  int cols[8], local_col[8];
  bool any_own = false;
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = 4 * g + e;
      cols[c] = col0 + g * 128 + tx * 4 + e;
      local_col[c] = -1;
      if (d.part_id > 0 && cols[c] < N && graph_id[cols[c]] == d.part_id) {
        local_col[c] = ed2st[cols[c]] - d.offset;
        any_own = true;
      }
    }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int i = row0 + ty * 8 + r;
    if (i >= d.M) break;
    if (any_own) {
      const FAP_GLOBAL double *frow = gptr(d.fw + (int64_t)(d.fw_row0 + i) * d.ldf);
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (local_col[c] >= 0) acc[r][c] = frow[local_col[c]];
    }
    FAP_GLOBAL float *crow = gptr(d.C + (int64_t)i * d.ldc);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int j = cols[4 * g];
      if (kVec && j + 3 < N) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        *reinterpret_cast<FAP_GLOBAL v4f *>(crow + j) =
            v4f{to_out(acc[r][4 * g]), to_out(acc[r][4 * g + 1]), to_out(acc[r][4 * g + 2]),
                to_out(acc[r][4 * g + 3])};
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (j + e < N) crow[j + e] = to_out(acc[r][4 * g + e]);
      }
    }
  }
}

// ---------------------------------------------------------------- boundary rows
struct Bdy64Desc {
  const double *rows;  // the part's SSSP rows (stride ld_rows)
  int64_t ld_rows;
  const double *fw;    // FW matrix (stride ldf), row i = boundary vertex i
  int64_t ldf;
  float *out;          // first output row of the part (stride ld_out)
  int64_t ld_out;
  int rows_n, offset, part_id;
  int64_t item_begin;  // prefix over parts of rows_n * n
};

__device__ int64_t bdy_begin(const Bdy64Desc &d) { return d.item_begin; }

__global__ void bdy64_kernel(const Bdy64Desc *__restrict__ descs, int nd, int64_t total, int n,
                             const int *__restrict__ graph_id, const int *__restrict__ ed2st) {
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= total) return;
  const Bdy64Desc d = descs[find_begin(descs, nd, item, bdy_begin)];
  const int64_t local = item - d.item_begin;
  const int i = (int)(local / n), j = (int)(local - (int64_t)i * n);
  const double v = graph_id[j] == d.part_id ? d.fw[(int64_t)i * d.ldf + ed2st[j] - d.offset]
                                            : d.rows[(int64_t)i * d.ld_rows + j];
  d.out[(int64_t)i * d.ld_out + j] = to_out(v);
}

Workspace g_ws_rows, g_ws_fw, g_ws_meta;
int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

// The parts tile [0, n) (every vertex in exactly one of them).
static bool covers_graph(std::vector<fap_part_desc> act, int n) {
  std::sort(act.begin(), act.end(),
            [](const fap_part_desc &a, const fap_part_desc &b) { return a.offset < b.offset; });
  int next = 0;
  for (const fap_part_desc &q : act) {
    if (q.offset != next) return false;
    next = q.offset + q.num;
  }
  return next == n;
}

static int solve_parts_f32(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                           float *d_out, int64_t ld_out, int32_t mode, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && nparts >= 0 && (nparts == 0 || parts) && d_out,
              "fap_hip_solve_parts_f32: null argument");
  const int form = mode & 0xFF;
  FAP_REQUIRE((form == FAP_MODE_PLAIN || form == FAP_MODE_BATCHED || form == FAP_MODE_OVERLAP) &&
                  (mode & ~(0xFF | FAP_MODE_FLAG_INT32_FW | FAP_MODE_FLAG_SSSP |
                            FAP_MODE_FLAG_FW)) == 0,
              "fap_hip_solve_parts_f32: unknown mode %d", mode);
  const int n = g->n;
  FAP_REQUIRE(n > 0 && g->m >= 0 && ld_out >= n && g->d_row_offset && g->d_graph_id &&
                  g->d_st2ed && g->d_ed2st && (g->m == 0 || g->d_col),
              "fap_hip_solve_parts_f32: bad graph view");
  hipStream_t st = resolve(stream);
  const bool batched = form != FAP_MODE_PLAIN;  // OVERLAP runs as BATCHED here
  const bool vec = (ld_out % 4 == 0) && ((uintptr_t)d_out % 16 == 0);

  std::vector<fap_part_desc> act;
  for (int p = 0; p < nparts; ++p) {
    const fap_part_desc &q = parts[p];
    FAP_REQUIRE(q.part_id >= 1 && q.num >= 0 && q.bdy >= 0 && q.bdy <= q.num && q.offset >= 0 &&
                    (int64_t)q.offset + q.num <= n && q.out_row >= 0,
                "fap_hip_solve_parts_f32: bad part %d", p);
    FAP_REQUIRE(q.num <= kMaxPart, "fap_hip_solve_parts_f32: part %d has %d vertices (limit %d)",
                q.part_id, q.num, kMaxPart);
    if (q.num > 0) act.push_back(q);
  }
  const int P = (int)act.size();
  FAP_TRY(timing_start(st));
  if (P == 0) return timing_finish(st);
  // Parts that tile the graph, batched form: the boundary rows come from the boundary
  // graph (the batch stages), not from one SSSP per boundary vertex.
  if (batched && !(mode & FAP_MODE_FLAG_SSSP) && covers_graph(act, n)) {
    FAP_TRY(solve_one_batched(g, act.data(), P, d_out, ld_out, true, 3, st));
    return timing_finish(st);
  }

  const int64_t ldr = align_up(n, 2);  // SSSP row stride (doubles)
  std::vector<int> S(P), nb(P), offsets(P), src_begin(P);
  std::vector<int64_t> fw_off(P), row_off64(P), out_rows(P);
  int64_t fw_total = 0, fw_max = 0;
  int src_total = 0, bdy_max = 0, max_nb = 0, rows_total = 0;
  for (int p = 0; p < P; ++p) {
    S[p] = (int)align_up(act[p].num, kT);
    nb[p] = S[p] / kT;
    max_nb = std::max(max_nb, nb[p]);
    fw_off[p] = batched ? fw_total : 0;
    fw_total += (int64_t)S[p] * S[p];
    fw_max = std::max<int64_t>(fw_max, (int64_t)S[p] * S[p]);
    offsets[p] = act[p].offset;
    src_begin[p] = batched ? src_total : 0;
    row_off64[p] = (int64_t)src_begin[p] * ldr;
    out_rows[p] = src_begin[p];
    src_total += act[p].bdy;
    bdy_max = std::max(bdy_max, act[p].bdy);
  }
  FAP_TRY(g_ws_fw.ensure((batched ? fw_total : fw_max) * 8));
  FAP_TRY(g_ws_rows.ensure(std::max<int64_t>(1, (int64_t)(batched ? src_total : bdy_max) * ldr) * 8));
  double *fw_base = static_cast<double *>(g_ws_fw.ptr);
  double *rows_base = static_cast<double *>(g_ws_rows.ptr);

  std::vector<Build64Desc> bd(P);
  std::vector<Fw64Desc> fd(P);
  std::vector<Merge64Desc> md(P);
  std::vector<Bdy64Desc> yd(P);
  std::vector<int> tiles(P, 0);
  int tile_total = 0;
  int64_t bdy_items = 0;
  for (int p = 0; p < P; ++p) {
    const fap_part_desc &q = act[p];
    double *fw = fw_base + fw_off[p];
    const double *rows = rows_base + row_off64[p];
    bd[p] = Build64Desc{fw, S[p], q.num, q.bdy, q.offset, q.part_id, batched ? rows_total : 0, rows,
                        ldr};
    rows_total += S[p];
    fd[p] = Fw64Desc{fw, nb[p]};
    Merge64Desc m{};
    m.A = fw + (int64_t)q.bdy * S[p];
    m.lda = S[p];
    m.B = rows;
    m.ldb = ldr;
    m.C = d_out + (q.out_row + q.bdy) * ld_out;
    m.ldc = ld_out;
    m.M = q.num - q.bdy;
    m.K = q.bdy;
    m.row_blocks = (m.M + kMRows - 1) / kMRows;
    tiles[p] = m.M > 0 ? m.row_blocks * ((n + kMCols - 1) / kMCols) : 0;
    m.tile_begin = batched ? tile_total : 0;
    tile_total += tiles[p];
    m.part_id = q.part_id;
    m.offset = q.offset;
    m.fw = fw;
    m.ldf = S[p];
    m.fw_row0 = q.bdy;
    md[p] = m;
    yd[p] = Bdy64Desc{rows, ldr, fw, S[p], d_out + q.out_row * ld_out, ld_out, q.bdy, q.offset,
                      q.part_id, batched ? bdy_items : 0};
    bdy_items += (int64_t)q.bdy * n;
  }
  FAP_REQUIRE(tile_total < (1 << 30), "fap_hip_solve_parts_f32: too many tiles");

  // Descriptor tables in one upload.
  int64_t off = 0;
  auto carve = [&](int64_t bytes) {
    const int64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const int64_t o_bd = carve(sizeof(Build64Desc) * P), o_fd = carve(sizeof(Fw64Desc) * P);
  const int64_t o_md = carve(sizeof(Merge64Desc) * P), o_yd = carve(sizeof(Bdy64Desc) * P);
  const int64_t o_off = carve(4LL * P), o_sb = carve(4LL * P), o_or = carve(8LL * P);
  const int64_t copy_end = off;
  const int64_t o_src = carve(4LL * std::max(src_total, 1)), o_rw = carve(8LL * std::max(src_total, 1));
  FAP_TRY(g_ws_meta.ensure(off));
  char *meta = static_cast<char *>(g_ws_meta.ptr);
  std::vector<char> host(copy_end);
  std::memcpy(host.data() + o_bd, bd.data(), sizeof(Build64Desc) * P);
  std::memcpy(host.data() + o_fd, fd.data(), sizeof(Fw64Desc) * P);
  std::memcpy(host.data() + o_md, md.data(), sizeof(Merge64Desc) * P);
  std::memcpy(host.data() + o_yd, yd.data(), sizeof(Bdy64Desc) * P);
  std::memcpy(host.data() + o_off, offsets.data(), 4LL * P);
  std::memcpy(host.data() + o_sb, src_begin.data(), 4LL * P);
  std::memcpy(host.data() + o_or, out_rows.data(), 8LL * P);
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host.data(), copy_end, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  const Build64Desc *d_bd = reinterpret_cast<const Build64Desc *>(meta + o_bd);
  const Fw64Desc *d_fd = reinterpret_cast<const Fw64Desc *>(meta + o_fd);
  const Merge64Desc *d_md = reinterpret_cast<const Merge64Desc *>(meta + o_md);
  const Bdy64Desc *d_yd = reinterpret_cast<const Bdy64Desc *>(meta + o_yd);
  const int *d_offsets = reinterpret_cast<const int *>(meta + o_off);
  const int *d_sb = reinterpret_cast<const int *>(meta + o_sb);
  const int64_t *d_or = reinterpret_cast<const int64_t *>(meta + o_or);
  int *d_src = reinterpret_cast<int *>(meta + o_src);
  int64_t *d_rw = reinterpret_cast<int64_t *>(meta + o_rw);

  auto span = [&](int stage, auto &&fn) -> int {
    hipEvent_t a, b;
    FAP_TRY(timing_record(st, &a));
    FAP_TRY(fn());
    FAP_TRY(timing_record(st, &b));
    timing_add_span(stage, a, b);
    return FAP_OK;
  };
  // Parts [p0, p1) in one launch per stage.
  auto run = [&](int p0, int p1, int nsrc, int nrows, int mnb, int ntiles, int64_t nbdy) -> int {
    FAP_TRY(span(0, [&]() -> int {
      if (nsrc == 0) return FAP_OK;
      FAP_TRY(launch_gather_sources(d_offsets + p0, d_sb + p0, d_or + p0, p1 - p0, nsrc,
                                    g->d_st2ed, 1, d_src, d_rw, st));
      hipLaunchKernelGGL(sssp64_kernel, dim3((nsrc + kSWaves - 1) / kSWaves), dim3(64 * kSWaves), 0,
                         st, g->d_row_offset, g->d_col, g->d_wf, n, d_src, nsrc,
                         rows_base + row_off64[p0], ldr);
      FAP_CHECK_LAUNCH();
      return FAP_OK;
    }));
    FAP_TRY(span(1, [&]() -> int {
      hipLaunchKernelGGL(build64_kernel, dim3(nrows), dim3(256), 0, st, d_bd + p0, p1 - p0, nrows,
                         g->d_row_offset, g->d_col, g->d_wf, g->d_graph_id, g->d_st2ed, g->d_ed2st);
      FAP_CHECK_LAUNCH();
      return FAP_OK;
    }));
    FAP_TRY(span(2, [&]() { return launch_fw64(d_fd + p0, p1 - p0, mnb, st); }));
    FAP_TRY(span(3, [&]() -> int {
      if (ntiles == 0) return FAP_OK;
      if (vec)
        hipLaunchKernelGGL(merge64_kernel<true>, dim3(ntiles), dim3(256), 0, st, d_md + p0, p1 - p0,
                           ntiles, n, g->d_graph_id, g->d_ed2st);
      else
        hipLaunchKernelGGL(merge64_kernel<false>, dim3(ntiles), dim3(256), 0, st, d_md + p0, p1 - p0,
                           ntiles, n, g->d_graph_id, g->d_ed2st);
      FAP_CHECK_LAUNCH();
      return FAP_OK;
    }));
    return span(4, [&]() -> int {
      if (nbdy == 0) return FAP_OK;
      hipLaunchKernelGGL(bdy64_kernel, dim3((unsigned)((nbdy + 255) / 256)), dim3(256), 0, st,
                         d_yd + p0, p1 - p0, nbdy, n, g->d_graph_id, g->d_ed2st);
      FAP_CHECK_LAUNCH();
      return FAP_OK;
    });
  };
  if (batched) {
    FAP_TRY(run(0, P, src_total, rows_total, max_nb, tile_total, bdy_items));
  } else {
    for (int p = 0; p < P; ++p)
      FAP_TRY(run(p, p + 1, act[p].bdy, S[p], nb[p], tiles[p], (int64_t)act[p].bdy * n));
  }
  return timing_finish(st);
}

extern "C" int fap_hip_solve_parts_f32(const fap_graph_dev *g, const fap_part_desc *parts,
                                       int32_t nparts, float *d_out, int64_t ld_out, int32_t mode,
                                       void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  const int s = solve_parts_f32(g, parts, nparts, d_out, ld_out, mode, stream);
  const int f = ws_fence_record(st);  // also after a failed call: its launches used the workspaces
  return s != FAP_OK ? s : f;
}
