// Many independent graphs, solved stage by stage: every stage is one launch over the
// parts (or boundary graphs) of all graphs of a batch, and the whole sequence is
// captured once as a hipGraph and replayed.
//
// Each graph follows fapGraph::run (fap/fap.cpp:195-503) for all its parts.  When an
// item's parts cover its graph (a full APSP) the boundary rows come from the graph's
// boundary graph instead of one SSSP per boundary vertex (SSSP is latency-bound: tens
// of thousands of label-correcting waves took 60 % of the batch time):
//   1 build           every part's s x s matrix from its in-part arcs (CSR order, a
//                     later arc wins, util_Myalgorithm.h:44-84), diagonal 0;
//   2 Floyd-Warshall  all matrices, batched blocked FW (floyd.hip / fp64.hip): the
//                     in-part distances d_q (floyd.h:38-49);
//   3 boundary graph  per graph, its boundary vertices (the first b_q sorted positions
//                     of every part, util_preprocess.h:66-71) with the in-part blocks
//                     d_q[B_q][B_q] and the cut arcs (both ends are boundary vertices,
//                     util_preprocess.h:32-58) as arcs;
//   4 Floyd-Warshall  all boundary-graph matrices: D_B = the graph's distances between
//                     boundary vertices (every path between them is a path of the
//                     boundary graph);
//   5 boundary rows   d(x, v) = min_{c in B_q} D_B(x, c) + d_q(c, v) for v in part q
//                     (the last entry into q): one min-plus product per (graph, part);
//   6 merge           inner rows: d(i, v) = min(d_p(i, v) [v in p],
//                     min_{k in B_p} d_p(i, k) + d(k, v)) (minplus.h:51-64 with the
//                     decode of util_Myalgorithm.h:236-247 folded in), a persistent
//                     grid pulling (part, row block, column block) units from a
//                     device counter, largest parts first.
// Items with a subset of their graph's parts (sub-graph requests) take the reference's
// form instead: 0 boundary SSSP (one wave per boundary source, label-correcting,
// batch_sssp.h:232-252), 1 build with those rows as the boundary rows, 2, 6.
// Every quantity is an exact min/plus expression of integer weights (int32) or of
// float32 weights summed in FP64 (float32 output with MAXVALUE), so the two forms agree
// bit for bit with each other and with the per-graph pipeline.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"
#include "mp_tile.h"
#include "pk16.h"

namespace fap_hip {

int launch_fw(const FwDesc *d_descs, int ndesc, int max_nb, hipStream_t stream);
int launch_fw64(const Fw64Desc *d_descs, int nd, int max_nb, hipStream_t st);
int launch_fw64_t32(const Fw64Desc *d_descs, int nd, int max_nb, hipStream_t st);
int launch_fw16(const Fw16Desc *d_descs, int ndesc, int max_nb, hipStream_t stream);
int launch_fw16_unpack(const UnpackDesc *d_descs, int nd, int64_t total_items,
                       const int *d_comp_sorted, int *d_flags, hipStream_t stream);
int launch_inpart_bfs(const InpartDesc *d_descs, int nd, int total_groups, int s_max,
                      hipStream_t stream);

namespace {

constexpr int kBT = 64;           // tile edge of the Floyd-Warshall matrices
constexpr int kBatchMaxS = kMaxPart;  // largest part
constexpr int kBQ = 1024;         // SSSP frontier queue entries per wave
constexpr int kSWaves = 4;        // waves (sources) per workgroup of the SSSP launch
constexpr int kMCols = 256;       // merge: columns per unit (four per lane)
constexpr int kWideN = 16384;     // FP64 merge: 8 x 8 register tiles from this graph size
constexpr int kKc = 64;           // k per LDS chunk of the products

// One graph of the batch.
struct BGraph {
  const int *row_off, *col, *w;  // CSR (w: int32 weights or NULL)
  const float *wf;               // float32 weights (FP64 path) or NULL
  const int *graph_id, *st2ed, *ed2st;
  void *out;                     // n x ld_out result (int32 or float32)
  int64_t ld_out;
  void *R;                       // boundary rows of a full item (nb x ldr, +inf padded)
  void *Rs;                      // the same rows with columns in sorted-position order
  int perm;                      // 1: boundary rows via Rs and bperm_kernel (scattered parts)
  int rlo, rcnt;                 // boundary-graph rows produced (all, or one part's)
  int tile;                      // Floyd-Warshall tile edge of its parts and boundary graph
  int pbase;                     // first of them in the row numbering of bperm_kernel
  int64_t ldr;
  void *MB;                      // boundary-graph matrix (SB x SB), full items only
  void *MB16;                    // int32 packed graphs: the same matrix capped at kCap16
                                 //   as uint16, closed by the 16-bit Floyd-Warshall
  int SB;
  int n, nb;                     // vertices; boundary vertices (full items)
  int bbase;                     // first entry of the graph in bpart / brow
  int pid_base, max_pid;         // pid2idx[pid_base + part id] = part index, or -1
  int gb_row_begin;              // first boundary-graph build row
  int full;                      // 1: parts cover the graph (boundary-graph form)
  int packed;                    // 1: merge on packed 16-bit distances (unit weights,
                                 //    n <= kCap16: every finite distance below the cap)
};

// One part of one graph.
struct BPart {
  void *D;                       // S x S matrix (ld = S)
  const void *Brows;             // the part's boundary rows (int32: rows of out)
  int64_t ldb;
  int64_t out_row;               // first result row of the part
  int S, g, part_id, off, s, b;
  int bs;                        // boundary-graph index of its first boundary vertex
  int row_begin;                 // first build row (prefix of S)
  int bfs;                       // in-part distances by bit-parallel BFS (no build, no FW)
  int nested;                    // in-part distances by the nested closure (no build, no FW)
};

struct BRowUnit {                // stage 5: boundary-graph rows r0 .. of graph parts[q].g,
  int q, r0, cb, pad;            // columns cb * 256 .. of part q
};

struct BMerge {                  // stage 6: one unit (part, row block, column block)
  int p, rb, cb, pad;
};

struct BSrc {                    // stage 0: boundary source q of part p
  int p, q;
};

template <typename T>
struct Dist;

template <>
struct Dist<int> {
  typedef int Out;
  static __device__ __forceinline__ int inf() { return kInf; }
  static __device__ __forceinline__ int mn(int a, int b) { return min(a, b); }
  static __device__ __forceinline__ int mn3(int a, int x, int y) { return min3_i32(a, x, y); }
  static __device__ __forceinline__ int weight(const BGraph &g, int e) {
    return g.w ? g.w[e] : 1;
  }
  static __device__ __forceinline__ int out(int v) { return v; }
  static __device__ __forceinline__ int load_agent(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ int atomic_min(int *p, int v) { return atomicMin(p, v); }
};

template <>
struct Dist<double> {
  typedef float Out;
  static __device__ __forceinline__ double inf() { return __builtin_inf(); }
  static __device__ __forceinline__ double mn(double a, double b) { return fmin(a, b); }
  static __device__ __forceinline__ double mn3(double a, double x, double y) {
    return fmin(a, fmin(x, y));
  }
  static __device__ __forceinline__ double weight(const BGraph &g, int e) {
    return g.wf ? (double)g.wf[e] : (g.w ? (double)g.w[e] : 1.0);
  }
  static __device__ __forceinline__ float out(double v) { return v < 1e8 ? (float)v : 1e8f; }
  static __device__ __forceinline__ double load_agent(const double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT));
  }
  static __device__ __forceinline__ double atomic_min(double *p, double v) {
    // Non-negative doubles and +inf order as their unsigned bit patterns.
    return __longlong_as_double((long long)atomicMin(
        reinterpret_cast<unsigned long long *>(p),
        (unsigned long long)__double_as_longlong(v)));
  }
};

// Value of lane j of the wave (j wave-uniform).
__device__ __forceinline__ int bcast_lane(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ __forceinline__ double bcast_lane(double v, int j) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, j);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Largest i in [0, n) with key(i) <= x (key ascending, key(0) <= x).
template <typename F>
__device__ __forceinline__ int upper_index(int n, int x, F key) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------- 0: boundary SSSP
template <typename T>
__device__ __forceinline__ void relax(const BGraph &g, T *row, int u, int *q_next, int *q_count) {
  const T du = Dist<T>::load_agent(row + u);
  if (!(du < Dist<T>::inf())) return;
  const int e1 = g.row_off[u + 1];
  for (int e = g.row_off[u]; e < e1; ++e) {
    const int v = g.col[e];
    const T nd = du + Dist<T>::weight(g, e);
    if (nd < row[v] && nd < Dist<T>::atomic_min(row + v, nd)) {
      const int pos = atomicAdd(q_count, 1);
      if (pos < kBQ) q_next[pos] = v;
    }
  }
}

// One wave computes the full row of source s (labels only decrease, so a stale plain
// read of row[v] can only let an unnecessary atomic through).  A frontier larger than
// the LDS queue is replaced by a scan of the whole row.
template <typename T>
__device__ void wave_sssp(const BGraph &g, int s, T *row, int *queue, int *qcount, int lane) {
  for (int v = lane; v < g.n; v += 64) row[v] = (v == s) ? T(0) : Dist<T>::inf();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) {
    queue[0] = s;
    qcount[0] = 1;
    qcount[1] = 0;
  }
  wave_sync();
  int cur = 0, qsize = 1;
  bool scan = false;
  while (true) {
    const int nxt = cur ^ 1;
    int *q_next = queue + nxt * kBQ;
    int *q_cnt = qcount + nxt;
    if (!scan) {
      for (int i = lane; i < qsize; i += 64) relax(g, row, queue[cur * kBQ + i], q_next, q_cnt);
    } else {
      for (int u = lane; u < g.n; u += 64) relax(g, row, u, q_next, q_cnt);
    }
    wave_sync();
    const int cnt = __hip_atomic_load(q_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync();
    if (cnt == 0) break;
    scan = cnt > kBQ;
    qsize = scan ? 0 : cnt;
    if (lane == 0) qcount[cur] = 0;
    wave_sync();
    cur = nxt;
  }
}

template <typename T>
__global__ __launch_bounds__(64 * kSWaves) void bsssp_kernel(const BGraph *__restrict__ graphs,
                                                             const BPart *__restrict__ parts,
                                                             const BSrc *__restrict__ srcs,
                                                             int nsrc) {
  __shared__ int queues[kSWaves][2 * kBQ + 2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kSWaves + wave;
  if (i >= nsrc) return;  // whole wave leaves; no workgroup barrier is used
  const BSrc sr = srcs[i];
  const BPart p = parts[sr.p];
  const BGraph g = graphs[p.g];
  T *row = static_cast<T *>(const_cast<void *>(p.Brows)) + (int64_t)sr.q * p.ldb;
  const int s = g.st2ed[p.off + sr.q];
  if (s < 0 || s >= g.n) return;  // a corrupt st2ed entry: no row
  wave_sssp<T>(g, s, row, queues[wave], queues[wave] + 2 * kBQ, lane);
}

// ---------------------------------------------------------------- 1: build
// One wave per padded row of every part.  Every entry is written once by the lane
// owning its column: INF, the weight of the last in-part arc to it (CSR order), 0 on
// the diagonal; for parts of partial items the boundary rows are the SSSP rows.
template <typename T>
__global__ __launch_bounds__(256) void bbuild_kernel(const BGraph *__restrict__ graphs,
                                                     const BPart *__restrict__ parts, int np,
                                                     int total_rows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= total_rows) return;
  const int pi = upper_index(np, row, [&](int k) { return parts[k].row_begin; });
  const BPart p = parts[pi];
  if (p.bfs || p.nested) return;  // the BFS or the nested closure writes the matrix
  const BGraph g = graphs[p.g];
  const int i = row - p.row_begin;
  FAP_GLOBAL T *D = gptr(static_cast<T *>(p.D)) + (int64_t)i * p.S;
  if (i >= p.s) {
    for (int t = lane; t < p.S; t += 64) D[t] = Dist<T>::inf();
    return;
  }
  if (!g.full && i < p.b) {  // boundary row of a partial item: its SSSP row
    const FAP_GLOBAL T *src = gptr(static_cast<const T *>(p.Brows)) + (int64_t)i * p.ldb;
    for (int t = lane; t < p.S; t += 64)
      D[t] = t < p.s ? src[g.st2ed[p.off + t]] : Dist<T>::inf();
    return;
  }
  const int v = g.st2ed[p.off + i];
  const int e0 = g.row_off[v], e1 = g.row_off[v + 1];
  // The row's arcs, 64 at a time: lane l holds arc e0 + 64c + l (its in-part column or
  // -1, and its weight); each lane then scans them in CSR order for its own columns
  // (readlane: no memory access per column; a later arc overwrites an earlier one).
  // Reading graph_id / ed2st of every arc for every column took 1.3 ms of the 512-graph
  // batch.
  constexpr int kCols = 16;  // columns per lane held in registers (S <= 1024 in one pass)
  for (int t0 = 0; t0 < p.S; t0 += 64 * kCols) {
    T val[kCols];
#pragma unroll
    for (int q = 0; q < kCols; ++q) val[q] = Dist<T>::inf();
    for (int c0 = e0; c0 < e1; c0 += 64) {
      const int e = c0 + lane;
      int tcol = -1;
      T w = Dist<T>::inf();
      if (e < e1) {
        const int u = g.col[e];
        if (g.graph_id[u] == p.part_id) {
          tcol = g.ed2st[u] - p.off;
          w = Dist<T>::weight(g, e);
        }
      }
      const int cnt = min(64, e1 - c0);
      for (int j = 0; j < cnt; ++j) {  // uniform
        const int tj = __builtin_amdgcn_readlane(tcol, j);
        if (tj < t0 || tj >= t0 + 64 * kCols) continue;  // uniform
        const T wj = bcast_lane(w, j);
#pragma unroll
        for (int q = 0; q < kCols; ++q)
          if (t0 + q * 64 + lane == tj) val[q] = wj;
      }
    }
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
      const int t = t0 + q * 64 + lane;
      if (t < p.S) D[t] = t == i ? T(0) : val[q];
    }
  }
}

// ---------------------------------------------------------------- 3: boundary graph
// One wave per padded row of every boundary-graph matrix.  Row r = boundary vertex x
// of part q (local c): the in-part block d_q[c][B_q], the cut arcs of x (minimum over
// parallel arcs), 0 on the diagonal.
template <typename T>
__global__ __launch_bounds__(256) void bgb_kernel(const BGraph *__restrict__ graphs,
                                                  const BPart *__restrict__ parts,
                                                  const int *__restrict__ gbg, int ngb,
                                                  const int *__restrict__ bpart,
                                                  const int *__restrict__ pid2idx,
                                                  int total_rows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= total_rows) return;
  const int gi = gbg[upper_index(ngb, row, [&](int k) { return graphs[gbg[k]].gb_row_begin; })];
  const BGraph g = graphs[gi];
  const int r = row - g.gb_row_begin;
  FAP_GLOBAL T *M = gptr(static_cast<T *>(g.MB)) + (int64_t)r * g.SB;
  // Packed graphs (every finite distance below kCap16): the row goes to the 16-bit
  // matrix, capped (the cap reads as unreachable).
  FAP_GLOBAL uint16_t *M16 =
      g.MB16 ? gptr(static_cast<uint16_t *>(g.MB16)) + (int64_t)r * g.SB : nullptr;
  auto put = [&](int t, T v) {
    if constexpr (sizeof(T) == 4) {
      if (M16) {
        M16[t] = (uint16_t)min((unsigned)v, (unsigned)kCap16);
        return;
      }
    }
    M[t] = v;
  };
  if (r >= g.nb) {
    for (int t = lane; t < g.SB; t += 64) put(t, Dist<T>::inf());
    return;
  }
  const BPart q = parts[bpart[g.bbase + r]];
  const int c = r - q.bs;
  const FAP_GLOBAL T *Dq = gptr(static_cast<const T *>(q.D)) + (int64_t)c * q.S;
  const int x = g.st2ed[q.off + c];
  const int e0 = g.row_off[x], e1 = g.row_off[x + 1];
  // Cross-part arcs resolved once per arc (lane l: arc e0 + 64c + l), then applied to
  // the lanes' columns by readlane, as in bbuild_kernel.
  constexpr int kCols = 16;
  for (int t0 = 0; t0 < g.SB; t0 += 64 * kCols) {
    T val[kCols];
#pragma unroll
    for (int qq = 0; qq < kCols; ++qq) {
      const int t = t0 + qq * 64 + lane;
      val[qq] = (t >= q.bs && t < q.bs + q.b) ? Dq[t - q.bs] : Dist<T>::inf();
    }
    for (int c0 = e0; c0 < e1; c0 += 64) {
      const int e = c0 + lane;
      int tcol = -1;
      T w = Dist<T>::inf();
      if (e < e1) {
        const int y = g.col[e];
        const int py = g.graph_id[y];
        if (py != q.part_id && py >= 1 && py <= g.max_pid) {
          const int qi = pid2idx[g.pid_base + py];
          if (qi >= 0) {
            const BPart qy = parts[qi];
            const int ly = g.ed2st[y] - qy.off;
            if (ly >= 0 && ly < qy.b) {
              tcol = qy.bs + ly;
              w = Dist<T>::weight(g, e);
            }
          }
        }
      }
      const int cnt = min(64, e1 - c0);
      for (int j = 0; j < cnt; ++j) {  // uniform
        const int tj = __builtin_amdgcn_readlane(tcol, j);
        if (tj < t0 || tj >= t0 + 64 * kCols) continue;  // uniform
        const T wj = bcast_lane(w, j);
#pragma unroll
        for (int qq = 0; qq < kCols; ++qq)
          if (t0 + qq * 64 + lane == tj) val[qq] = Dist<T>::mn(val[qq], wj);
      }
    }
#pragma unroll
    for (int qq = 0; qq < kCols; ++qq) {
      const int t = t0 + qq * 64 + lane;
      if (t < g.SB) put(t, t == r ? T(0) : val[qq]);
    }
  }
}

// Partial items: the SSSP rows (scratch) to the result rows (int32, or float32 from
// FP64).
template <typename T>
__global__ __launch_bounds__(256) void bconv_kernel(const BGraph *__restrict__ graphs,
                                                    const BPart *__restrict__ parts,
                                                    const BSrc *__restrict__ srcs) {
  typedef typename Dist<T>::Out Out;
  const BSrc sr = srcs[blockIdx.x];
  const BPart p = parts[sr.p];
  const BGraph g = graphs[p.g];
  const FAP_GLOBAL T *src = gptr(static_cast<const T *>(p.Brows)) + (int64_t)sr.q * p.ldb;
  FAP_GLOBAL Out *dst = gptr(static_cast<Out *>(g.out)) + (p.out_row + sr.q) * g.ld_out;
  for (int v = threadIdx.x; v < g.n; v += 256) dst[v] = Dist<T>::out(src[v]);
}

// ---------------------------------------------------------------- 6: merge
// Persistent grid; a unit is (part, row block of 4 * RB inner rows, column block of 256
// columns), the row blocks of one column block consecutive so its boundary-row panel
// is reused from L2.  Wave w owns rows w*RB .. w*RB + RB - 1 of the block and lane l
// the columns l + 64e (e < 4), so every load and store of a wave covers 256
// contiguous bytes.  The 4*RB x 64 A block is staged k-major in LDS (broadcast reads);
// the four waves read the same B rows (L1).  Own columns take min(d_p(i, v), product).
constexpr int kMWaves = 4;
// Row padding of the staged A block: 16 bytes, so the staging writes (64 lanes on 64
// consecutive k of one row) spread over 16 banks instead of one, and the broadcast
// reads of a row's entries stay 16-byte aligned.
template <typename T>
constexpr int kAPad = 16 / (int)sizeof(T);

// Operands of one min-plus unit: rows 0 .. 4*RB - 1 of A (the first `rows` valid, the
// rest +inf) against columns col0 + l + 64e (e < 4) of B, over k < kdim.  B columns
// above colmax are read at colmax (their results are never stored), so every B load
// is in bounds without a mask.
template <typename T>
struct MpArgs {
  const FAP_GLOBAL T *A;
  int64_t lda;
  int rows, kdim;
  const FAP_GLOBAL T *B;
  int64_t ldb;
  int col0, colmax;
};

// Stages A[r][k0 + k] (r < 4*RB, k < kKc; +inf outside rows < m.rows, k < kc) k-major
// in LDS; packed: capped at kCap16 and broadcast to both halves.
template <typename T, int RB, bool kPacked>
__device__ __forceinline__ void mp_stage(const MpArgs<T> &m, T (*As)[kMWaves * RB + kAPad<T>],
                                         int k0, int kc, int tid) {
  constexpr int kRows = kMWaves * RB;
  for (int x = tid; x < kKc * kRows; x += 256) {
    const int r = x / kKc, k = x - r * kKc;
    T a = (k < kc && r < m.rows) ? m.A[(int64_t)r * m.lda + k0 + k] : Dist<T>::inf();
    if constexpr (kPacked) {
      const unsigned u = (unsigned)a < (unsigned)kCap16 ? (unsigned)a : (unsigned)kCap16;
      a = (T)(u * 0x10001u);
    }
    As[k][r] = a;
  }
}

template <typename T, int RB, bool kPacked, int kE = 4>
__device__ __forceinline__ void mp_init(T (&acc)[RB][kPacked ? 2 : kE]) {
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int e = 0; e < (kPacked ? 2 : kE); ++e) acc[r][e] = kPacked ? (T)kPk16Top : Dist<T>::inf();
}

// acc[r][.] = min(acc, min_{k < kc} As[k][wave*RB + r] + B[k0 + k][col]) for the wave's
// RB rows and the lane's kE (four unless noted) columns col0 + l + 64e.  kPacked (int32, unit weights,
// n <= kCap16): two columns per register, a step on two entries one v_pk_add_u16 and
// half a v_pk_minimum3_f16 (pk16.h); every finite distance is below the cap, so a
// result at the cap is unreachable.  The four waves read the same B rows (L1).
template <typename T, int RB, bool kPacked, int kE = 4>
__device__ __forceinline__ void mp_chunk(const MpArgs<T> &m, const T (*As)[kMWaves * RB + kAPad<T>],
                                         int k0, int kc, int col0, T (&acc)[RB][kPacked ? 2 : kE],
                                         int tid) {
  static_assert(!kPacked || kE == 4, "packed units hold four columns per lane");
  constexpr int kLdA = kMWaves * RB + kAPad<T>;
  const int wave = tid >> 6, lane = tid & 63;
  const FAP_GLOBAL T *B = m.B;
  int col[kE];
#pragma unroll
  for (int e = 0; e < kE; ++e) col[e] = min(col0 + e * 64 + lane, m.colmax);
  const T *Aw = &As[0][wave * RB];
  if constexpr (sizeof(T) == 8) {
    // FP64 has no three-operand minimum: one k per step (fewer live registers), row
    // k + 1 in flight while row k is used (two register buffers).
    auto load_b = [&](int k, T (&x)[kE]) {
      const FAP_GLOBAL T *rp = B + (int64_t)(k0 + k) * m.ldb;
#pragma unroll
      for (int e = 0; e < kE; ++e) x[e] = rp[col[e]];
    };
    auto step = [&](int k, const T (&bv)[kE]) {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const T a = Aw[k * kLdA + r];
#pragma unroll
        for (int e = 0; e < kE; ++e) acc[r][e] = Dist<T>::mn(acc[r][e], a + bv[e]);
      }
    };
    T b0[kE], b1[kE];
    load_b(0, b0);
    int k = 0;
    for (; k + 1 < kc; k += 2) {
      load_b(k + 1, b1);
      step(k, b0);
      // Unconditional (the last row again at the end): a load behind a branch makes
      // the wait before step k + 1 cover it too.
      load_b(min(k + 2, kc - 1), b0);
      step(k + 1, b1);
    }
    if (k < kc) step(k, b0);
  } else {
    // Rows k, k + 1 of the chunk as one pair (the three-operand minimum), the next
    // pair in flight while this one is used; rows past kc read as +inf (row kc of As
    // is +inf when kc is odd, and a pair past the end loads row kc - 1 again).
    auto load_b = [&](int k, T (&x)[kE]) {
      const FAP_GLOBAL T *rp = B + (int64_t)(k0 + min(k, kc - 1)) * m.ldb;
#pragma unroll
      for (int e = 0; e < kE; ++e) x[e] = rp[col[e]];
    };
    T n0[kE], n1[kE];
    load_b(0, n0);
    load_b(1, n1);
    for (int k = 0; k < kc; k += 2) {
      T b0[kE], b1[kE];
#pragma unroll
      for (int e = 0; e < kE; ++e) {
        b0[e] = n0[e];
        b1[e] = n1[e];
      }
      load_b(k + 2, n0);
      load_b(k + 3, n1);
      T a0[RB], a1[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        a0[r] = Aw[k * kLdA + r];
        a1[r] = Aw[(k + 1) * kLdA + r];  // row kc of As is INF when kc is odd
      }
      if constexpr (kPacked) {
        auto cap = [](T v) -> unsigned {
          const unsigned u = (unsigned)v;
          return u < (unsigned)kCap16 ? u : (unsigned)kCap16;
        };
        const unsigned p0[2] = {cap(b0[0]) | (cap(b0[1]) << 16u),
                                cap(b0[2]) | (cap(b0[3]) << 16u)};
        const unsigned p1[2] = {cap(b1[0]) | (cap(b1[1]) << 16u),
                                cap(b1[2]) | (cap(b1[3]) << 16u)};
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const unsigned s0 = __builtin_bit_cast(
                unsigned,
                (us2)(__builtin_bit_cast(us2, (unsigned)a0[r]) + __builtin_bit_cast(us2, p0[h])));
            const unsigned s1 = __builtin_bit_cast(
                unsigned,
                (us2)(__builtin_bit_cast(us2, (unsigned)a1[r]) + __builtin_bit_cast(us2, p1[h])));
            acc[r][h] = (T)pk_min3((unsigned)acc[r][h], s0, s1);
          }
      } else {
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
          for (int e = 0; e < kE; ++e)
            acc[r][e] = Dist<T>::mn3(acc[r][e], a0[r] + b0[e], a1[r] + b1[e]);
      }
    }
  }
}

// One unit: acc over every k chunk, then epi(acc) (after the workgroup's last LDS read).
template <typename T, int RB, bool kPacked, int kE = 4, typename Epi>
__device__ __forceinline__ void mp_unit(const MpArgs<T> &m, T (*As)[kMWaves * RB + kAPad<T>],
                                        int tid, Epi &&epi) {
  T acc[RB][kPacked ? 2 : kE];
  mp_init<T, RB, kPacked, kE>(acc);
  for (int k0 = 0; k0 < m.kdim; k0 += kKc) {
    const int kc = min(kKc, m.kdim - k0);
    mp_stage<T, RB, kPacked>(m, As, k0, kc, tid);
    __syncthreads();
    mp_chunk<T, RB, kPacked, kE>(m, As, k0, kc, m.col0, acc, tid);
    __syncthreads();
  }
  epi(acc);
}

// Entry e (column col0 + l + 64e) of accumulator row r, unpacked.
template <typename T, int RB, bool kPacked>
__device__ __forceinline__ T mp_value(const T (&acc)[RB][kPacked ? 2 : 4], int r, int e) {
  if constexpr (kPacked) {
    const unsigned h = ((unsigned)acc[r][e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
    return h >= (unsigned)kCap16 ? Dist<T>::inf() : (T)h;
  } else {
    return acc[r][e];
  }
}

// One merge unit: inner rows r0 .. r0 + 4*RB - 1 of part p against the 256 columns of
// column block cb.  Own columns take min(d_p(i, v), product) (the decode).
template <typename T, int RB, bool kPacked>
__device__ __forceinline__ void merge_unit(const BGraph &g, const BPart &p, int rb, int cb,
                                           T (*As)[kMWaves * RB + kAPad<T>], int tid) {
  typedef typename Dist<T>::Out Out;
  constexpr int kRows = kMWaves * RB;
  const int wave = tid >> 6, lane = tid & 63;
  const int inner = p.s - p.b, r0 = rb * kRows;
  const FAP_GLOBAL T *D = gptr(static_cast<const T *>(p.D));
  MpArgs<T> m;
  m.A = D + (int64_t)(p.b + r0) * p.S;
  m.lda = p.S;
  m.rows = inner - r0;
  m.kdim = p.b;
  m.B = gptr(static_cast<const T *>(p.Brows));
  m.ldb = p.ldb;
  m.col0 = cb * kMCols;
  m.colmax = INT32_MAX;  // boundary rows are padded with +inf to whole column blocks
  FAP_GLOBAL Out *out = gptr(static_cast<Out *>(g.out));
  mp_unit<T, RB, kPacked>(m, As, tid, [&](const T (&acc)[RB][kPacked ? 2 : 4]) {
    int col[4], own[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      col[e] = cb * kMCols + e * 64 + lane;
      own[e] = col[e] < g.n && g.graph_id[col[e]] == p.part_id ? g.ed2st[col[e]] - p.off : -1;
      if (own[e] >= p.s) own[e] = -1;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int i = r0 + wave * RB + r;
      if (i >= inner) continue;
      FAP_GLOBAL Out *orow = out + (p.out_row + p.b + i) * g.ld_out;
      const FAP_GLOBAL T *drow = D + (int64_t)(p.b + i) * p.S;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (col[e] >= g.n) continue;
        T val = mp_value<T, RB, kPacked>(acc, r, e);
        if (own[e] >= 0) val = Dist<T>::mn(val, drow[own[e]]);
        orow[col[e]] = Dist<T>::out(val);
      }
    }
  });
}

// ---------------------------------------------------------------- 5: boundary rows
// Unit (q, r0, cb): boundary-graph rows x = r0 .. r0 + 4*RB - 1 of graph g against
// columns j = cb*256 .. of part q: C[x][j] = min_{c < b_q} D_B[x][bs_q + c] + d_q[c][j]
// (the same min-plus unit as the merge: D_B rows staged k-major in LDS, d_q rows read
// four columns per lane, three-operand or packed 16-bit minima), written to the result
// at row brow[x], column st2ed[off_q + j], and to the padded boundary rows R.  The
// first form (one column per thread, two VALU instructions per step, 32 broadcast LDS
// reads per k) spent as many LDS cycles as VALU cycles.
template <typename T, int RB, bool kPacked>
__device__ __forceinline__ void brows_unit(const BGraph &g, const BPart &q, int r0, int cb,
                                           const int64_t *__restrict__ brow,
                                           T (*As)[kMWaves * RB + kAPad<T>], int tid) {
  typedef typename Dist<T>::Out Out;
  const int wave = tid >> 6, lane = tid & 63;
  MpArgs<T> m;
  m.A = gptr(static_cast<const T *>(g.MB)) + (int64_t)r0 * g.SB + q.bs;
  m.lda = g.SB;
  const int rend = g.rlo + g.rcnt;  // rows r0 .. rend - 1 are produced
  m.rows = rend - r0;
  m.kdim = q.b;
  m.B = gptr(static_cast<const T *>(q.D));
  m.ldb = q.S;
  m.col0 = cb * kMCols;
  m.colmax = q.S - 1;  // padded columns read +inf; clamped ones are never stored
  FAP_GLOBAL T *Rs = gptr(static_cast<T *>(g.Rs));
  FAP_GLOBAL Out *out = gptr(static_cast<Out *>(g.out));
  FAP_GLOBAL T *R = gptr(static_cast<T *>(g.R));
  mp_unit<T, RB, kPacked>(m, As, tid, [&](const T (&acc)[RB][kPacked ? 2 : 4]) {
    if (!g.perm) {  // straight to the vertex columns
      int v[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = cb * kMCols + e * 64 + lane;
        v[e] = j < q.s ? g.st2ed[q.off + j] : -1;
        if (v[e] >= g.n) v[e] = -1;  // (a corrupt st2ed entry writes nothing)
      }
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int x = r0 + wave * RB + r;
        if (x >= rend) continue;
        FAP_GLOBAL Out *orow = out + brow[g.bbase + x] * g.ld_out;
        FAP_GLOBAL T *rrow = R + (int64_t)(x - g.rlo) * g.ldr;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (v[e] < 0) continue;
          const T val = mp_value<T, RB, kPacked>(acc, r, e);
          orow[v[e]] = Dist<T>::out(val);
          rrow[v[e]] = val;
        }
      }
      return;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int x = r0 + wave * RB + r;
      if (x >= rend) continue;
      FAP_GLOBAL T *rrow = Rs + (int64_t)(x - g.rlo) * g.ldr + q.off;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = cb * kMCols + e * 64 + lane;
        if (j < q.s) rrow[j] = mp_value<T, RB, kPacked>(acc, r, e);
      }
    }
  });
}

// Boundary row x of graph g from sorted-position order to vertex order: the result row
// brow[x] and the padded row R[x] take Rs[x][ed2st[v]] at column v (gathered reads of
// one row, coalesced writes).  For graphs whose parts hold scattered vertex ids (kNN
// point sets): writing the boundary-row product straight to vertex columns scattered
// every store (6 of 33 ms of a 60k-vertex kNN graph's solve); on grid-numbered road
// graphs the direct writes are cheaper (the extra pass cost 1.2 ms of the 512-graph
// batch).
template <typename T>
__global__ __launch_bounds__(256) void bperm_kernel(const BGraph *__restrict__ graphs,
                                                    const int *__restrict__ gbg, int ngb,
                                                    const int64_t *__restrict__ brow) {
  typedef typename Dist<T>::Out Out;
  const int row = blockIdx.x;
  const int gi = gbg[upper_index(ngb, row, [&](int k) { return graphs[gbg[k]].pbase; })];
  const BGraph g = graphs[gi];
  const int xr = row - g.pbase;  // row of R / Rs
  if (!g.perm || xr < 0 || xr >= g.rcnt) return;
  const int x = g.rlo + xr;
  const FAP_GLOBAL T *src = gptr(static_cast<const T *>(g.Rs)) + (int64_t)xr * g.ldr;
  FAP_GLOBAL T *rrow = gptr(static_cast<T *>(g.R)) + (int64_t)xr * g.ldr;
  FAP_GLOBAL Out *orow = gptr(static_cast<Out *>(g.out)) + brow[g.bbase + x] * g.ld_out;
  for (int v = threadIdx.x; v < g.n; v += 256) {
    const int pos = g.ed2st[v];
    const T val = (pos >= 0 && pos < g.n) ? src[pos] : Dist<T>::inf();
    rrow[v] = val;
    orow[v] = Dist<T>::out(val);
  }
}

template <typename T, int RB>
__global__ __launch_bounds__(256) void brows_kernel(const BGraph *__restrict__ graphs,
                                                    const BPart *__restrict__ parts,
                                                    const BRowUnit *__restrict__ units,
                                                    const int64_t *__restrict__ brow) {
  __shared__ __attribute__((aligned(16))) T As[kKc][kMWaves * RB + kAPad<T>];
  const BRowUnit un = units[blockIdx.x];
  if (un.q < 0) return;  // past the end of its XCD's list
  const BPart q = parts[un.q];
  const BGraph g = graphs[q.g];
  if constexpr (sizeof(T) == 4) {
    if (g.packed) {
      brows_unit<T, RB, true>(g, q, un.r0, un.cb, brow, As, threadIdx.x);
      return;
    }
  }
  brows_unit<T, RB, false>(g, q, un.r0, un.cb, brow, As, threadIdx.x);
}

// ---------------------------------------------------------------- 2': nested closure
// In-part distances of a weighted part of a full item (distances from its own arcs
// alone) with at least kNestMinS vertices, without Floyd-Warshall on its s x s matrix
// (s^3 steps).  The host cuts the part into sub-parts of at most kNestT vertices
// (recursive BFS bisection) and orders each sub-part's vertices crossing-first, where a
// crossing vertex has an in-part arc to or from another sub-part.  Then:
//   a  every sub-part's kNestT x kNestT matrix from its arcs (diagonal 0), closed by one
//      Floyd-Warshall tile: w_q, distances inside sub-part q;
//   b  the crossing graph X of the part: its crossing vertices, the closed blocks
//      w_q[crossing][crossing] and the arcs between sub-parts; closed by the batched
//      Floyd-Warshall: X* = in-part distances between crossing vertices;
//   c  crossing rows: d(x, v) = min_{y crossing in q(v)} X*(x, y) + w_q(v)(y, v)
//      (the last entry into v's sub-part; y = x covers paths that never leave it);
//   d  every row: d(i, v) = min(w_q(i)(i, v) [v in q(i)],
//                               min_{x crossing in q(i)} w_q(i)(i, x) + d(x, v))
//      (the first exit from i's sub-part), stored at (sigma(i), sigma(v)) of the part's
//      matrix, sigma = sub-part position -> the part's own vertex order.
// Each is an exact min over the same in-part paths as Floyd-Warshall's closure of the
// same arcs (parallel arcs: the later one, as in the build), so the matrices are equal
// entry for entry.  Work for a 530-vertex road part: about 3e7 steps against 1.9e8.
constexpr int kNestT = 64;      // largest sub-part: one Floyd-Warshall tile
template <typename T>
constexpr int kNestRB = sizeof(T) == 8 ? 8 : 16;  // rows per wave of the part-row units
constexpr int kNestMinS = 160;  // smallest part taking the nested closure

struct NPart {
  void *D;         // the part's matrix (S x S), as for Floyd-Warshall
  void *X;         // crossing graph (SX x SX)
  void *R;         // crossing rows (nx x ldr, columns in sub-part positions)
  int64_t ldr;
  int S, s, nx, SX;
  int pos0, pad;   // first entry of the part in sigma / aoff
};

struct NSub {
  void *W;         // closed sub-part matrix (kNestT x kNestT)
  int part;        // index into NPart
  int st, n, b;    // first position, vertices, crossing vertices (positions st .. st+b-1)
  int x0;          // crossing-graph index of its first crossing vertex
  int row0, xrow0; // first row in the build / crossing-graph row numbering
};

struct NUnit {
  int sub, rb, cb, pad;
};

// a: row `a` of sub-part k's matrix.  acode >= 0: arc to sub-part column acode; < 0: to
// crossing index -1 - acode (another sub-part).  Arcs are unique per (tail, head).
template <typename T>
__global__ __launch_bounds__(256) void nbuild_kernel(const NPart *__restrict__ parts,
                                                     const NSub *__restrict__ subs, int nsubs,
                                                     const int *__restrict__ aoff,
                                                     const int *__restrict__ acode,
                                                     const T *__restrict__ aw, int total_rows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= total_rows) return;
  const NSub q = subs[upper_index(nsubs, row, [&](int k) { return subs[k].row0; })];
  const int a = row - q.row0;
  const int gp = parts[q.part].pos0 + q.st + a;
  T val = Dist<T>::inf();
  const int e1 = aoff[gp + 1];
  for (int e = aoff[gp]; e < e1; ++e)  // uniform
    if (acode[e] == lane) val = aw[e];
  if (lane == a) val = T(0);
  if (lane < q.n) gptr(static_cast<T *>(q.W))[(int64_t)a * kNestT + lane] = val;
}

// b: row of crossing vertex c of sub-part k in its part's crossing graph.
template <typename T>
__global__ __launch_bounds__(256) void nxgraph_kernel(const NPart *__restrict__ parts,
                                                      const NSub *__restrict__ subs, int nsubs,
                                                      const int *__restrict__ aoff,
                                                      const int *__restrict__ acode,
                                                      const T *__restrict__ aw, int total_rows) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= total_rows) return;
  const NSub q = subs[upper_index(nsubs, row, [&](int k) { return subs[k].xrow0; })];
  const NPart P = parts[q.part];
  const int c = row - q.xrow0, g = q.x0 + c;
  const int gp = P.pos0 + q.st + c;
  const int e0 = aoff[gp], e1 = aoff[gp + 1];
  const FAP_GLOBAL T *wr = gptr(static_cast<const T *>(q.W)) + (int64_t)c * kNestT;
  FAP_GLOBAL T *xr = gptr(static_cast<T *>(P.X)) + (int64_t)g * P.SX;
  for (int j = lane; j < P.nx; j += 64) {
    T val = (j >= q.x0 && j < q.x0 + q.b) ? wr[j - q.x0] : Dist<T>::inf();
    for (int e = e0; e < e1; ++e)  // uniform
      if (acode[e] == -1 - j) val = aw[e];
    if (j == g) val = T(0);
    xr[j] = val;
  }
}

// c: crossing rows r0 .. r0 + 4*RB - 1 of one part against the columns of one sub-part
// (one column per lane: sub-parts hold at most 64 vertices).
template <typename T, int RB>
__global__ __launch_bounds__(256) void nxrows_kernel(const NPart *__restrict__ parts,
                                                     const NSub *__restrict__ subs,
                                                     const NUnit *__restrict__ units) {
  constexpr int kRows = kMWaves * RB;
  __shared__ __attribute__((aligned(16))) T As[kKc][kRows + kAPad<T>];
  const NUnit un = units[blockIdx.x];
  const NSub q = subs[un.sub];
  const NPart P = parts[q.part];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = un.rb * kRows;
  MpArgs<T> m;
  m.A = gptr(static_cast<const T *>(P.X)) + (int64_t)r0 * P.SX + q.x0;
  m.lda = P.SX;
  m.rows = P.nx - r0;
  m.kdim = q.b;
  m.B = gptr(static_cast<const T *>(q.W));
  m.ldb = kNestT;
  m.col0 = 0;
  m.colmax = kNestT - 1;
  FAP_GLOBAL T *R = gptr(static_cast<T *>(P.R));
  mp_unit<T, RB, false, 1>(m, As, tid, [&](const T (&acc)[RB][1]) {
    if (lane >= q.n) return;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int g = r0 + wave * RB + r;
      if (g < P.nx) R[(int64_t)g * P.ldr + q.st + lane] = acc[r][0];
    }
  });
}

// d: rows r0 .. of sub-part q against positions cb * 256 .. of its part.
template <typename T, int RB>
__global__ __launch_bounds__(256) void nrows_kernel(const NPart *__restrict__ parts,
                                                    const NSub *__restrict__ subs,
                                                    const NUnit *__restrict__ units,
                                                    const int *__restrict__ sigma) {
  constexpr int kRows = kMWaves * RB;
  __shared__ __attribute__((aligned(16))) T As[kKc][kRows + kAPad<T>];
  const NUnit un = units[blockIdx.x];
  const NSub q = subs[un.sub];
  const NPart P = parts[q.part];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = un.rb * kRows;
  const FAP_GLOBAL T *W = gptr(static_cast<const T *>(q.W));
  MpArgs<T> m;
  m.A = W + (int64_t)r0 * kNestT;
  m.lda = kNestT;
  m.rows = q.n - r0;
  m.kdim = q.b;
  m.B = gptr(static_cast<const T *>(P.R)) + (int64_t)q.x0 * P.ldr;
  m.ldb = P.ldr;
  m.col0 = un.cb * kMCols;
  m.colmax = INT32_MAX;  // crossing rows are padded with +inf to whole column blocks
  FAP_GLOBAL T *D = gptr(static_cast<T *>(P.D));
  mp_unit<T, RB, false>(m, As, tid, [&](const T (&acc)[RB][4]) {
    int dcol[4], own[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int pos = un.cb * kMCols + e * 64 + lane;
      dcol[e] = pos < P.s ? sigma[P.pos0 + pos] : -1;
      own[e] = pos >= q.st && pos < q.st + q.n ? pos - q.st : -1;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int a = r0 + wave * RB + r;
      if (a >= q.n) continue;
      FAP_GLOBAL T *drow = D + (int64_t)sigma[P.pos0 + q.st + a] * P.S;
      const FAP_GLOBAL T *wrow = W + (int64_t)a * kNestT;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (dcol[e] < 0) continue;
        T val = acc[r][e];
        if (own[e] >= 0) val = Dist<T>::mn(val, wrow[own[e]]);
        drow[dcol[e]] = val;
      }
    }
  });
}

// Units arrive in launch slots: slot 8 * i + x holds the i-th unit of XCD x (or none),
// where all row blocks of one (part, column block) are on one XCD, so its boundary-row
// panel is read from HBM once and then from that XCD's L2 (with units pulled in global
// order every XCD fetched every panel: 7x the algorithmic reads).  Persistent grid
// (`counters`, one per XCD): workgroup b starts on XCD b % 8, takes the units of its own
// list and then helps the others; plain grid (no counters): workgroup L runs slot L.
template <typename Run>
__device__ __forceinline__ void for_each_unit(const BMerge *__restrict__ slots, int nslots,
                                              int *counters, const int *__restrict__ lens,
                                              int *unit_s, Run &&run) {
  const int tid = threadIdx.x;
  if (!counters) {  // plain grid: one slot per workgroup
    if ((int)blockIdx.x < nslots && slots[blockIdx.x].p >= 0) run(slots[blockIdx.x]);
    return;
  }
  const int home = blockIdx.x & 7;
  while (true) {
    if (tid == 0) {
      int u = -1;
      for (int t = 0; t < 8 && u < 0; ++t) {  // own list first, then the others
        const int x = (home + t) & 7;
        if (__hip_atomic_load(counters + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= lens[x])
          continue;
        const int i = atomicAdd(counters + x, 1);
        if (i < lens[x]) u = 8 * i + x;
      }
      *unit_s = u;
    }
    __syncthreads();
    // Uniform by construction; readfirstlane lets the compiler keep the unit's
    // descriptors and pointers in scalar registers (scalar loads, scalar address math).
    const int u = __builtin_amdgcn_readfirstlane(*unit_s);
    __syncthreads();  // every thread has read unit_s before thread 0 takes the next one
    if (u < 0) break;  // uniform: every list is drained
    if (slots[u].p >= 0) run(slots[u]);
  }
}

// int32 merge: a unit is (part, row block of 4 * RB inner rows, column block of 256
// columns), on the min-plus unit above.
template <typename T, int RB>
__global__ __launch_bounds__(256, 4) void bmerge_kernel(const BGraph *__restrict__ graphs,
                                                     const BPart *__restrict__ parts,
                                                     const BMerge *__restrict__ slots,
                                                     int nslots, int *counters,
                                                     const int *__restrict__ lens) {
  __shared__ __attribute__((aligned(16))) T As[kKc][kMWaves * RB + kAPad<T>];
  __shared__ int unit_s;
  const int tid = threadIdx.x;
  for_each_unit(slots, nslots, counters, lens, &unit_s, [&](const BMerge &m) {
    const BPart p = parts[m.p];
    const BGraph g = graphs[p.g];
    if (g.packed) {
      merge_unit<T, RB, true>(g, p, m.rb, m.cb, As, tid);
      return;
    }
    merge_unit<T, RB, false>(g, p, m.rb, m.cb, As, tid);
  });
}

// FP64 merge: a unit is (part, row block of 64 inner rows, column block of 256
// columns); thread t holds an 8 x 8 register tile, rows (t / 32) * 8 + {0..7} and
// columns 4 (t % 32) + {0..3}, 128 + 4 (t % 32) + {0..3}; A (64 x 16) and B (16 x 256)
// staged in LDS per k chunk of 16.  Per k: 16 LDS values for 64 steps.  The first
// form (8 rows x 4 columns per lane, B read from L1 per wave) ran at half the rate.
constexpr int kM64Rows = 64, kM64Kc = 16;

__device__ __forceinline__ void merge_unit64(const BGraph &g, const BPart &p, int rb, int cb,
                                             double (*As)[kM64Rows], double (*Bs)[kMCols],
                                             int tid) {
  const int inner = p.s - p.b, row0 = rb * kM64Rows, col0 = cb * kMCols;
  const FAP_GLOBAL double *D = gptr(static_cast<const double *>(p.D));
  const FAP_GLOBAL double *B = gptr(static_cast<const double *>(p.Brows));
  const int tx = tid & 31, ty = tid >> 5;
  double acc[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[r][c] = __builtin_inf();
  for (int k0 = 0; k0 < p.b; k0 += kM64Kc) {
    const int kc = min(kM64Kc, p.b - k0);
    {  // As[kk][i] = A[row0 + i][k0 + kk], four per thread
      const int i = tid & 63, kq = (tid >> 6) * 4;
      const bool row_ok = row0 + i < inner;
      const FAP_GLOBAL double *arow = D + (int64_t)(p.b + row0 + i) * p.S + k0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        As[kq + q][i] = (row_ok && kq + q < kc) ? arow[kq + q] : __builtin_inf();
    }
    // Bs[kk][c] = B[k0 + kk][col0 + c]: the boundary rows are +inf-padded to whole
    // column blocks; rows past kc read as +inf.
#pragma unroll
    for (int q = 0; q < kM64Kc; ++q)
      Bs[q][tid] = q < kc ? B[(int64_t)(k0 + q) * p.ldb + col0 + tid] : __builtin_inf();
    __syncthreads();
    for (int k = 0; k < kc; ++k) {
      double av[8], bv[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) av[r] = As[k][ty * 8 + r];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) bv[4 * h + e] = Bs[k][h * 128 + tx * 4 + e];
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = fmin(acc[r][c], av[r] + bv[c]);
    }
    __syncthreads();
  }
  int col[8], own[8];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int c = 4 * h + e;
      col[c] = col0 + h * 128 + tx * 4 + e;
      own[c] = col[c] < g.n && g.graph_id[col[c]] == p.part_id ? g.ed2st[col[c]] - p.off : -1;
      if (own[c] >= p.s) own[c] = -1;
    }
  FAP_GLOBAL float *out = gptr(static_cast<float *>(g.out));
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int i = row0 + ty * 8 + r;
    if (i >= inner) break;
    const FAP_GLOBAL double *drow = D + (int64_t)(p.b + i) * p.S;
    FAP_GLOBAL float *orow = out + (p.out_row + p.b + i) * g.ld_out;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (col[c] >= g.n) continue;
      double v = acc[r][c];
      if (own[c] >= 0) v = fmin(v, drow[own[c]]);
      orow[col[c]] = Dist<double>::out(v);
    }
  }
}

__global__ __launch_bounds__(256) void bmerge64_kernel(const BGraph *__restrict__ graphs,
                                                       const BPart *__restrict__ parts,
                                                       const BMerge *__restrict__ slots,
                                                       int nslots, int *counters,
                                                       const int *__restrict__ lens) {
  __shared__ __attribute__((aligned(16))) double As[kM64Kc][kM64Rows];
  __shared__ __attribute__((aligned(16))) double Bs[kM64Kc][kMCols];
  __shared__ int unit_s;
  const int tid = threadIdx.x;
  for_each_unit(slots, nslots, counters, lens, &unit_s, [&](const BMerge &m) {
    const BPart p = parts[m.p];
    merge_unit64(graphs[p.g], p, m.rb, m.cb, As, Bs, tid);
  });
}

__global__ void fill_f64_kernel(double *__restrict__ p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void reset_counter_kernel(int *c, int n) {
  const int i = threadIdx.x;
  if (i < n) c[i] = 0;
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Launch slots for units grouped by the data they share: group g is units
// [gb[g], gb[g + 1]); groups are dealt to the 8 XCDs, largest cost first, each to the
// XCD with the least cost so far, and slot 8 * i + x holds the i-th unit of XCD x
// (`none` past the end of its list).  A grid launched over the slots puts workgroup L on
// XCD L % 8, so the units of a group share one L2.
template <typename U>
std::vector<U> xcd_slots(const std::vector<U> &units, const std::vector<int> &gb,
                         const std::vector<double> &gcost, U none, int (&len)[8]) {
  const int ng = (int)gcost.size();
  std::vector<int> order(ng);
  for (int g = 0; g < ng; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gcost[a] > gcost[b]; });
  std::vector<int> lists[8];
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int g : order) {
    const int x = (int)(std::min_element(load, load + 8) - load);
    for (int u = gb[g]; u < gb[g + 1]; ++u) lists[x].push_back(u);
    load[x] += gcost[g];
  }
  size_t maxlen = 0;
  for (int x = 0; x < 8; ++x) {
    len[x] = (int)lists[x].size();
    maxlen = std::max(maxlen, lists[x].size());
  }
  std::vector<U> slots(8 * maxlen, none);
  for (int x = 0; x < 8; ++x)
    for (size_t i = 0; i < lists[x].size(); ++i) slots[8 * i + x] = units[lists[x][i]];
  return slots;
}

// Row-block sizes (template instances): stage 5 sources per unit, stage 6 rows per unit.
int pick_rows(int code, int dflt) {
  switch (code) {
    case 1: return 8;
    case 2: return 16;
    case 3: return 32;
    default: return dflt;
  }
}

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

// One kind of entries (int32 or FP64) of a batch.
struct fap_batch_kind {
  int ngraphs = 0, nparts = 0, build_rows = 0, max_nb = 0;
  int ngb = 0, gb_rows = 0, gb_max_nb = 0, nunits5 = 0, nmerge = 0, merge_units = 0, nsrc = 0;
  int rows6 = 16;  // merge rows per wave (FP64: 8, register budget)
  int rb5 = 8;     // boundary-row unit rows per wave (4 * rb5 boundary-graph rows per unit)
  int nbrows = 0;  // boundary rows of full items (one workgroup each in bperm_kernel)
  // Floyd-Warshall descriptors: the 64-tile ones first, then (FP64) the 32-tile ones,
  // each group with its own rounds.
  int nfw64 = 0, ngb64 = 0, nfw32 = 0, ngb32 = 0, max_nb32 = 0, gb_max_nb32 = 0;
  int mlen[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // merge units per XCD list
  int *d_mlen = nullptr;
  // FP64 parts of graphs with at least kWideN vertices: the 8 x 8 register-tile merge.
  BMerge *merge_w = nullptr;
  int merge_units_w = 0;
  int mlen_w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int *d_mlen_w = nullptr;
  int grid = 0;    // persistent merge workgroups
  int grid_w = 0;  // persistent workgroups of the 8 x 8 FP64 merge
  int nfw = 0;     // parts whose in-part matrix comes from Floyd-Warshall
  // In-part BFS descriptors: parts of at most 1024 vertices first (n2 of them, g2
  // source groups), then those up to kInpartBfsMaxS (n4, g4).
  int nip2 = 0, gip2 = 0, nip4 = 0, gip4 = 0;
  InpartDesc *ip = nullptr;
  char *d_mem = nullptr;  // tables and scratch in one allocation
  BGraph *graphs = nullptr;
  BPart *parts = nullptr;
  int *gbg = nullptr, *bpart = nullptr, *pid2idx = nullptr;
  int64_t *brow = nullptr;
  BRowUnit *units5 = nullptr;
  BMerge *merge = nullptr;
  BSrc *srcs = nullptr;
  void *fw_descs = nullptr, *gb_descs = nullptr;
  // int32 boundary graphs of packed graphs: 16-bit Floyd-Warshall, then expansion into
  // the int32 matrix (items of 8 entries).
  int ngb16 = 0, gb16_max_nb = 0;
  int64_t gb16_items = 0;
  Fw16Desc *gb16_descs = nullptr;
  UnpackDesc *gb16_unpack = nullptr;
  // Nested closure: parts, sub-parts, build rows (sum of sub-part sizes), crossing-graph
  // rows, product units (crossing rows, part rows), crossing-graph FW descriptors (FP64:
  // 64-tile ones, then 32-tile ones).
  int nnest = 0, nnsub = 0, nest_rows = 0, nx_rows = 0, nu3 = 0, nu4 = 0;
  int nx64 = 0, nx32 = 0, nx_max_nb = 0, nx_max_nb32 = 0;
  NPart *npart = nullptr;
  NSub *nsub = nullptr;
  int *naoff = nullptr, *nacode = nullptr, *nsigma = nullptr;
  void *naw = nullptr;
  NUnit *nu3d = nullptr, *nu4d = nullptr;
  void *nw_descs = nullptr, *nx_descs = nullptr;
};

struct fap_batch {
  fap_batch_kind kind[2];  // 0 int32, 1 FP64
  int *d_counter = nullptr;
  int per_cu = 0, rows5 = 16, rows6 = 16;
  bool packed = true;  // packed 16-bit merge where exact
  bool bfs = true;     // in-part BFS where exact
  bool unit_grid = false;  // merge as a plain grid instead of the persistent queue
  bool fw32 = false;       // FP64 Floyd-Warshall on 32 x 32 tiles
  bool nested = true;      // nested closure of large weighted parts
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Both kinds present: the FP64 kind runs on `side`, forked from and joined back into
  // the caller's stream (two branches of the captured graph).
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {

// Batched Floyd-Warshall over nd descriptors (the launchers put descriptors on the
// grid's y dimension: at most 65,535 per launch).
template <typename T>
int launch_fw_chunks(void *descs, int nd, int max_nb, int tile, hipStream_t st) {
  for (int d0 = 0; d0 < nd; d0 += 65535) {
    const int cnt = std::min(65535, nd - d0);
    if (sizeof(T) == 4) FAP_TRY(launch_fw(static_cast<FwDesc *>(descs) + d0, cnt, max_nb, st));
    else if (tile == 32)
      FAP_TRY(launch_fw64_t32(static_cast<Fw64Desc *>(descs) + d0, cnt, max_nb, st));
    else FAP_TRY(launch_fw64(static_cast<Fw64Desc *>(descs) + d0, cnt, max_nb, st));
  }
  return FAP_OK;
}

// Nested closure of the kind's large weighted parts (stages a-d above).
template <typename T>
int launch_nested(const fap_batch_kind &q, hipStream_t st) {
  hipLaunchKernelGGL(nbuild_kernel<T>, dim3((q.nest_rows + 3) / 4), dim3(256), 0, st, q.npart,
                     q.nsub, q.nnsub, q.naoff, q.nacode, static_cast<const T *>(q.naw),
                     q.nest_rows);
  FAP_CHECK_LAUNCH();
  FAP_TRY(launch_fw_chunks<T>(q.nw_descs, q.nnsub, 1, kBT, st));
  if (q.nx_rows > 0) {
    hipLaunchKernelGGL(nxgraph_kernel<T>, dim3((q.nx_rows + 3) / 4), dim3(256), 0, st, q.npart,
                       q.nsub, q.nnsub, q.naoff, q.nacode, static_cast<const T *>(q.naw),
                       q.nx_rows);
    FAP_CHECK_LAUNCH();
    FAP_TRY(launch_fw_chunks<T>(q.nx_descs, q.nx64, q.nx_max_nb, kBT, st));
    if (q.nx32 > 0)
      FAP_TRY(launch_fw_chunks<T>(static_cast<Fw64Desc *>(q.nx_descs) + q.nx64, q.nx32,
                                  q.nx_max_nb32, 32, st));
    hipLaunchKernelGGL((nxrows_kernel<T, 16>), dim3(q.nu3), dim3(256), 0, st, q.npart, q.nsub,
                       q.nu3d);
    FAP_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL((nrows_kernel<T, kNestRB<T>>), dim3(q.nu4), dim3(256), 0, st, q.npart, q.nsub,
                     q.nu4d, q.nsigma);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

template <typename T>
int launch_kind(const fap_batch *b, const fap_batch_kind &q, int *counter, hipStream_t st) {
  if (q.nparts == 0) return FAP_OK;
  if (q.nsrc > 0) {
    hipLaunchKernelGGL(bsssp_kernel<T>, dim3((q.nsrc + kSWaves - 1) / kSWaves),
                       dim3(64 * kSWaves), 0, st, q.graphs, q.parts, q.srcs, q.nsrc);
    FAP_CHECK_LAUNCH();
  }
  if (q.nfw > 0) {
    hipLaunchKernelGGL(bbuild_kernel<T>, dim3((q.build_rows + 3) / 4), dim3(256), 0, st,
                       q.graphs, q.parts, q.nparts, q.build_rows);
    FAP_CHECK_LAUNCH();
    FAP_TRY(launch_fw_chunks<T>(q.fw_descs, q.nfw64, q.max_nb, 64, st));
    if (q.nfw32 > 0)
      FAP_TRY(launch_fw_chunks<T>(static_cast<Fw64Desc *>(q.fw_descs) + q.nfw64, q.nfw32,
                                  q.max_nb32, 32, st));
  }
  if (q.nnest > 0) FAP_TRY(launch_nested<T>(q, st));
  if (q.nip2 > 0) FAP_TRY(launch_inpart_bfs(q.ip, q.nip2, q.gip2, 1024, st));
  if (q.nip4 > 0) FAP_TRY(launch_inpart_bfs(q.ip + q.nip2, q.nip4, q.gip4, kInpartBfsMaxS, st));
  if (q.ngb > 0) {
    hipLaunchKernelGGL(bgb_kernel<T>, dim3((q.gb_rows + 3) / 4), dim3(256), 0, st, q.graphs,
                       q.parts, q.gbg, q.ngb, q.bpart, q.pid2idx, q.gb_rows);
    FAP_CHECK_LAUNCH();
    if (q.ngb64 > 0) FAP_TRY(launch_fw_chunks<T>(q.gb_descs, q.ngb64, q.gb_max_nb, 64, st));
    if (q.ngb16 > 0) {
      FAP_TRY(launch_fw16(q.gb16_descs, q.ngb16, q.gb16_max_nb, st));
      FAP_TRY(launch_fw16_unpack(q.gb16_unpack, q.ngb16, q.gb16_items, nullptr, nullptr, st));
    }
    if (q.ngb32 > 0)
      FAP_TRY(launch_fw_chunks<T>(static_cast<Fw64Desc *>(q.gb_descs) + q.ngb64, q.ngb32,
                                  q.gb_max_nb32, 32, st));
  }
  if (q.nunits5 > 0) {
    const dim3 g5(q.nunits5), t5(256);
    if (q.rb5 == 4) hipLaunchKernelGGL((brows_kernel<T, 4>), g5, t5, 0, st, q.graphs, q.parts, q.units5, q.brow);
    else if (q.rb5 == 16) hipLaunchKernelGGL((brows_kernel<T, 16>), g5, t5, 0, st, q.graphs, q.parts, q.units5, q.brow);
    else hipLaunchKernelGGL((brows_kernel<T, 8>), g5, t5, 0, st, q.graphs, q.parts, q.units5, q.brow);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(bperm_kernel<T>, dim3(q.nbrows), dim3(256), 0, st, q.graphs, q.gbg, q.ngb,
                       q.brow);
    FAP_CHECK_LAUNCH();
  }
  if (q.nsrc > 0) {
    hipLaunchKernelGGL(bconv_kernel<T>, dim3(q.nsrc), dim3(256), 0, st, q.graphs, q.parts, q.srcs);
    FAP_CHECK_LAUNCH();
  }
  if (q.merge_units > 0) {
    // Persistent work queue, or (unit_grid) a plain grid of one workgroup per unit whose
    // slots free up as units end, so the other kind's launches interleave.
    const dim3 gm(b->unit_grid ? q.merge_units : std::min(q.grid, q.merge_units)), tm(256);
    int *c = b->unit_grid ? nullptr : counter;
    if (q.rows6 == 8)
      hipLaunchKernelGGL((bmerge_kernel<T, 8>), gm, tm, 0, st, q.graphs, q.parts, q.merge,
                         q.merge_units, c, q.d_mlen);
    else
      hipLaunchKernelGGL((bmerge_kernel<T, 16>), gm, tm, 0, st, q.graphs, q.parts, q.merge,
                         q.merge_units, c, q.d_mlen);
    FAP_CHECK_LAUNCH();
  }
  if (q.merge_units_w > 0) {
    const dim3 gm(b->unit_grid ? q.merge_units_w : std::min(q.grid_w, q.merge_units_w)), tm(256);
    hipLaunchKernelGGL(bmerge64_kernel, gm, tm, 0, st, q.graphs, q.parts, q.merge_w,
                       q.merge_units_w, b->unit_grid ? nullptr : counter + 16, q.d_mlen_w);
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

// The two kinds share nothing but the counter array (eight counters per merge launch:
// int32 at 0, FP64 at 8 and 24), so they run
// side by side: the Floyd-Warshall rounds of one (short, latency-bound launches) fill
// in around the other's long launches.
int batch_launch(fap_batch *b, hipStream_t st) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, st, b->d_counter, 32);
  FAP_CHECK_LAUNCH();
  if (!b->side) {
    FAP_TRY(launch_kind<int>(b, b->kind[0], b->d_counter, st));
    return launch_kind<double>(b, b->kind[1], b->d_counter + 8, st);
  }
  FAP_CHECK_HIP(hipEventRecord(b->ev_fork, st));
  FAP_CHECK_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
  FAP_TRY(launch_kind<double>(b, b->kind[1], b->d_counter + 8, b->side));
  FAP_TRY(launch_kind<int>(b, b->kind[0], b->d_counter, st));
  FAP_CHECK_HIP(hipEventRecord(b->ev_join, b->side));
  FAP_CHECK_HIP(hipStreamWaitEvent(st, b->ev_join, 0));
  return FAP_OK;
}

// Host copy of one graph's CSR and layout (the nested closure's planning input).
struct HostCsr {
  std::vector<int> row_off, col, graph_id, st2ed, ed2st, w;
  std::vector<float> wf;
};

int fetch_csr(const fap_graph_dev &d, HostCsr &h) {
  auto get = [](auto &v, const void *src, int64_t count) {
    v.resize(count);
    return count == 0 || hipMemcpy(v.data(), src, count * sizeof(v[0]), hipMemcpyDeviceToHost) ==
                             hipSuccess;
  };
  const bool ok = get(h.row_off, d.d_row_offset, (int64_t)d.n + 1) && get(h.col, d.d_col, d.m) &&
                  get(h.graph_id, d.d_graph_id, d.n) && get(h.st2ed, d.d_st2ed, d.n) &&
                  get(h.ed2st, d.d_ed2st, d.n) && (!d.d_w || get(h.w, d.d_w, d.m)) &&
                  (!d.d_wf || get(h.wf, d.d_wf, d.m));
  if (!ok) {
    (void)hipGetLastError();
    set_error("fap_hip_batch_create: graph copy to the host failed");
    return FAP_ERR_HIP;
  }
  return FAP_OK;
}

// One part's nested-closure plan: sub-parts of at most kNestT vertices, positions
// (sub-part by sub-part, crossing vertices first), and every in-part arc coded from its
// tail's position (see nbuild_kernel).
struct NestLocal {
  std::vector<int> sigma;              // position -> the part's vertex (0 .. s-1)
  std::vector<int> st, n, b, x0;       // per sub-part
  std::vector<int> aoff{0}, acode;     // arcs per position
  std::vector<double> aw;
  int nx = 0;                          // crossing vertices
};

// Sub-parts of at most kNestT vertices (ids 0 ..) by recursive bisection along BFS
// orders: a set is ordered by hop distance from a pseudo-peripheral vertex (the last one
// reached from its first vertex; further components follow in the same way) and cut
// where the first floor(k/2) of its k = ceil(size / kNestT) sub-parts end.  Linear work
// per level: a multilevel partitioner took 1-3 ms per 400-600-vertex part.
std::vector<int> bfs_sub_parts(const std::vector<std::vector<int>> &adj) {
  const int s = (int)adj.size();
  std::vector<int> sp(s, 0), in_set(s, -1), seen(s, -1), order, queue;
  std::vector<std::vector<int>> stack{std::vector<int>(s)};
  for (int i = 0; i < s; ++i) stack[0][i] = i;
  int stamp = 0, next_id = 0;
  auto bfs = [&](int root, int tag, std::vector<int> &out) {  // within the current set
    seen[root] = tag;
    queue.assign(1, root);
    for (size_t h = 0; h < queue.size(); ++h)
      for (int u : adj[queue[h]])
        if (in_set[u] == stamp && seen[u] != tag) {
          seen[u] = tag;
          queue.push_back(u);
        }
    out.insert(out.end(), queue.begin(), queue.end());
    return queue.back();
  };
  int tag = 0;
  while (!stack.empty()) {
    std::vector<int> V = std::move(stack.back());
    stack.pop_back();
    if ((int)V.size() <= kNestT) {
      for (int v : V) sp[v] = next_id;
      ++next_id;
      continue;
    }
    ++stamp;
    for (int v : V) in_set[v] = stamp;
    order.clear();
    const int probe_tag = ++tag, order_tag = ++tag;
    for (int v : V) {
      if (seen[v] == order_tag) continue;
      std::vector<int> scratch;
      const int far = bfs(v, probe_tag, scratch);  // component of v, farthest vertex
      bfs(far, order_tag, order);
    }
    const int k = ((int)V.size() + kNestT - 1) / kNestT;
    const int left = (int)((int64_t)V.size() * (k / 2) / k);
    stack.emplace_back(order.begin() + left, order.end());
    stack.emplace_back(order.begin(), order.begin() + left);
  }
  return sp;
}

bool nest_local(const HostCsr &h, int off, int s, int part_id, NestLocal &L) {
  // In-part arcs in CSR order; a later arc between the same ends replaces the earlier
  // one (the build's rule), loops are dropped (the diagonal is 0).
  std::vector<std::vector<std::pair<int, double>>> out(s);
  for (int i = 0; i < s; ++i) {
    const int v = h.st2ed[off + i];
    for (int e = h.row_off[v]; e < h.row_off[v + 1]; ++e) {
      const int u = h.col[e];
      if (h.graph_id[u] != part_id) continue;
      const int j = h.ed2st[u] - off;
      if (j == i || j < 0 || j >= s) continue;
      const double w = !h.wf.empty() ? (double)h.wf[e] : !h.w.empty() ? (double)h.w[e] : 1.0;
      auto &l = out[i];
      auto it = std::find_if(l.begin(), l.end(), [&](const std::pair<int, double> &x) {
        return x.first == j;
      });
      if (it != l.end()) it->second = w;
      else l.push_back({j, w});
    }
  }
  std::vector<std::vector<int>> und(s);
  for (int i = 0; i < s; ++i)
    for (const auto &x : out[i]) {
      und[i].push_back(x.first);
      und[x.first].push_back(i);
    }
  const std::vector<int> sp = bfs_sub_parts(und);
  std::vector<char> cross(s, 0);
  for (int i = 0; i < s; ++i)
    for (const auto &x : out[i])
      if (sp[i] != sp[x.first]) cross[i] = cross[x.first] = 1;
  L.sigma.resize(s);
  for (int i = 0; i < s; ++i) L.sigma[i] = i;
  std::sort(L.sigma.begin(), L.sigma.end(), [&](int x, int y) {
    if (sp[x] != sp[y]) return sp[x] < sp[y];
    if (cross[x] != cross[y]) return cross[x] > cross[y];
    return x < y;
  });
  std::vector<int> inv(s), subof(s);
  for (int p = 0; p < s; ++p) {
    const int i = L.sigma[p];
    inv[i] = p;
    if (p == 0 || sp[i] != sp[L.sigma[p - 1]]) {
      L.st.push_back(p);
      L.n.push_back(0);
      L.b.push_back(0);
      L.x0.push_back(L.nx);
    }
    subof[p] = (int)L.st.size() - 1;
    ++L.n.back();
    if (cross[i]) {
      ++L.b.back();
      ++L.nx;
    }
  }
  for (int p = 0; p < s; ++p) {
    const int q = subof[p];
    for (const auto &x : out[L.sigma[p]]) {
      const int pj = inv[x.first], qj = subof[pj];
      L.acode.push_back(qj == q ? pj - L.st[q] : -1 - (L.x0[qj] + pj - L.st[qj]));
      L.aw.push_back(x.second);
    }
    L.aoff.push_back((int)L.acode.size());
  }
  return true;
}

// Host plan of one kind: the tables and the scratch layout, uploaded in one block.
int plan_kind(const fap_batch *b, fap_batch_kind &q, const std::vector<const fap_batch_item *> &items,
              bool fp64) {
  const int elem = fp64 ? 8 : 4;
  q.rows6 = fp64 ? 8 : b->rows6;  // rows per unit: 4 * rows6 (FP64: register budget)
  // Boundary-source batch width (8, 16 or 32 sources) -> rows per wave of a unit of
  // 4 waves (4, 8 or 16; FP64 at most 8, register budget).
  q.rb5 = std::min(b->rows5 / 2, fp64 ? 8 : 16);
  // Persistent merge grid: as many workgroups as the device keeps resident (its
  // register and LDS use), unless the caller fixed the workgroups per CU.
  auto grid_of = [&](const void *fn) {
    int per_cu = b->per_cu;
    if (per_cu <= 0 &&
        (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
         per_cu <= 0)) {
      (void)hipGetLastError();
      per_cu = 2;
    }
    return per_cu * std::max(device_cus(), 1);
  };
  q.grid = grid_of(fp64 ? (const void *)bmerge_kernel<double, 8>
                        : (q.rows6 == 8 ? (const void *)bmerge_kernel<int, 8>
                                        : (const void *)bmerge_kernel<int, 16>));
  if (fp64) q.grid_w = grid_of((const void *)bmerge64_kernel);
  std::vector<BGraph> graphs;
  std::vector<BPart> parts;
  std::vector<int> gbg, bpart, pid2idx;
  std::vector<int64_t> brow;
  std::vector<BRowUnit> units5;
  std::vector<BSrc> srcs;
  std::vector<double> pcost;
  // Scratch offsets (bytes) of every part matrix, boundary-graph matrix and FP64 row
  // block; resolved to pointers once the allocation exists.
  std::vector<int64_t> d_at, mb_at, mb16_at, r_at, rs_at, brows_at;
  std::vector<char> emitted;  // per part: its rows are produced (merge units)
  int nbrows = 0;             // boundary-graph rows produced over all graphs
  int64_t scratch = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = scratch;
    scratch = align_up(scratch + bytes, 256);
    return o;
  };
  int build_rows = 0, max_nb = 0, gb_rows = 0, gb_max_nb = 0;
  // Nested closure tables (host), with the scratch offsets of their matrices.
  std::vector<NPart> nps;
  std::vector<NSub> nss;
  std::vector<int> np_src, np_tile, nsig, naoff{0}, nacode;
  std::vector<double> naw;
  std::vector<NUnit> u3, u4;
  std::vector<int64_t> w_at, x_at, rr_at;
  int nest_rows = 0, nx_rows = 0;
  for (const fap_batch_item *it : items) {
    const fap_graph_dev &d = it->g;
    BGraph G{};
    G.row_off = d.d_row_offset;
    G.col = d.d_col;
    G.w = d.d_w;
    G.wf = d.d_wf;
    G.graph_id = d.d_graph_id;
    G.st2ed = d.d_st2ed;
    G.ed2st = d.d_ed2st;
    G.out = it->d_out;
    G.ld_out = it->ld_out;
    G.n = d.n;
    G.ldr = align_up(d.n, kMCols);  // boundary rows: +inf padding to whole column blocks
    G.packed = !fp64 && b->packed && d.d_w == nullptr && d.d_wf == nullptr && d.n <= kCap16;
    G.tile = fp64 && (b->fw32 || (it->flags & FAP_ITEM_FW_TILE32)) ? 32 : kBT;
    G.bbase = (int)bpart.size();
    G.pid_base = (int)pid2idx.size();
    const int gi = (int)graphs.size();
    std::vector<const fap_part_desc *> ps;
    int max_pid = 0;
    for (int k = 0; k < it->nparts; ++k) {
      const fap_part_desc &pd = it->parts[k];
      FAP_REQUIRE(pd.part_id >= 1 && pd.num >= 0 && pd.bdy >= 0 && pd.bdy <= pd.num &&
                      pd.offset >= 0 && (int64_t)pd.offset + pd.num <= d.n && pd.out_row >= 0,
                  "fap_hip_batch_create: bad part %d (id %d)", k, pd.part_id);
      FAP_REQUIRE(pd.num <= kBatchMaxS, "fap_hip_batch_create: part with %d vertices (limit %d)",
                  pd.num, kBatchMaxS);
      max_pid = std::max(max_pid, pd.part_id);
      if (pd.num > 0) ps.push_back(&pd);
    }
    // Full item: the parts tile [0, n) (then the boundary-graph form).
    {
      std::vector<std::pair<int, int>> span;
      for (const fap_part_desc *pd : ps) span.push_back({pd->offset, pd->num});
      std::sort(span.begin(), span.end());
      int next = 0;
      bool cover = true;
      for (const auto &x : span) {
        cover = cover && x.first == next;
        next = x.first + x.second;
      }
      G.full = cover && next == d.n && !(it->flags & FAP_ITEM_BOUNDARY_SSSP);
    }
    G.max_pid = max_pid;
    pid2idx.resize(pid2idx.size() + max_pid + 1, -1);
    int nb = 0;
    for (const fap_part_desc *pd : ps) {
      FAP_REQUIRE(pid2idx[G.pid_base + pd->part_id] < 0,
                  "fap_hip_batch_create: part id %d repeated", pd->part_id);
      pid2idx[G.pid_base + pd->part_id] = (int)parts.size();
      BPart P{};
      P.S = (int)align_up(pd->num, G.tile);
      P.g = gi;
      P.part_id = pd->part_id;
      P.off = pd->offset;
      P.s = pd->num;
      P.b = pd->bdy;
      P.bs = nb;
      P.row_begin = build_rows;
      P.out_row = pd->out_row;
      // Unit weights on paired arcs (component labels present): the in-part distances
      // of a full item are BFS distances, symmetric.
      P.bfs = !fp64 && b->bfs && G.full && d.d_w == nullptr && d.d_wf == nullptr &&
              d.d_comp != nullptr && P.s <= kInpartBfsMaxS;
      build_rows += P.S;
      if (G.tile == kBT) max_nb = std::max(max_nb, P.S / kBT);
      else q.max_nb32 = std::max(q.max_nb32, P.S / G.tile);
      d_at.push_back(take((int64_t)P.S * P.S * elem));
      if (G.full) {
        for (int c = 0; c < P.b; ++c) {
          bpart.push_back((int)parts.size());
          brow.push_back(P.out_row + c);
        }
        nb += P.b;
        brows_at.push_back(-1);
      } else {
        // Partial item: SSSP rows (int32: the part's rows of the result).
        brows_at.push_back(take((int64_t)P.b * G.ldr * elem));
        for (int c = 0; c < P.b; ++c) srcs.push_back(BSrc{(int)parts.size(), c});
      }
      const double s = P.s, bb = P.b;
      pcost.push_back((s - bb) * bb * d.n + (s - bb) * d.n);
      parts.push_back(P);
    }
    G.nb = G.full ? nb : 0;
    // Boundary-graph rows produced: all of them, or those of the one part asked for.
    G.rlo = 0;
    G.rcnt = G.nb;
    const int pfirst = (int)parts.size() - (int)ps.size();
    emitted.resize(parts.size(), 1);
    if (it->only_part > 0) {
      FAP_REQUIRE(G.full, "fap_hip_batch_create: only_part needs parts that tile the graph");
      FAP_REQUIRE(it->only_part <= max_pid && pid2idx[G.pid_base + it->only_part] >= 0,
                  "fap_hip_batch_create: only_part %d is not a part of the item", it->only_part);
      const BPart &O = parts[pid2idx[G.pid_base + it->only_part]];
      G.rlo = O.bs;
      G.rcnt = O.b;
      for (int pi = pfirst; pi < (int)parts.size(); ++pi)
        emitted[pi] = parts[pi].part_id == it->only_part;
    }
    // Vertex-id locality of the parts: 64-byte lines (16 ids) a part's vertices touch,
    // over all parts, per vertex; scattered numbering takes the sorted-order rows.
    if (G.nb > 0) {
      std::vector<int> st(d.n);
      FAP_CHECK_HIP(hipMemcpy(st.data(), d.d_st2ed, 4LL * d.n, hipMemcpyDeviceToHost));
      int64_t lines = 0;
      std::vector<int> blk;
      for (const fap_part_desc *pd : ps) {
        blk.clear();
        for (int j = 0; j < pd->num; ++j) blk.push_back(st[pd->offset + j] >> 4);
        std::sort(blk.begin(), blk.end());
        lines += std::unique(blk.begin(), blk.end()) - blk.begin();
      }
      G.perm = lines * 4 > d.n;
    }
    if (G.nb > 0) {
      G.SB = (int)align_up(G.nb, G.tile);
      G.gb_row_begin = gb_rows;
      gb_rows += G.SB;
      if (G.tile == kBT) gb_max_nb = std::max(gb_max_nb, G.SB / kBT);
      else q.gb_max_nb32 = std::max(q.gb_max_nb32, G.SB / G.tile);
      gbg.push_back(gi);
      mb_at.push_back(take((int64_t)G.SB * G.SB * elem));
      mb16_at.push_back(!fp64 && G.packed ? take((int64_t)G.SB * G.SB * 2) : -1);
      r_at.push_back(take((int64_t)std::max(G.rcnt, 1) * G.ldr * elem));
      rs_at.push_back(G.perm ? take((int64_t)std::max(G.rcnt, 1) * G.ldr * elem) : -1);
      G.pbase = nbrows;
      nbrows += G.rcnt;
      for (int pi = pfirst; pi < (int)parts.size(); ++pi)
        for (int r0 = G.rlo; r0 < G.rlo + G.rcnt; r0 += kMWaves * q.rb5)
          for (int cb = 0; cb * kMCols < parts[pi].s; ++cb) units5.push_back(BRowUnit{pi, r0, cb, 0});
    } else {
      mb_at.push_back(-1);
      mb16_at.push_back(-1);
      r_at.push_back(-1);
      rs_at.push_back(-1);
    }
    // Nested closure of the item's large weighted parts (full items: a partial item
    // closes its matrices around SSSP rows).
    if (b->nested && G.full) {
      HostCsr hc;
      bool fetched = false;
      const int xt = fp64 ? G.tile : kBT;
      for (int pi = pfirst; pi < (int)parts.size(); ++pi) {
        BPart &P = parts[pi];
        if (P.bfs || P.s < kNestMinS) continue;
        if (!fetched) {
          FAP_TRY(fetch_csr(d, hc));
          fetched = true;
        }
        NestLocal L;
        if (!nest_local(hc, P.off, P.s, P.part_id, L)) continue;
        P.nested = 1;
        NPart N{};
        N.S = P.S;
        N.s = P.s;
        N.nx = L.nx;
        N.SX = L.nx > 0 ? (int)align_up(L.nx, xt) : 0;
        N.ldr = align_up(P.s, kMCols);
        N.pos0 = (int)nsig.size();
        const int npi = (int)nps.size();
        np_src.push_back(pi);
        np_tile.push_back(xt);
        x_at.push_back(L.nx > 0 ? take((int64_t)N.SX * N.SX * elem) : -1);
        rr_at.push_back(L.nx > 0 ? take((int64_t)L.nx * N.ldr * elem) : -1);
        if (L.nx > 0) {
          if (xt == kBT) q.nx_max_nb = std::max(q.nx_max_nb, N.SX / kBT);
          else q.nx_max_nb32 = std::max(q.nx_max_nb32, N.SX / xt);
        }
        const int rows4 = kMWaves * (fp64 ? kNestRB<double> : kNestRB<int>);
        for (size_t k = 0; k < L.st.size(); ++k) {
          NSub Q{};
          Q.part = npi;
          Q.st = L.st[k];
          Q.n = L.n[k];
          Q.b = L.b[k];
          Q.x0 = L.x0[k];
          Q.row0 = nest_rows;
          nest_rows += Q.n;
          Q.xrow0 = nx_rows;
          nx_rows += Q.b;
          w_at.push_back(take((int64_t)kNestT * kNestT * elem));
          const int qi = (int)nss.size();
          for (int rb = 0; rb * kMWaves * 16 < L.nx; ++rb) u3.push_back(NUnit{qi, rb, 0, 0});
          for (int rb = 0; rb * rows4 < Q.n; ++rb)
            for (int cb = 0; cb * kMCols < P.s; ++cb) u4.push_back(NUnit{qi, rb, cb, 0});
          nss.push_back(Q);
        }
        const int abase = (int)nacode.size();
        for (int pp = 0; pp < P.s; ++pp) {
          nsig.push_back(L.sigma[pp]);
          naoff.push_back(abase + L.aoff[pp + 1]);
        }
        nacode.insert(nacode.end(), L.acode.begin(), L.acode.end());
        naw.insert(naw.end(), L.aw.begin(), L.aw.end());
        nps.push_back(N);
      }
    }
    graphs.push_back(G);
  }
  // Merge entries: parts with inner rows, largest first.
  std::vector<int> order;
  for (int k = 0; k < (int)parts.size(); ++k)
    if (parts[k].s > parts[k].b && emitted[k]) order.push_back(k);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pcost[x] > pcost[y]; });
  // Units of a part: the row blocks of one column block consecutive (its boundary-row
  // panel stays in L2 while they run).  FP64 parts of graphs with at least kWideN
  // vertices take the 8 x 8 register-tile merge (64-row units), the others the 8 x 4
  // form (32-row units): measured on MI355X, the 512-graph batch (n <= 8463) took 26.2
  // against 31.2 ms with the 8 x 4 form, a 60k-vertex kNN graph 12 against 29 ms with
  // the 8 x 8 form.
  auto build_units = [&](bool wide, int (&len)[8]) -> std::vector<BMerge> {
    const int rows = wide ? kM64Rows : kMWaves * q.rows6;
    std::vector<BMerge> munits;
    std::vector<int> mgb{0};
    std::vector<double> mcost;
    for (int k : order) {
      const BPart &P = parts[k];
      if (fp64 && (graphs[P.g].n >= kWideN) != wide) continue;
      const int rbs = (P.s - P.b + rows - 1) / rows;
      const int cbs = (graphs[P.g].n + kMCols - 1) / kMCols;
      for (int cb = 0; cb < cbs; ++cb) {
        for (int rb = 0; rb < rbs; ++rb) munits.push_back(BMerge{k, rb, cb, 0});
        mgb.push_back((int)munits.size());
        mcost.push_back((double)rbs * rows * (P.b + 8));
      }
    }
    return xcd_slots(munits, mgb, mcost, BMerge{-1, 0, 0, 0}, len);
  };
  int64_t nunits_all = 0;
  for (int k : order)
    nunits_all += (int64_t)((parts[k].s - parts[k].b + 31) / 32) *
                  ((graphs[parts[k].g].n + kMCols - 1) / kMCols);
  FAP_REQUIRE(nunits_all < (1LL << 27), "fap_hip_batch_create: too many merge units");
  const std::vector<BMerge> merge = build_units(false, q.mlen);
  const std::vector<BMerge> merge_w = fp64 ? build_units(true, q.mlen_w) : std::vector<BMerge>();
  const int64_t units = (int64_t)merge.size();
  FAP_REQUIRE(units5.size() < (1u << 31) && srcs.size() < (1u << 31),
              "fap_hip_batch_create: batch too large");
  // Tables after the scratch, one allocation.
  auto table = [&](int64_t bytes) { return take(std::max<int64_t>(bytes, 16)); };
  const int64_t o_graphs = table(sizeof(BGraph) * graphs.size());
  const int64_t o_parts = table(sizeof(BPart) * parts.size());
  const int64_t o_gbg = table(4 * gbg.size()), o_bpart = table(4 * bpart.size());
  const int64_t o_pid = table(4 * pid2idx.size()), o_brow = table(8 * brow.size());
  const int64_t o_u5 = table(sizeof(BRowUnit) * units5.size());
  const int64_t o_merge = table(sizeof(BMerge) * merge.size());
  const int64_t o_merge_w = table(sizeof(BMerge) * merge_w.size());
  const int64_t o_mlen = table(sizeof(q.mlen)), o_mlen_w = table(sizeof(q.mlen_w));
  const int64_t o_src = table(sizeof(BSrc) * srcs.size());
  const int64_t o_fw = table((fp64 ? sizeof(Fw64Desc) : sizeof(FwDesc)) * parts.size());
  const int64_t o_ip = table(sizeof(InpartDesc) * parts.size());
  const int64_t o_gbd = table((fp64 ? sizeof(Fw64Desc) : sizeof(FwDesc)) * gbg.size());
  const int64_t o_gb16 = table(sizeof(Fw16Desc) * gbg.size());
  const int64_t o_ud16 = table(sizeof(UnpackDesc) * gbg.size());
  const size_t fwd = fp64 ? sizeof(Fw64Desc) : sizeof(FwDesc);
  const int64_t o_np = table(sizeof(NPart) * nps.size()), o_ns = table(sizeof(NSub) * nss.size());
  const int64_t o_sig = table(4 * nsig.size()), o_aoff = table(4 * naoff.size());
  const int64_t o_acode = table(4 * nacode.size()), o_aw = table((int64_t)elem * naw.size());
  const int64_t o_u3 = table(sizeof(NUnit) * u3.size()), o_u4 = table(sizeof(NUnit) * u4.size());
  const int64_t o_nwd = table(fwd * nss.size()), o_nxd = table(fwd * nps.size());
  FAP_TRY(dev_alloc((void **)&q.d_mem, std::max<int64_t>(scratch, 256)));
  char *base = q.d_mem;
  if (o_graphs > 0) {
    // Scratch to +inf (FAP_INF for int32) once: the padding columns of the boundary
    // rows are never written by the kernels and must read as unreachable; everything
    // else is rewritten on every run.
    if (fp64)
      hipLaunchKernelGGL(fill_f64_kernel, dim3(1024), dim3(256), 0, nullptr,
                         reinterpret_cast<double *>(base), o_graphs / 8, __builtin_huge_val());
    else
      FAP_TRY(fap_hip_fill_i32(reinterpret_cast<int32_t *>(base), kInf, o_graphs / 4, nullptr));
    FAP_CHECK_LAUNCH();
    FAP_CHECK_HIP(hipDeviceSynchronize());
  }
  for (size_t k = 0; k < parts.size(); ++k) {
    BPart &P = parts[k];
    P.D = base + d_at[k];
    const BGraph &G = graphs[P.g];
    P.ldb = G.ldr;
    if (brows_at[k] >= 0) P.Brows = base + brows_at[k];  // full items: set below from R
  }
  for (size_t gi = 0; gi < graphs.size(); ++gi) {
    BGraph &G = graphs[gi];
    if (mb_at[gi] >= 0) G.MB = base + mb_at[gi];
    if (mb16_at[gi] >= 0) G.MB16 = base + mb16_at[gi];
    if (r_at[gi] >= 0) G.R = base + r_at[gi];
    if (rs_at[gi] >= 0) G.Rs = base + rs_at[gi];
  }
  for (BPart &P : parts) {
    const BGraph &G = graphs[P.g];
    // (only the emitted part's rows exist when the item asks for one part)
    if (G.full && G.R && P.bs >= G.rlo && P.bs + P.b <= G.rlo + G.rcnt)
      P.Brows = static_cast<char *>(G.R) + (int64_t)(P.bs - G.rlo) * G.ldr * elem;
  }
  for (size_t k = 0; k < nps.size(); ++k) {
    NPart &N = nps[k];
    N.D = base + d_at[np_src[k]];
    if (x_at[k] >= 0) N.X = base + x_at[k];
    if (rr_at[k] >= 0) N.R = base + rr_at[k];
  }
  for (size_t k = 0; k < nss.size(); ++k) nss[k].W = base + w_at[k];
  std::vector<char> host(scratch - o_graphs);
  auto put = [&](int64_t o, const void *src, size_t bytes) {
    if (bytes) std::memcpy(host.data() + (o - o_graphs), src, bytes);
  };
  put(o_graphs, graphs.data(), sizeof(BGraph) * graphs.size());
  put(o_parts, parts.data(), sizeof(BPart) * parts.size());
  put(o_gbg, gbg.data(), 4 * gbg.size());
  put(o_bpart, bpart.data(), 4 * bpart.size());
  put(o_pid, pid2idx.data(), 4 * pid2idx.size());
  put(o_brow, brow.data(), 8 * brow.size());
  put(o_u5, units5.data(), sizeof(BRowUnit) * units5.size());
  put(o_merge, merge.data(), sizeof(BMerge) * merge.size());
  put(o_merge_w, merge_w.data(), sizeof(BMerge) * merge_w.size());
  put(o_mlen, q.mlen, sizeof(q.mlen));
  put(o_mlen_w, q.mlen_w, sizeof(q.mlen_w));
  put(o_src, srcs.data(), sizeof(BSrc) * srcs.size());
  put(o_np, nps.data(), sizeof(NPart) * nps.size());
  put(o_ns, nss.data(), sizeof(NSub) * nss.size());
  put(o_sig, nsig.data(), 4 * nsig.size());
  put(o_aoff, naoff.data(), 4 * naoff.size());
  put(o_acode, nacode.data(), 4 * nacode.size());
  put(o_u3, u3.data(), sizeof(NUnit) * u3.size());
  put(o_u4, u4.data(), sizeof(NUnit) * u4.size());
  if (fp64) {
    put(o_aw, naw.data(), 8 * naw.size());
    std::vector<Fw64Desc> wd, xd;
    for (const NSub &Q : nss) wd.push_back(Fw64Desc{static_cast<double *>(Q.W), 1});
    for (int t32 = 0; t32 < 2; ++t32)
      for (size_t k = 0; k < nps.size(); ++k)
        if (nps[k].nx > 0 && (np_tile[k] == 32) == (t32 == 1))
          xd.push_back(Fw64Desc{static_cast<double *>(nps[k].X), nps[k].SX / np_tile[k]});
    q.nx32 = 0;
    for (size_t k = 0; k < nps.size(); ++k) q.nx32 += nps[k].nx > 0 && np_tile[k] == 32;
    q.nx64 = (int)xd.size() - q.nx32;
    put(o_nwd, wd.data(), sizeof(Fw64Desc) * wd.size());
    put(o_nxd, xd.data(), sizeof(Fw64Desc) * xd.size());
  } else {
    std::vector<int> awi(naw.size());
    for (size_t e = 0; e < naw.size(); ++e) awi[e] = (int)naw[e];
    put(o_aw, awi.data(), 4 * awi.size());
    std::vector<FwDesc> wd, xd;
    for (const NSub &Q : nss) wd.push_back(FwDesc{static_cast<int *>(Q.W), kNestT, 1});
    for (const NPart &N : nps)
      if (N.nx > 0) xd.push_back(FwDesc{static_cast<int *>(N.X), N.SX, N.SX / kBT});
    q.nx64 = (int)xd.size();
    put(o_nwd, wd.data(), sizeof(FwDesc) * wd.size());
    put(o_nxd, xd.data(), sizeof(FwDesc) * xd.size());
  }
  if (fp64) {
    std::vector<Fw64Desc> fd, gd;
    // 64-tile descriptors first, then the 32-tile ones (their own launch sequence).
    for (int t32 = 0; t32 < 2; ++t32) {
      for (const BPart &P : parts)
        if (!P.nested && (graphs[P.g].tile == 32) == (t32 == 1))
          fd.push_back(Fw64Desc{static_cast<double *>(P.D), P.S / graphs[P.g].tile});
      for (int gi : gbg)
        if ((graphs[gi].tile == 32) == (t32 == 1))
          gd.push_back(Fw64Desc{static_cast<double *>(graphs[gi].MB), graphs[gi].SB / graphs[gi].tile});
      if (t32 == 0) {
        q.nfw64 = (int)fd.size();
        q.ngb64 = (int)gd.size();
      }
    }
    q.nfw32 = (int)fd.size() - q.nfw64;
    q.ngb32 = (int)gd.size() - q.ngb64;
    put(o_fw, fd.data(), sizeof(Fw64Desc) * fd.size());
    put(o_gbd, gd.data(), sizeof(Fw64Desc) * gd.size());
    q.nfw = (int)fd.size();
  } else {
    std::vector<FwDesc> fd, gd;
    std::vector<InpartDesc> ip2, ip4;
    for (const BPart &P : parts) {
      if (P.nested) continue;
      if (!P.bfs) {
        fd.push_back(FwDesc{static_cast<int *>(P.D), P.S, P.S / kBT});
        continue;
      }
      const BGraph &G = graphs[P.g];
      std::vector<InpartDesc> &l = P.s <= 1024 ? ip2 : ip4;
      int &groups = P.s <= 1024 ? q.gip2 : q.gip4;
      l.push_back(InpartDesc{static_cast<int *>(P.D), P.S, P.s, P.off, P.part_id, groups, 0,
                             G.row_off, G.col, G.graph_id, G.st2ed, G.ed2st});
      groups += (P.s + 63) / 64;
    }
    // Boundary graphs of packed graphs on 16 bits (half the VALU steps of int32).
    std::vector<Fw16Desc> g16;
    std::vector<UnpackDesc> u16;
    gb_max_nb = 0;
    for (int gi : gbg) {
      const BGraph &G = graphs[gi];
      if (G.MB16) {
        g16.push_back(Fw16Desc{static_cast<uint16_t *>(G.MB16), G.SB, G.SB / kBT});
        u16.push_back(UnpackDesc{static_cast<const uint16_t *>(G.MB16), static_cast<int *>(G.MB),
                                 G.SB, G.nb, 0, 0, q.gb16_items});
        q.gb16_items += (int64_t)G.SB * G.SB / 8;
        q.gb16_max_nb = std::max(q.gb16_max_nb, G.SB / kBT);
        continue;
      }
      gd.push_back(FwDesc{static_cast<int *>(G.MB), G.SB, G.SB / kBT});
      gb_max_nb = std::max(gb_max_nb, G.SB / kBT);
    }
    put(o_fw, fd.data(), sizeof(FwDesc) * fd.size());
    put(o_gbd, gd.data(), sizeof(FwDesc) * gd.size());
    put(o_gb16, g16.data(), sizeof(Fw16Desc) * g16.size());
    put(o_ud16, u16.data(), sizeof(UnpackDesc) * u16.size());
    q.ngb16 = (int)g16.size();
    q.nfw = q.nfw64 = (int)fd.size();
    q.ngb64 = (int)gd.size();
    q.nip2 = (int)ip2.size();
    q.nip4 = (int)ip4.size();
    ip2.insert(ip2.end(), ip4.begin(), ip4.end());
    put(o_ip, ip2.data(), sizeof(InpartDesc) * ip2.size());
  }
  if (hipMemcpy(base + o_graphs, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    set_error("fap_hip_batch_create: table upload failed");
    return FAP_ERR_HIP;
  }
  q.ngraphs = (int)graphs.size();
  q.nparts = (int)parts.size();
  q.build_rows = build_rows;
  q.max_nb = max_nb;
  q.ngb = (int)gbg.size();
  q.gb_rows = gb_rows;
  q.gb_max_nb = gb_max_nb;
  q.nunits5 = (int)units5.size();
  q.nbrows = nbrows;
  q.nmerge = (int)merge.size();
  q.merge_units = (int)units;
  q.nsrc = (int)srcs.size();
  q.graphs = reinterpret_cast<BGraph *>(base + o_graphs);
  q.parts = reinterpret_cast<BPart *>(base + o_parts);
  q.gbg = reinterpret_cast<int *>(base + o_gbg);
  q.bpart = reinterpret_cast<int *>(base + o_bpart);
  q.pid2idx = reinterpret_cast<int *>(base + o_pid);
  q.brow = reinterpret_cast<int64_t *>(base + o_brow);
  q.units5 = reinterpret_cast<BRowUnit *>(base + o_u5);
  q.merge = reinterpret_cast<BMerge *>(base + o_merge);
  q.d_mlen = reinterpret_cast<int *>(base + o_mlen);
  q.merge_w = reinterpret_cast<BMerge *>(base + o_merge_w);
  q.merge_units_w = (int)merge_w.size();
  q.d_mlen_w = reinterpret_cast<int *>(base + o_mlen_w);
  q.srcs = reinterpret_cast<BSrc *>(base + o_src);
  q.fw_descs = base + o_fw;
  q.ip = reinterpret_cast<InpartDesc *>(base + o_ip);
  q.gb_descs = base + o_gbd;
  q.gb16_descs = reinterpret_cast<Fw16Desc *>(base + o_gb16);
  q.gb16_unpack = reinterpret_cast<UnpackDesc *>(base + o_ud16);
  q.nnest = (int)nps.size();
  q.nnsub = (int)nss.size();
  q.nest_rows = nest_rows;
  q.nx_rows = nx_rows;
  q.nu3 = (int)u3.size();
  q.nu4 = (int)u4.size();
  q.npart = reinterpret_cast<NPart *>(base + o_np);
  q.nsub = reinterpret_cast<NSub *>(base + o_ns);
  q.nsigma = reinterpret_cast<int *>(base + o_sig);
  q.naoff = reinterpret_cast<int *>(base + o_aoff);
  q.nacode = reinterpret_cast<int *>(base + o_acode);
  q.naw = base + o_aw;
  q.nu3d = reinterpret_cast<NUnit *>(base + o_u3);
  q.nu4d = reinterpret_cast<NUnit *>(base + o_u4);
  q.nw_descs = base + o_nwd;
  q.nx_descs = base + o_nxd;
  return FAP_OK;
}

}  // namespace

extern "C" int fap_hip_batch_destroy(fap_batch *b) {
  if (!b) return FAP_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->side) {
    (void)hipStreamSynchronize(b->side);
    (void)hipStreamDestroy(b->side);
  }
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  for (auto &q : b->kind)
    if (q.d_mem) dev_free(q.d_mem);
  if (b->d_counter) dev_free(b->d_counter);
  delete b;
  return FAP_OK;
}

extern "C" int fap_hip_batch_create(fap_batch **out, const fap_batch_item *items, int32_t nitems,
                                    int32_t flags) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(out && nitems >= 0 && (nitems == 0 || items), "fap_hip_batch_create: null argument");
  FAP_REQUIRE((flags & ~(FAP_BATCH_GRAPH | 0xFF00 | 0xF0000 | FAP_BATCH_NO_PACKED |
                         FAP_BATCH_NO_BFS | FAP_BATCH_UNIT_GRID | FAP_BATCH_FW_TILE32 |
                         FAP_BATCH_NO_NESTED)) == 0,
              "fap_hip_batch_create: unknown flags %d", flags);
  *out = nullptr;
  std::vector<const fap_batch_item *> kinds[2];
  for (int g = 0; g < nitems; ++g) {
    const fap_batch_item &it = items[g];
    const fap_graph_dev &d = it.g;
    FAP_REQUIRE(d.n > 0 && d.d_row_offset && d.d_graph_id && d.d_st2ed && d.d_ed2st &&
                    (d.m == 0 || d.d_col) && it.d_out && it.ld_out >= d.n &&
                    (it.nparts == 0 || it.parts),
                "fap_hip_batch_create: bad item %d", g);
    kinds[it.is_float ? 1 : 0].push_back(&it);
  }
  fap_batch *b = new fap_batch;
  b->use_graph = (flags & FAP_BATCH_GRAPH) != 0;
  b->per_cu = (flags >> 8) & 0xFF;  // merge workgroups per CU (0: as many as fit)
  b->rows5 = pick_rows((flags >> 16) & 0x3, 16);
  b->rows6 = pick_rows((flags >> 18) & 0x3, 16) == 8 ? 8 : 16;
  b->packed = (flags & FAP_BATCH_NO_PACKED) == 0;
  b->bfs = (flags & FAP_BATCH_NO_BFS) == 0;
  b->unit_grid = (flags & FAP_BATCH_UNIT_GRID) != 0;
  b->fw32 = (flags & FAP_BATCH_FW_TILE32) != 0;
  b->nested = (flags & FAP_BATCH_NO_NESTED) == 0;
  int s = dev_alloc((void **)&b->d_counter, 256);
  for (int k = 0; k < 2 && s == FAP_OK; ++k)
    if (!kinds[k].empty()) s = plan_kind(b, b->kind[k], kinds[k], k == 1);
  if (s == FAP_OK && b->kind[0].nparts > 0 && b->kind[1].nparts > 0) {
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      set_error("fap_hip_batch_create: stream or event creation failed");
      s = FAP_ERR_HIP;
    }
  }
  if (s == FAP_OK && b->use_graph) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      set_error("fap_hip_batch_create: stream creation failed");
      s = FAP_ERR_HIP;
    } else {
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        s = batch_launch(b, st);
        hipGraph_t gr = nullptr;
        e = hipStreamEndCapture(st, &gr);
        if (s == FAP_OK && e == hipSuccess) {
          b->graph = gr;
          e = hipGraphInstantiate(&b->exec, gr, nullptr, nullptr, 0);
        } else if (gr) {
          (void)hipGraphDestroy(gr);
        }
      }
      if (s == FAP_OK && e != hipSuccess) {
        set_error("fap_hip_batch_create: graph capture failed: %s", hipGetErrorString(e));
        s = FAP_ERR_HIP;
      }
      (void)hipStreamDestroy(st);
    }
  }
  if (s != FAP_OK) {
    fap_hip_batch_destroy(b);
    return s;
  }
  *out = b;
  return FAP_OK;
}

extern "C" int fap_hip_batch_nested_parts(const fap_batch *b) {
  FAP_REQUIRE(b, "fap_hip_batch_nested_parts: null batch");
  return b->kind[0].nnest + b->kind[1].nnest;
}

extern "C" int fap_hip_batch_run(fap_batch *b, void *stream) {
  FAP_REQUIRE(b, "fap_hip_batch_run: null batch");
  hipStream_t st = resolve(stream);
  if (b->exec) {
    FAP_CHECK_HIP(hipGraphLaunch(b->exec, st));
    return FAP_OK;
  }
  return batch_launch(b, st);
}

namespace fap_hip {

// A full solve of one graph through the batch stages: in-part distances, the boundary
// graph closed by Floyd-Warshall (no limit on its size), boundary rows by one min-plus
// pass, merge.  The single-graph pipeline takes this route for full solves its own
// boundary-graph form does not cover (float weights; more than its boundary limit):
// the alternative there is one label-correcting SSSP per boundary vertex, which on a
// weighted road graph of 62k vertices took 1.8-3 s.  Timed as one span (stage `stage`).
int solve_one_batched(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                      void *d_out, int64_t ld_out, bool is_float, int stage, hipStream_t st) {
  fap_batch_item it{};
  it.g = *g;
  it.parts = parts;
  it.nparts = nparts;
  it.is_float = is_float ? 1 : 0;
  it.d_out = d_out;
  it.ld_out = ld_out;
  fap_batch *b = nullptr;
  // One run: the nested closure's host planning would cost more than it saves.
  FAP_TRY(fap_hip_batch_create(&b, &it, 1, FAP_BATCH_NO_NESTED));
  hipEvent_t e0, e1;
  int s = timing_record(st, &e0);
  if (s == FAP_OK) s = fap_hip_batch_run(b, st);
  if (s == FAP_OK) s = timing_record(st, &e1);
  if (s == FAP_OK) {
    timing_add_span(stage, e0, e1);
    timing_note_bgraph();
  }
  // The batch's scratch is freed by destroy: wait for its work first.
  if (hipStreamSynchronize(st) != hipSuccess && s == FAP_OK) {
    set_error("solve_one_batched: stream synchronize failed");
    s = FAP_ERR_HIP;
  }
  fap_hip_batch_destroy(b);
  return s;
}

}  // namespace fap_hip

extern "C" int fap_hip_solve_batch(const fap_batch_item *items, int32_t nitems, int32_t mode,
                                   void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(nitems >= 0 && (nitems == 0 || items), "fap_hip_solve_batch: null argument");
  const int form = mode & 0xFF;
  FAP_REQUIRE(form == FAP_MODE_PLAIN || form == FAP_MODE_BATCHED || form == FAP_MODE_OVERLAP,
              "fap_hip_solve_batch: unknown mode %d", mode);
  hipStream_t st = resolve(stream);
  if (form == FAP_MODE_PLAIN) {
    // The reference's structure: graph after graph, each a per-part launch sequence.
    for (int g = 0; g < nitems; ++g) {
      const fap_batch_item &it = items[g];
      if (it.is_float)
        FAP_TRY(fap_hip_solve_parts_f32(&it.g, it.parts, it.nparts, static_cast<float *>(it.d_out),
                                        it.ld_out, FAP_MODE_PLAIN, st));
      else
        FAP_TRY(fap_hip_solve_parts(&it.g, it.parts, it.nparts, static_cast<int32_t *>(it.d_out),
                                    it.ld_out, FAP_MODE_PLAIN | FAP_MODE_FLAG_INT32_FW, st));
    }
    return FAP_OK;
  }
  fap_batch *b = nullptr;
  FAP_TRY(fap_hip_batch_create(&b, items, nitems, FAP_BATCH_NO_NESTED));  // one run
  // Stage times of the call (fap_hip_last_stage_times): the whole batch as the merge span.
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int s = timing_start(st);
  if (s == FAP_OK) s = timing_record(st, &e0);
  if (s == FAP_OK) s = fap_hip_batch_run(b, st);
  if (s == FAP_OK) s = timing_record(st, &e1);
  if (s == FAP_OK) {
    timing_add_span(3, e0, e1);
    for (int i = 0; i < nitems; ++i)
      if (items[i].only_part > 0) timing_note_bgraph();
    s = timing_finish(st);
  }
  if (s == FAP_OK && hipStreamSynchronize(st) != hipSuccess) {
    set_error("fap_hip_solve_batch: stream synchronize failed");
    s = FAP_ERR_HIP;
  }
  fap_hip_batch_destroy(b);
  return s;
}
