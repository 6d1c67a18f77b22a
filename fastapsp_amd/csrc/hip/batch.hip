// Many independent graphs in one persistent launch fed by a device work queue.
//
// The work table holds one entry per part of every graph (the unit fapGraph::run
// processes, fap/fap.cpp:195-503).  Per distance type two launches:
//   1 boundary SSSP   one wave per boundary source of every part of every graph,
//                     label-correcting relaxation with atomicMin on the row,
//                     frontier queues in LDS (batch_sssp.h:232-252).  SSSP is
//                     latency-bound (levels x memory round trips), so it runs as
//                     one wide launch where thousands of waves hide each other's
//                     latency, not inside the per-part workgroups;
// then a persistent grid of 512-thread workgroups pulls parts from an atomic
// counter, largest first (the host sorts the table by a cost model), and runs the
// rest of the per-part pipeline inside the workgroup (eight waves):
//   2 build           the s x s matrix in the workgroup's scratch slot
//                     (util_Myalgorithm.h:44-84, later arc wins, diagonal 0);
//   3 Floyd-Warshall  blocked, 64 x 64 tiles: the pivot closed in LDS by the whole
//                     workgroup, then the panel and remaining tiles shared by the
//                     waves, each tile a register-tiled min-plus product staged
//                     through the wave's own LDS slab (floyd.h:38-49);
//   4 merge           inner rows x n columns = FW(inner, boundary) (x) SSSP rows,
//                     64 x 64 output tiles per wave, the part's own columns taken
//                     from the FW matrix in the epilogue (minplus.h:51-64 and
//                     util_Myalgorithm.h:236-247);
//   5 boundary rows   the SSSP rows with their own columns from the FW matrix.
// Parts never depend on each other, so the only synchronisation is the workgroup
// barrier between stages; a workgroup leaves when the counter passes the table end,
// which every workgroup reaches.  Templated on the distance type: int32 (integer
// weights, FAP_INF) and double (float weights, float32 output with MAXVALUE), the
// arithmetic of fp64.hip.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "gmem.h"

namespace fap_hip {
namespace {

constexpr int kBT = 64;        // tile edge
constexpr int kBKc = 16;       // k per LDS chunk of a tile product
constexpr int kBQ = 1024;      // frontier queue entries per wave
constexpr int kBWaves = 8;     // waves per workgroup of the persistent kernel
constexpr int kBThreads = 64 * kBWaves;
constexpr int kP1Per = kBT * kBT / kBThreads;  // pivot entries per thread in phase 1
constexpr int kSWaves = 4;     // waves (sources) per workgroup of the SSSP launch
constexpr int kBatchMaxS = 4096;  // largest part the batch form takes

struct BatchWork {
  const int *row_off, *col, *w;  // CSR (w: int32 weights or NULL)
  const float *wf;               // float weights (double path) or NULL
  const int *graph_id, *st2ed, *ed2st;
  void *out;                     // n x ld_out result of the graph (int32 or float32)
  int64_t ld_out;
  void *rows;                    // the part's SSSP rows (int32: its boundary rows of out)
  int64_t ldr;
  int64_t out_row;               // the part's first row in out
  int n, part_id, offset, num, bdy, pad;
};

// One boundary source: part `work`, its q-th boundary vertex.
struct BatchSrc {
  int work, q;
};

template <typename T>
struct Dist;

template <>
struct Dist<int> {
  typedef int Out;
  static __device__ __forceinline__ int inf() { return kInf; }
  static __device__ __forceinline__ int mn(int a, int b) { return min(a, b); }
  static __device__ __forceinline__ int weight(const BatchWork &wk, int e) {
    return wk.w ? wk.w[e] : 1;
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
  static __device__ __forceinline__ double weight(const BatchWork &wk, int e) {
    return wk.wf ? (double)wk.wf[e] : 1.0;
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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ SSSP
template <typename T>
__device__ __forceinline__ void relax(const BatchWork &wk, T *row, int u, int *q_next,
                                      int *q_count) {
  const T du = Dist<T>::load_agent(row + u);
  if (!(du < Dist<T>::inf())) return;
  const int e1 = wk.row_off[u + 1];
  for (int e = wk.row_off[u]; e < e1; ++e) {
    const int v = wk.col[e];
    const T nd = du + Dist<T>::weight(wk, e);
    if (nd < row[v] && nd < Dist<T>::atomic_min(row + v, nd)) {
      const int pos = atomicAdd(q_count, 1);
      if (pos < kBQ) q_next[pos] = v;
    }
  }
}

// One wave: the full row of source s (labels decrease only, so a stale plain read
// of row[v] can only let an unnecessary atomic through).
template <typename T>
__device__ void wave_sssp(const BatchWork &wk, int s, T *row, int *queue, int *qcount, int lane) {
  for (int v = lane; v < wk.n; v += 64) row[v] = (v == s) ? T(0) : Dist<T>::inf();
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
      for (int i = lane; i < qsize; i += 64) relax(wk, row, queue[cur * kBQ + i], q_next, q_cnt);
    } else {
      for (int u = lane; u < wk.n; u += 64) relax(wk, row, u, q_next, q_cnt);
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

// ------------------------------------------------------------------ tile products
__device__ __forceinline__ int orow(int t, int q) { return (q < 4 ? 0 : 32) + t * 4 + (q & 3); }

// acc = min(acc, A (x) B) for a 64-row A block (rows at a, stride lda, columns
// k in [0, kdim)) and a B block (rows k at b, stride ldb, 64 columns); entries
// outside [0, kdim) or outside the valid rows/columns read as +inf.  One wave,
// staged through its LDS slab As[kBKc][64], Bs[kBKc][64].
template <typename T>
__device__ __forceinline__ void wave_product(T (&acc)[8][8], const FAP_GLOBAL T *a, int64_t lda,
                                             int arows, const FAP_GLOBAL T *b, int64_t ldb,
                                             int bcols, int kdim, T *As, T *Bs, int lane) {
  const int tr = lane >> 3, tc = lane & 7;
  typedef T vec4 __attribute__((ext_vector_type(4)));
  for (int k0 = 0; k0 < kdim; k0 += kBKc) {
    if (lane < arows && k0 + kBKc <= kdim) {
      // A rows start 64-element aligned and k0 is a multiple of 16: vector loads.
      const FAP_GLOBAL vec4 *src = reinterpret_cast<const FAP_GLOBAL vec4 *>(a + (int64_t)lane * lda + k0);
#pragma unroll
      for (int c = 0; c < kBKc / 4; ++c) {
        const vec4 v = src[c];
#pragma unroll
        for (int e = 0; e < 4; ++e) As[(4 * c + e) * kBT + lane] = v[e];
      }
    } else {
#pragma unroll
      for (int q = 0; q < kBKc; ++q) {
        const int k = k0 + q;
        As[q * kBT + lane] =
            (lane < arows && k < kdim) ? a[(int64_t)lane * lda + k] : Dist<T>::inf();
      }
    }
#pragma unroll
    for (int q = 0; q < kBKc; ++q) {
      const int k = k0 + q;
      Bs[q * kBT + lane] = (lane < bcols && k < kdim) ? b[(int64_t)k * ldb + lane] : Dist<T>::inf();
    }
    wave_sync();
    const int kc = min(kBKc, kdim - k0);
    for (int k = 0; k < kc; ++k) {
      T av[8], bv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        av[q] = As[k * kBT + orow(tr, q)];
        bv[q] = Bs[k * kBT + orow(tc, q)];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[q][c] = Dist<T>::mn(acc[q][c], av[q] + bv[c]);
    }
    wave_sync();  // the slab is refilled next chunk
  }
}

template <typename T>
__device__ __forceinline__ void load_tile(T (&acc)[8][8], const FAP_GLOBAL T *t, int64_t ld,
                                          int lane) {
  const int tr = lane >> 3, tc = lane & 7;
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[q][c] = t[(int64_t)orow(tr, q) * ld + orow(tc, c)];
}

template <typename T>
__device__ __forceinline__ void store_tile(const T (&acc)[8][8], FAP_GLOBAL T *t, int64_t ld,
                                           int lane) {
  const int tr = lane >> 3, tc = lane & 7;
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int c = 0; c < 8; ++c) t[(int64_t)orow(tr, q) * ld + orow(tc, c)] = acc[q][c];
}

// ------------------------------------------------------------------ one part
template <typename T>
__device__ void process_part(const BatchWork &wk, T *Mgen, char *lds) {
  // The work record and the scratch slot hold generic pointers: every access goes
  // through the global address space (flat accesses also count against the LDS counter,
  // so the LDS staging writes after them waited for memory).
  FAP_GLOBAL T *M = gptr(Mgen);
  const FAP_GLOBAL T *rows = gptr(static_cast<const T *>(wk.rows));
  const FAP_GLOBAL int *st2ed = gptr(wk.st2ed), *row_off = gptr(wk.row_off), *col = gptr(wk.col);
  const FAP_GLOBAL int *graph_id = gptr(wk.graph_id), *ed2st = gptr(wk.ed2st);
  const int64_t ldr = wk.ldr;
  typedef typename Dist<T>::Out Out;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int num = wk.num, bdy = wk.bdy, off = wk.offset, n = wk.n;
  const int S = (num + kBT - 1) / kBT * kBT, nb = S / kBT;
  FAP_GLOBAL Out *out = gptr(static_cast<Out *>(wk.out));

  // 2. build (the SSSP rows were written by batch_sssp_kernel)
  for (int i = 0; i < S; ++i)
    for (int j = tid; j < S; j += kBThreads)
      M[(int64_t)i * S + j] = (i < bdy && j < num) ? rows[(int64_t)i * ldr + st2ed[off + j]]
                                                   : Dist<T>::inf();
  __syncthreads();
  for (int i = bdy + tid; i < num; i += kBThreads) {
    const int ver = st2ed[off + i];
    for (int e = row_off[ver]; e < row_off[ver + 1]; ++e) {
      const int v = col[e];
      if (graph_id[v] != wk.part_id) continue;
      const int idx = ed2st[v] - off;
      if (idx >= 0 && idx < num) M[(int64_t)i * S + idx] = Dist<T>::weight(wk, e);
    }
  }
  __syncthreads();
  for (int i = tid; i < num; i += kBThreads) M[(int64_t)i * S + i] = T(0);
  __syncthreads();

  // 3. Floyd-Warshall
  T *As = reinterpret_cast<T *>(lds) + wave * 2 * kBKc * kBT;
  T *Bs = As + kBKc * kBT;
  for (int r = 0; r < nb; ++r) {
    {  // phase 1: pivot tile closed in LDS, kP1Per consecutive entries per thread
      T *D = reinterpret_cast<T *>(lds);
      FAP_GLOBAL T *g = M + (int64_t)r * kBT * S + r * kBT;
      const int row = tid / (kBT / kP1Per), c0 = (tid % (kBT / kP1Per)) * kP1Per;
#pragma unroll
      for (int c = 0; c < kP1Per; ++c) D[row * kBT + c0 + c] = g[(int64_t)row * S + c0 + c];
      for (int k = 0; k < kBT; ++k) {
        __syncthreads();  // row k and column k do not change during step k
        const T dik = D[row * kBT + k];
#pragma unroll
        for (int c = 0; c < kP1Per; ++c)
          D[row * kBT + c0 + c] = Dist<T>::mn(D[row * kBT + c0 + c], dik + D[k * kBT + c0 + c]);
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < kP1Per; ++c) g[(int64_t)row * S + c0 + c] = D[row * kBT + c0 + c];
    }
    __syncthreads();
    // phase 2: row tiles (r, o) and column tiles (o, r), o != r
    const FAP_GLOBAL T *P = M + (int64_t)r * kBT * S + r * kBT;
    for (int t = wave; t < 2 * (nb - 1); t += kBWaves) {
      const bool is_row = t < nb - 1;
      int o = is_row ? t : t - (nb - 1);
      o += (o >= r);
      FAP_GLOBAL T *tile =
          is_row ? M + (int64_t)r * kBT * S + o * kBT : M + (int64_t)o * kBT * S + r * kBT;
      T acc[8][8];
      load_tile(acc, tile, S, lane);
      if (is_row) wave_product(acc, P, S, kBT, tile, S, kBT, kBT, As, Bs, lane);
      else wave_product(acc, tile, S, kBT, P, S, kBT, kBT, As, Bs, lane);
      store_tile(acc, tile, S, lane);
    }
    __syncthreads();
    // phase 3: tiles outside the pivot row and column
    for (int t = wave; t < (nb - 1) * (nb - 1); t += kBWaves) {
      int i = t / (nb - 1), j = t - i * (nb - 1);
      i += (i >= r);
      j += (j >= r);
      FAP_GLOBAL T *tile = M + (int64_t)i * kBT * S + j * kBT;
      T acc[8][8];
      load_tile(acc, tile, S, lane);
      wave_product(acc, M + (int64_t)i * kBT * S + r * kBT, S, kBT,
                   M + (int64_t)r * kBT * S + j * kBT, S, kBT, kBT, As, Bs, lane);
      store_tile(acc, tile, S, lane);
    }
    __syncthreads();
  }

  // 4. merge: inner rows x all columns, own columns from the FW matrix
  const int mrows = num - bdy;
  if (mrows > 0) {
    const int rbs = (mrows + kBT - 1) / kBT, cbs = (n + kBT - 1) / kBT;
    const int tr = lane >> 3, tc = lane & 7;
    for (int t = wave; t < rbs * cbs; t += kBWaves) {
      const int rb = t / cbs, cb = t - rb * cbs;
      const int row0 = rb * kBT, col0 = cb * kBT;
      T acc[8][8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[q][c] = Dist<T>::inf();
      if (bdy > 0)
        wave_product(acc, M + (int64_t)(bdy + row0) * S, S, min(kBT, mrows - row0), rows + col0,
                     ldr, min(kBT, n - col0), bdy, As, Bs, lane);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int j = col0 + orow(tc, c);
        if (j >= n) continue;
        const int own = graph_id[j] == wk.part_id ? ed2st[j] - off : -1;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int i = row0 + orow(tr, q);
          if (i >= mrows) continue;
          const T v = own >= 0 ? M[(int64_t)(bdy + i) * S + own] : acc[q][c];
          out[(wk.out_row + bdy + i) * wk.ld_out + j] = Dist<T>::out(v);
        }
      }
    }
  }
  __syncthreads();

  // 5. boundary rows (after the merge, which reads them as its B operand)
  for (int i = 0; i < bdy; ++i)
    for (int j = tid; j < n; j += kBThreads) {
      const T v = graph_id[j] == wk.part_id ? M[(int64_t)i * S + ed2st[j] - off]
                                               : rows[(int64_t)i * ldr + j];
      out[(wk.out_row + i) * wk.ld_out + j] = Dist<T>::out(v);
    }
  __syncthreads();
}

// Boundary SSSP of every part: one wave per source.
template <typename T>
__global__ __launch_bounds__(64 * kSWaves) void batch_sssp_kernel(const BatchWork *__restrict__ work,
                                                                  const BatchSrc *__restrict__ srcs,
                                                                  int nsrc) {
  __shared__ int queues[kSWaves][2 * kBQ + 2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kSWaves + wave;
  if (i >= nsrc) return;  // whole wave leaves; no workgroup barrier is used
  const BatchSrc sr = srcs[i];
  const BatchWork wk = work[sr.work];
  T *row = static_cast<T *>(wk.rows) + (int64_t)sr.q * wk.ldr;
  wave_sssp<T>(wk, wk.st2ed[wk.offset + sr.q], row, queues[wave], queues[wave] + 2 * kBQ, lane);
}

// Persistent grid: each workgroup owns one scratch slot (an S_max x S_max matrix)
// and pulls parts until the counter passes nwork.
template <typename T>
__global__ __launch_bounds__(kBThreads) void batch_kernel(const BatchWork *__restrict__ work,
                                                          int nwork, int *counter, char *scratch,
                                                          int64_t slot_bytes) {
  extern __shared__ __attribute__((aligned(16))) char lds[];
  __shared__ int item_s;
  T *M = reinterpret_cast<T *>(scratch + (int64_t)blockIdx.x * slot_bytes);
  while (true) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int item = item_s;
    __syncthreads();
    if (item >= nwork) break;  // uniform over the workgroup
    process_part<T>(work[item], M, lds);
  }
}

__global__ void reset_counter_kernel(int *c, int n) {
  const int i = threadIdx.x;
  if (i < n) c[i] = 0;
}

size_t lds_bytes(int elem) {
  const size_t pivot = (size_t)kBT * kBT * elem;
  const size_t slabs = (size_t)kBWaves * 2 * kBKc * kBT * elem;
  return std::max(pivot, slabs);
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

// One kind of entries (int32 or double) of a batch.
struct fap_batch_kind {
  int nwork = 0, grid = 0, nsrc = 0;
  BatchWork *d_work = nullptr;
  BatchSrc *d_src = nullptr;
  char *d_scratch = nullptr;  // scratch slots, then (double path) the SSSP rows
  int64_t slot_bytes = 0;
};

struct fap_batch {
  fap_batch_kind kind[2];  // 0 int32, 1 double
  int *d_counter = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t captured_on = nullptr;
};

static int batch_launch(fap_batch *b, hipStream_t st) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, st, b->d_counter, 2);
  FAP_CHECK_LAUNCH();
  for (int k = 0; k < 2; ++k) {
    fap_batch_kind &q = b->kind[k];
    if (q.nwork == 0) continue;
    const dim3 sgrid((q.nsrc + kSWaves - 1) / kSWaves);
    if (k == 0) {
      if (q.nsrc > 0)
        hipLaunchKernelGGL(batch_sssp_kernel<int>, sgrid, dim3(64 * kSWaves), 0, st, q.d_work,
                           q.d_src, q.nsrc);
      hipLaunchKernelGGL(batch_kernel<int>, dim3(q.grid), dim3(kBThreads), lds_bytes(4), st,
                         q.d_work, q.nwork, b->d_counter + k, q.d_scratch, q.slot_bytes);
    } else {
      if (q.nsrc > 0)
        hipLaunchKernelGGL(batch_sssp_kernel<double>, sgrid, dim3(64 * kSWaves), 0, st, q.d_work,
                           q.d_src, q.nsrc);
      hipLaunchKernelGGL(batch_kernel<double>, dim3(q.grid), dim3(kBThreads), lds_bytes(8), st,
                         q.d_work, q.nwork, b->d_counter + k, q.d_scratch, q.slot_bytes);
    }
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

extern "C" int fap_hip_batch_destroy(fap_batch *b) {
  if (!b) return FAP_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto &q : b->kind) {
    if (q.d_work) dev_free(q.d_work);
    if (q.d_src) dev_free(q.d_src);
    if (q.d_scratch) dev_free(q.d_scratch);
  }
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
  FAP_REQUIRE((flags & ~(FAP_BATCH_GRAPH | 0xFF00)) == 0, "fap_hip_batch_create: unknown flags %d",
              flags);
  const int per_cu = ((flags >> 8) & 0xFF) ? ((flags >> 8) & 0xFF) : 2;  // workgroups per CU
  *out = nullptr;
  std::vector<BatchWork> work[2];
  std::vector<double> cost[2];
  int smax[2] = {0, 0};
  for (int g = 0; g < nitems; ++g) {
    const fap_batch_item &it = items[g];
    const fap_graph_dev &d = it.g;
    FAP_REQUIRE(d.n > 0 && d.d_row_offset && d.d_graph_id && d.d_st2ed && d.d_ed2st &&
                    (d.m == 0 || d.d_col) && it.d_out && it.ld_out >= d.n &&
                    (it.nparts == 0 || it.parts),
                "fap_hip_batch_create: bad item %d", g);
    const int k = it.is_float ? 1 : 0;
    for (int p = 0; p < it.nparts; ++p) {
      const fap_part_desc &q = it.parts[p];
      FAP_REQUIRE(q.part_id >= 1 && q.num >= 0 && q.bdy >= 0 && q.bdy <= q.num && q.offset >= 0 &&
                      (int64_t)q.offset + q.num <= d.n && q.out_row >= 0,
                  "fap_hip_batch_create: item %d part %d bad", g, p);
      FAP_REQUIRE(q.num <= kBatchMaxS, "fap_hip_batch_create: part with %d vertices (limit %d)",
                  q.num, kBatchMaxS);
      if (q.num == 0) continue;
      BatchWork w{d.d_row_offset, d.d_col, d.d_w, d.d_wf, d.d_graph_id, d.d_st2ed, d.d_ed2st,
                  it.d_out, it.ld_out, nullptr, 0, q.out_row, d.n, q.part_id, q.offset, q.num,
                  q.bdy, 0};
      work[k].push_back(w);
      const double s = q.num, b = q.bdy, inner = q.num - q.bdy;
      cost[k].push_back(s * s * s + inner * b * d.n + b * (d.n + (double)d.m) * 4.0);
      smax[k] = std::max(smax[k], (int)align_up(q.num, kBT));
    }
  }
  fap_batch *b = new fap_batch;
  b->use_graph = (flags & FAP_BATCH_GRAPH) != 0;
  int s = dev_alloc((void **)&b->d_counter, 256);
  static bool attr_set = false;
  if (s == FAP_OK && !attr_set) {
    if (hipFuncSetAttribute((const void *)batch_kernel<int>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(4)) !=
            hipSuccess ||
        hipFuncSetAttribute((const void *)batch_kernel<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(8)) !=
            hipSuccess) {
      set_error("fap_hip_batch_create: cannot raise the LDS limit");
      s = FAP_ERR_HIP;
    }
    attr_set = s == FAP_OK;
  }
  int dev = 0, cus = 0;
  if (s == FAP_OK) {
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
  }
  for (int k = 0; k < 2 && s == FAP_OK; ++k) {
    const int nw = (int)work[k].size();
    if (nw == 0) continue;
    // Largest parts first (longest-processing-time order).
    std::vector<int> order(nw);
    for (int i = 0; i < nw; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return cost[k][x] > cost[k][y]; });
    std::vector<BatchWork> sorted(nw);
    for (int i = 0; i < nw; ++i) sorted[i] = work[k][order[i]];
    fap_batch_kind &q = b->kind[k];
    const int elem = k ? 8 : 4;
    q.nwork = nw;
    q.grid = std::min(nw, per_cu * cus);
    q.slot_bytes = align_up((int64_t)smax[k] * smax[k] * elem, 256);
    // Scratch slots: at most 16 GiB in all (fewer resident workgroups for big parts).
    q.grid = (int)std::max<int64_t>(1, std::min<int64_t>(q.grid, (16LL << 30) / q.slot_bytes));
    // SSSP rows: the int32 path writes a part's boundary rows of the result in place;
    // the double path gets b_p x roundup(n, 2) doubles per part after the slots.
    const int64_t slots_bytes = q.slot_bytes * q.grid;
    int64_t rows_bytes = 0;
    std::vector<int64_t> rows_at(nw, 0);
    std::vector<BatchSrc> srcs;
    for (int i = 0; i < nw; ++i) {
      BatchWork &w = sorted[i];
      if (k) {
        w.ldr = align_up(w.n, 2);
        rows_at[i] = slots_bytes + rows_bytes;
        rows_bytes += align_up((int64_t)w.bdy * w.ldr * 8, 256);
      } else {
        w.ldr = w.ld_out;
        w.rows = static_cast<int32_t *>(w.out) + w.out_row * w.ld_out;
      }
      for (int j = 0; j < w.bdy; ++j) srcs.push_back(BatchSrc{i, j});
    }
    q.nsrc = (int)srcs.size();
    s = dev_alloc((void **)&q.d_scratch, std::max<int64_t>(slots_bytes + rows_bytes, 256));
    if (s == FAP_OK && k)
      for (int i = 0; i < nw; ++i) sorted[i].rows = q.d_scratch + rows_at[i];
    if (s == FAP_OK) s = dev_alloc((void **)&q.d_work, sizeof(BatchWork) * nw);
    if (s == FAP_OK) s = dev_alloc((void **)&q.d_src, sizeof(BatchSrc) * std::max(q.nsrc, 1));
    if (s == FAP_OK &&
        (hipMemcpy(q.d_work, sorted.data(), sizeof(BatchWork) * nw, hipMemcpyHostToDevice) !=
             hipSuccess ||
         (q.nsrc > 0 && hipMemcpy(q.d_src, srcs.data(), sizeof(BatchSrc) * q.nsrc,
                                  hipMemcpyHostToDevice) != hipSuccess))) {
      set_error("fap_hip_batch_create: work table upload failed");
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

extern "C" int fap_hip_batch_run(fap_batch *b, void *stream) {
  FAP_REQUIRE(b, "fap_hip_batch_run: null batch");
  hipStream_t st = resolve(stream);
  if (b->exec) {
    FAP_CHECK_HIP(hipGraphLaunch(b->exec, st));
    return FAP_OK;
  }
  return batch_launch(b, st);
}

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
  FAP_TRY(fap_hip_batch_create(&b, items, nitems, 0));
  int s = fap_hip_batch_run(b, st);
  if (s == FAP_OK && hipStreamSynchronize(st) != hipSuccess) {
    set_error("fap_hip_solve_batch: stream synchronize failed");
    s = FAP_ERR_HIP;
  }
  fap_hip_batch_destroy(b);
  return s;
}
