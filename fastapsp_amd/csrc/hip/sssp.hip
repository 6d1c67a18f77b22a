// Boundary SSSP: one wave64 per source, frontier relaxation over the CSR.
//
// Replaces batched_sssp_path (fap/kernel/batch_sssp.h:232-252), which runs one
// binary-heap Dijkstra (dijkstra_path, :79-114) per boundary vertex, and its GPU
// twin multi_source_Nvidia_sssp_data_on_gpu (fap/kernel/cuda/batch_sssp.cu:361-414),
// which walks one source at a time with a host round trip per level.
//
// Design (gfx950):
//  * Every source owns one wave64 and one output row; the output row itself is the
//    visited / label array, so the only HBM traffic beyond the CSR is the row.
//  * Frontier queues live in LDS (two ping-pong queues per wave).  Waves never
//    synchronise with each other, so one workgroup holds several independent
//    sources and there is no grid-wide or host synchronisation per level.
//  * Unit weights: level-synchronous BFS.  A vertex is claimed with one
//    atomicCAS(INF -> level+1); a plain load filters already-claimed neighbours
//    (labels only ever decrease, so a stale read can only cause a failed CAS).
//  * Integer weights: label-correcting relaxation with atomicMin; a vertex whose
//    label dropped is queued for the next round.
//  * A frontier larger than the LDS queue switches the next round to a scan of
//    the row (BFS: label == level; weighted: every vertex) with agent-coherent
//    loads, so no vertex is lost and memory use stays bounded.
#include "fap_internal.h"

namespace fap_hip {

constexpr int kSsspWaves = 4;     // waves (sources) per workgroup
constexpr int kSsspQcap = 1024;   // entries per frontier queue per wave

__device__ __forceinline__ int load_agent(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders LDS traffic between lanes of one wave (the LDS unit executes a wave's
// operations in order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool kVecFill>
__device__ __forceinline__ void fill_row(int *row, int n, int src, int lane) {
  if (kVecFill) {
    const int n4 = n >> 2;
    int4 *r4 = reinterpret_cast<int4 *>(row);
    for (int i = lane; i < n4; i += 64) {
      int4 v = make_int4(kInf, kInf, kInf, kInf);
      const int b = i << 2;
      if (src >= b && src < b + 4) {
        if (src == b) v.x = 0;
        else if (src == b + 1) v.y = 0;
        else if (src == b + 2) v.z = 0;
        else v.w = 0;
      }
      r4[i] = v;
    }
    for (int v = (n4 << 2) + lane; v < n; v += 64) row[v] = (v == src) ? 0 : kInf;
  } else {
    for (int v = lane; v < n; v += 64) row[v] = (v == src) ? 0 : kInf;
  }
}

// Claims every unvisited out-neighbour of u with label `d` (BFS).
__device__ __forceinline__ void bfs_expand(const int *__restrict__ row_off,
                                           const int *__restrict__ col, int *row, int u, int d,
                                           int *q_next, int *q_count) {
  const int e0 = row_off[u];
  const int e1 = row_off[u + 1];
  for (int e = e0; e < e1; e += 4) {
    int v[4], dv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = (e + t < e1) ? col[e + t] : -1;
#pragma unroll
    for (int t = 0; t < 4; ++t) dv[t] = (v[t] >= 0) ? row[v[t]] : 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (v[t] >= 0 && dv[t] == kInf) {
        if (atomicCAS(row + v[t], kInf, d) == kInf) {
          const int pos = atomicAdd(q_count, 1);
          if (pos < kSsspQcap) q_next[pos] = v[t];
        }
      }
    }
  }
}

// Relaxes every out-arc of u with its current label (integer weights).
__device__ __forceinline__ void sp_expand(const int *__restrict__ row_off,
                                          const int *__restrict__ col,
                                          const int *__restrict__ w, int *row, int u,
                                          int *q_next, int *q_count) {
  const int du = load_agent(row + u);
  if (du >= kInf) return;
  const int e0 = row_off[u];
  const int e1 = row_off[u + 1];
  for (int e = e0; e < e1; e += 4) {
    int v[4], nd[4], dv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = e + t < e1;
      v[t] = ok ? col[e + t] : -1;
      nd[t] = ok ? du + w[e + t] : kInf;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) dv[t] = (v[t] >= 0) ? row[v[t]] : 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (v[t] >= 0 && nd[t] < dv[t]) {
        if (nd[t] < atomicMin(row + v[t], nd[t])) {
          const int pos = atomicAdd(q_count, 1);
          if (pos < kSsspQcap) q_next[pos] = v[t];
        }
      }
    }
  }
}

// One wave per source.  rows[i] (or i * ld when rows == nullptr) is the output
// row of source i.
template <bool kWeighted, bool kVecFill>
__global__ __launch_bounds__(64 * kSsspWaves) void sssp_rows_kernel(
    const int *__restrict__ row_off, const int *__restrict__ col, const int *__restrict__ w,
    int n, const int *__restrict__ src, int nsrc, const int64_t *__restrict__ rows,
    int *__restrict__ out, int64_t ld) {
  __shared__ int queue[kSsspWaves][2][kSsspQcap];
  __shared__ int qcount[kSsspWaves][2];

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int sidx = blockIdx.x * kSsspWaves + wave;
  if (sidx >= nsrc) return;  // whole wave leaves; no workgroup barrier is used below

  const int s = src[sidx];
  int *row = out + (rows ? rows[sidx] : (int64_t)sidx * ld);

  fill_row<kVecFill>(row, n, s, lane);
  // The fill must be complete at L2 before the first atomic on this row.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");

  if (lane == 0) {
    queue[wave][0][0] = s;
    qcount[wave][0] = 1;
    qcount[wave][1] = 0;
  }
  wave_sync();

  int cur = 0;
  int qsize = 1;
  bool scan = false;
  int level = 0;
  while (true) {
    const int nxt = cur ^ 1;
    int *q_next = queue[wave][nxt];
    int *q_cnt = &qcount[wave][nxt];
    if (!scan) {
      const int *q_cur = queue[wave][cur];
      for (int i = lane; i < qsize; i += 64) {
        const int u = q_cur[i];
        if (kWeighted) sp_expand(row_off, col, w, row, u, q_next, q_cnt);
        else bfs_expand(row_off, col, row, u, level + 1, q_next, q_cnt);
      }
    } else {
      for (int u = lane; u < n; u += 64) {
        if (kWeighted) {
          sp_expand(row_off, col, w, row, u, q_next, q_cnt);
        } else if (load_agent(row + u) == level) {
          bfs_expand(row_off, col, row, u, level + 1, q_next, q_cnt);
        }
      }
    }
    wave_sync();
    const int cnt = __hip_atomic_load(q_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync();
    if (cnt == 0) break;
    scan = cnt > kSsspQcap;
    qsize = scan ? 0 : cnt;
    if (lane == 0) qcount[wave][cur] = 0;
    wave_sync();
    cur = nxt;
    ++level;
  }
}

// ---------------------------------------------------------------------------
// Unit-weight BFS with the visited set in LDS (n <= kBfsLdsMaxN).
//
// Per level the only global round trip is one 16-byte ELL load per frontier
// vertex (4 neighbour slots; a vertex of degree > 4 stores -2 in slot 3 and its
// remaining arcs are read from the CSR).  Claims are ds_or on the LDS bitmap; the
// label is written once with a plain store.  Each lane takes up to 4 frontier
// vertices per pass and issues all their loads before using any.
constexpr int kEllW = 4;
constexpr int kBfsQ = 1024;
constexpr int kBfsU = 4;  // frontier vertices per lane per pass (256 per wave)
constexpr int kBfsLdsMaxN = 1 << 20;

__global__ void ell_build_kernel(const int *__restrict__ row_off, const int *__restrict__ col, int n,
                                 int4 *__restrict__ ell) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int e0 = row_off[v], deg = row_off[v + 1] - e0;
  int s[kEllW];
#pragma unroll
  for (int t = 0; t < kEllW; ++t) s[t] = t < deg ? col[e0 + t] : -1;
  if (deg > kEllW) s[kEllW - 1] = -2;
  ell[v] = make_int4(s[0], s[1], s[2], s[3]);
}

// Exclusive prefix sum over the wave of x in [0, 63] (six ballots), and the total.
__device__ __forceinline__ int wave_excl_prefix(int x, int *total) {
  int pre = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    const unsigned long long m = __ballot((x >> b) & 1);
    pre += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
    tot += __popcll(m) << b;
  }
  *total = tot;
  return pre;
}

// The frontier of one level: q[0, main) and q[kBfsQ - tail, kBfsQ).
struct Frontier {
  int main, tail;
  __device__ int size() const { return main + tail; }
  __device__ int at(const int *q, int idx) const {
    return idx < main ? q[idx] : q[kBfsQ - tail + (idx - main)];
  }
};

// Expands up to kBfsU vertices per lane (u[t] < 0 = none).  All ELL loads are
// issued first, then all bitmap claims; queue slots of the main region come from a
// wave prefix sum (the queue belongs to this wave).  Arcs past an ELL row (degree
// > 4) are claimed one by one into the tail region with an LDS counter.
__device__ __forceinline__ void bfs_expand_batch(const int4 *__restrict__ ell,
                                                 const int *__restrict__ row_off,
                                                 const int *__restrict__ col, unsigned *vis,
                                                 int *row, const int (&u)[kBfsU], int d,
                                                 int *q_next, int &next_main, int *tail_cnt) {
  int4 nb[kBfsU];
#pragma unroll
  for (int t = 0; t < kBfsU; ++t) nb[t] = u[t] >= 0 ? ell[u[t]] : make_int4(-1, -1, -1, -1);
  int cand[4 * kBfsU];
#pragma unroll
  for (int t = 0; t < kBfsU; ++t) {
    cand[4 * t + 0] = nb[t].x;
    cand[4 * t + 1] = nb[t].y;
    cand[4 * t + 2] = nb[t].z;
    cand[4 * t + 3] = nb[t].w;
  }
  unsigned old[4 * kBfsU];
#pragma unroll
  for (int c = 0; c < 4 * kBfsU; ++c)
    old[c] = cand[c] >= 0 ? atomicOr(&vis[cand[c] >> 5], 1u << (cand[c] & 31)) : ~0u;
  unsigned fresh = 0;
#pragma unroll
  for (int c = 0; c < 4 * kBfsU; ++c)
    if (cand[c] >= 0 && !(old[c] & (1u << (cand[c] & 31)))) fresh |= 1u << c;
  int total;
  int pos = next_main + wave_excl_prefix(__popc(fresh), &total);
#pragma unroll
  for (int c = 0; c < 4 * kBfsU; ++c) {
    if (fresh & (1u << c)) {
      row[cand[c]] = d;
      if (pos < kBfsQ) q_next[pos] = cand[c];
      ++pos;
    }
  }
  next_main += total;
#pragma unroll
  for (int t = 0; t < kBfsU; ++t) {
    if (nb[t].w != -2) continue;
    const int e1 = row_off[u[t] + 1];
    for (int e = row_off[u[t]] + kEllW - 1; e < e1; ++e) {
      const int v = col[e];
      const unsigned bit = 1u << (v & 31);
      if (atomicOr(&vis[v >> 5], bit) & bit) continue;
      row[v] = d;
      const int k = atomicAdd(tail_cnt, 1);
      if (k < kBfsQ) q_next[kBfsQ - 1 - k] = v;
    }
  }
}

template <bool kVecFill>
__global__ __launch_bounds__(64) void bfs_lds_kernel(const int4 *__restrict__ ell,
                                                     const int *__restrict__ row_off,
                                                     const int *__restrict__ col, int n,
                                                     const int *__restrict__ src, int nsrc,
                                                     const int64_t *__restrict__ rows,
                                                     int *__restrict__ out, int64_t ld) {
  extern __shared__ int smem[];
  const int nwords = (n + 31) >> 5;
  unsigned *vis = reinterpret_cast<unsigned *>(smem);
  int *queue = smem + ((nwords + 3) & ~3);  // [2][kBfsQ]
  int *tail_cnt = queue + 2 * kBfsQ;        // [2]
  const int lane = threadIdx.x;
  const int sidx = blockIdx.x;
  if (sidx >= nsrc) return;
  const int s = src[sidx];
  int *row = out + (rows ? rows[sidx] : (int64_t)sidx * ld);

  fill_row<kVecFill>(row, n, s, lane);
  // The fill must be complete before any label store to the same row.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  for (int w = lane; w < nwords; w += 64) vis[w] = 0;
  wave_sync();
  if (lane == 0) {
    vis[s >> 5] = 1u << (s & 31);
    queue[0] = s;
    tail_cnt[0] = 0;
    tail_cnt[1] = 0;
  }
  wave_sync();

  int cur = 0, level = 0;
  Frontier f{1, 0};
  bool scan = false;
  while (true) {
    const int nxt = cur ^ 1;
    int *q_next = queue + nxt * kBfsQ;
    int *t_cnt = &tail_cnt[nxt];
    const int d = level + 1;
    int next_main = 0;
    if (!scan) {
      const int *q_cur = queue + cur * kBfsQ;
      const int fs = f.size();
      for (int base = 0; base < fs; base += 64 * kBfsU) {
        int u[kBfsU];
#pragma unroll
        for (int t = 0; t < kBfsU; ++t) {
          const int idx = base + t * 64 + lane;
          u[t] = idx < fs ? f.at(q_cur, idx) : -1;
        }
        bfs_expand_batch(ell, row_off, col, vis, row, u, d, q_next, next_main, t_cnt);
      }
    } else {
      // The frontier did not fit in LDS: find it in the row.  The release fence
      // waits for this wave's label stores, the agent-scope loads read them at L2.
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      for (int base = 0; base < n; base += 64 * kBfsU) {
        int u[kBfsU];
#pragma unroll
        for (int t = 0; t < kBfsU; ++t) {
          const int v = base + t * 64 + lane;
          u[t] = (v < n && load_agent(row + v) == level) ? v : -1;
        }
        bfs_expand_batch(ell, row_off, col, vis, row, u, d, q_next, next_main, t_cnt);
      }
    }
    wave_sync();
    const int tail = __hip_atomic_load(t_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    wave_sync();
    const int cnt = next_main + tail;
    if (cnt == 0) break;
    scan = cnt > kBfsQ;
    f = Frontier{scan ? 0 : next_main, scan ? 0 : tail};
    if (lane == 0) tail_cnt[cur] = 0;
    wave_sync();
    cur = nxt;
    ++level;
  }
}

int bfs_lds_bytes(int n) { return 4 * (((n + 31) / 32 + 3) / 4 * 4 + 2 * kBfsQ + 4); }

int launch_ell_build(const int *d_row_off, const int *d_col, int n, int4 *d_ell,
                     hipStream_t stream) {
  if (n <= 0) return FAP_OK;
  hipLaunchKernelGGL(ell_build_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_row_off,
                     d_col, n, d_ell);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// Unit weights with an ELL adjacency (built by launch_ell_build) and n within
// the LDS bitmap limit; returns FAP_ERR_INVALID when the LDS path does not apply.
int launch_bfs_lds(const int4 *d_ell, const int *d_row_off, const int *d_col, int n,
                   const int *d_src, int nsrc, const int64_t *d_rows, int *d_out, int64_t ld,
                   bool vec_fill, hipStream_t stream) {
  if (nsrc <= 0) return FAP_OK;
  if (n > kBfsLdsMaxN) return FAP_ERR_INVALID;
  const int lds = bfs_lds_bytes(n);
  static bool attr_set = false;
  if (!attr_set) {
    FAP_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&bfs_lds_kernel<true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    FAP_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&bfs_lds_kernel<false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  if (vec_fill)
    hipLaunchKernelGGL(bfs_lds_kernel<true>, dim3(nsrc), dim3(64), lds, stream, d_ell, d_row_off,
                       d_col, n, d_src, nsrc, d_rows, d_out, ld);
  else
    hipLaunchKernelGGL(bfs_lds_kernel<false>, dim3(nsrc), dim3(64), lds, stream, d_ell, d_row_off,
                       d_col, n, d_src, nsrc, d_rows, d_out, ld);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_sssp(const int *d_row_off, const int *d_col, const int *d_w, int n, const int *d_src,
                int nsrc, const int64_t *d_rows, int *d_out, int64_t ld, bool vec_fill,
                hipStream_t stream) {
  if (nsrc <= 0) return FAP_OK;
  dim3 grid((nsrc + kSsspWaves - 1) / kSsspWaves), block(64 * kSsspWaves);
  if (d_w) {
    if (vec_fill)
      hipLaunchKernelGGL((sssp_rows_kernel<true, true>), grid, block, 0, stream, d_row_off, d_col,
                         d_w, n, d_src, nsrc, d_rows, d_out, ld);
    else
      hipLaunchKernelGGL((sssp_rows_kernel<true, false>), grid, block, 0, stream, d_row_off, d_col,
                         d_w, n, d_src, nsrc, d_rows, d_out, ld);
  } else {
    if (vec_fill)
      hipLaunchKernelGGL((sssp_rows_kernel<false, true>), grid, block, 0, stream, d_row_off, d_col,
                         d_w, n, d_src, nsrc, d_rows, d_out, ld);
    else
      hipLaunchKernelGGL((sssp_rows_kernel<false, false>), grid, block, 0, stream, d_row_off,
                         d_col, d_w, n, d_src, nsrc, d_rows, d_out, ld);
  }
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_sssp_batch(const int32_t *d_row_offset, const int32_t *d_col,
                                  const int32_t *d_w, int32_t n, int64_t m,
                                  const int32_t *d_src, int32_t nsrc, int32_t *d_out,
                                  int64_t ld_out, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(n >= 0 && m >= 0 && nsrc >= 0 && ld_out >= n,
              "fap_hip_sssp_batch: bad sizes n=%d m=%lld nsrc=%d ld=%lld", n, (long long)m, nsrc,
              (long long)ld_out);
  if (nsrc == 0 || n == 0) return FAP_OK;
  FAP_REQUIRE(d_row_offset && d_src && d_out && (m == 0 || d_col),
              "fap_hip_sssp_batch: null pointer");
  const bool vec = (ld_out % 4 == 0) && ((uintptr_t)d_out % 16 == 0);
  hipStream_t st = resolve(stream);
  if (!d_w && n <= kBfsLdsMaxN) {
    static Workspace ws_ell;
    FAP_TRY(ws_ell.ensure((int64_t)n * sizeof(int4)));
    int4 *ell = static_cast<int4 *>(ws_ell.ptr);
    FAP_TRY(launch_ell_build(d_row_offset, d_col, n, ell, st));
    FAP_TRY(launch_bfs_lds(ell, d_row_offset, d_col, n, d_src, nsrc, nullptr, d_out, ld_out, vec,
                           st));
    // The ELL workspace is shared between calls: finish before it can be reused.
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    return FAP_OK;
  }
  return launch_sssp(d_row_offset, d_col, d_w, n, d_src, nsrc, nullptr, d_out, ld_out, vec, st);
}
