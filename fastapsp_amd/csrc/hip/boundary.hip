// Boundary rows of every part from the boundary graph, in place of one SSSP per
// boundary vertex over the whole graph.
//
// The reference computes the b_p rows of part p (distances from p's boundary
// vertices to every vertex) with one Dijkstra per boundary vertex
// (handle_boundry_path, fap/kernel/batch_sssp.h:280-303 -> dijkstra_path :79-114).
// When the pipeline solves every part of a graph, the same rows follow from the
// parts' own Floyd-Warshall results:
//   * any path leaves or enters a part through its boundary vertices (both ends of
//     every cut arc are boundary vertices, fap/utils/util_preprocess.h:32-58), so a
//     path between boundary vertices splits into cut arcs and stretches inside one
//     part between two of its boundary vertices;
//   * hence the distances between all boundary vertices are the all-pairs distances
//     of the boundary graph G_B: vertices = every part's boundary, arcs = each part's
//     in-part distances between its boundary vertices (the leading b_q x b_q block
//     of its in-part Floyd-Warshall matrix) plus the cut arcs with their weights;
//   * a path from boundary vertex g to vertex v of part q enters q for the last time
//     at some boundary vertex c of q and stays inside q afterwards, so
//        d(g, v) = min_{c in bdy(q)} D_B(g, c) + d_q(c, v)
//     with d_q the in-part distance (c = v's own part entry included: D_B(g, g) = 0).
// Every quantity is an exact integer min/plus expression, so the rows equal the
// SSSP rows bit for bit.  The inner rows then follow as in the SSSP form: for inner
// vertex i of part p, d(i, j) = min_b d_p(i, b) + d(b, j) over p's boundary, and for
// j in p additionally d_p(i, j) (min-plus merge, minplus.hip).
//
// Kernels:
//   bg_vert_kernel   one thread per G_B vertex: its clique in-weights (a column of
//                    its part's boundary block) and its cut arcs;
//   bg_closure_kernel one workgroup per G_B vertex s: row s of the closed G_B
//                    matrix by Bellman-Ford sweeps in LDS (clique pulls, cut-arc
//                    pushes), as many sweeps as shortest paths cross parts;
//   bg_rows_kernel   d(g, j) for 1024 output columns per workgroup and a share of
//                    the G_B rows, staged through LDS; each lane keeps its columns'
//                    in-part distances from their part's boundary in registers and
//                    the wave's stores are coalesced, so the rows go to HBM once
//                    (bg_rows_wide_kernel: parts with more than 32 boundary
//                    vertices, walked in register chunks).  Both also record the
//                    largest finite distance written (the merge's 16-bit bound).
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"
#include "mp_tile.h"

namespace fap_hip {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_part_by_bs(const BgPart *__restrict__ parts, int np, int g) {
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (parts[mid].bs <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// G_B target of cut arc e out of a vertex of part `pid`, or -1 for an in-part arc.
__device__ __forceinline__ int cut_target(int e, int pid, const BgPart *__restrict__ parts,
                                          const int *__restrict__ pidx_by_pid, int npid,
                                          const int *__restrict__ col,
                                          const int *__restrict__ graph_id,
                                          const int *__restrict__ ed2st) {
  const int u = col[e];
  const int pu = graph_id[u];
  if (pu == pid || pu < 0 || pu >= npid) return -1;  // in-part arc, or a label outside the table
  const BgPart q = parts[pidx_by_pid[pu]];
  const int lk = ed2st[u] - q.offset;
  return (lk >= 0 && lk < q.bdy) ? q.bs + lk : -1;
}

// One thread per G_B vertex g: its BgVert record, its clique in-weights
// (wt[woff + c] = d_q(c, k), column k of the part's boundary block) and the output
// row that holds its distance row.
__global__ void bg_vert_kernel(const BgPart *__restrict__ parts, int np, int nb_total,
                               const int *__restrict__ pidx_by_pid, int npid, const int *__restrict__ row_off,
                               const int *__restrict__ col, const int *__restrict__ w,
                               const int *__restrict__ graph_id, const int *__restrict__ st2ed,
                               const int *__restrict__ ed2st, BgVert *__restrict__ verts,
                               int *__restrict__ wt, int64_t *__restrict__ orow) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nb_total) return;
  const BgPart p = parts[find_part_by_bs(parts, np, g)];
  const int k = g - p.bs;
  orow[g] = p.out_row + k;
  BgVert bv;
  bv.bs = p.bs;
  bv.bdy = p.bdy;
  bv.woff = p.cb + k * p.bdy;
  for (int c = 0; c < p.bdy; ++c) wt[bv.woff + c] = gptr(p.fw)[(int64_t)c * p.S + k];
  int nc = 0;
  const int v = st2ed[p.offset + k];
  for (int e = row_off[v]; e < row_off[v + 1]; ++e) {
    const int t = cut_target(e, p.part_id, parts, pidx_by_pid, npid, col, graph_id, ed2st);
    if (t < 0) continue;
    if (nc < kBgCutW) {
      bv.cut_t[nc] = t;
      bv.cut_w[nc] = w ? w[e] : 1;
    }
    ++nc;
  }
  for (int i = nc; i < kBgCutW; ++i) {
    bv.cut_t[i] = g;  // unused slots: a zero-weight self arc, which never improves
    bv.cut_w[i] = 0;
  }
  bv.ncut = nc;
  verts[g] = bv;
}

constexpr int kClThreads = 256;
constexpr int kClPad = 8;  // slack entries after the clique weights and the LDS row

// Row s of the closed G_B matrix by Bellman-Ford sweeps, one workgroup per source,
// distances in LDS.  A sweep (a) pulls every vertex v over its part's clique,
// dist[v] = min(dist[v], min_c dist[bs + c] + d_q(c, k)) -- the in-part distances
// between boundary vertices are already closed, so one pull settles a part -- and
// (b) pushes every cut arc with an LDS atomicMin.  Values only decrease and are
// always lengths of real paths; a sweep that changes nothing leaves every arc
// satisfied, so dist is then the exact distance row (at most nb_total sweeps with
// non-negative weights).  Thread t owns vertices t + 256 i (i < VPT) and keeps their
// records in registers for all sweeps.
template <int VPT>
__global__ __launch_bounds__(kClThreads) void bg_closure_kernel(
    const BgVert *__restrict__ verts, const int *__restrict__ wt, int nb_total, int Sb,
    const BgPart *__restrict__ parts, int np, const int *__restrict__ pidx_by_pid, int npid,
    const int *__restrict__ row_off, const int *__restrict__ col, const int *__restrict__ w,
    const int *__restrict__ graph_id, const int *__restrict__ st2ed,
    const int *__restrict__ ed2st, int *__restrict__ out) {
  extern __shared__ int dist[];
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  BgVert mine[VPT];
#pragma unroll
  for (int i = 0; i < VPT; ++i) {
    const int v = tid + i * kClThreads;
    if (v < nb_total) {
      mine[i] = verts[v];
    } else {
      mine[i].bs = mine[i].bdy = mine[i].woff = mine[i].ncut = 0;
#pragma unroll
      for (int c = 0; c < kBgCutW; ++c) {
        mine[i].cut_t[c] = 0;
        mine[i].cut_w[c] = kInf;  // never improves anything
      }
    }
  }
  for (int j = tid; j < Sb + kClPad; j += kClThreads) dist[j] = j == s ? 0 : kInf;
  __syncthreads();
  for (int sweep = 0; sweep <= nb_total; ++sweep) {
    bool changed = false;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = tid + i * kClThreads;
      if (v >= nb_total) continue;
      const int *wv = wt + mine[i].woff;
      const int *dv = dist + mine[i].bs;
      int best = dist[v];
      // Eight steps per iteration, loads first (wt and dist hold kClPad slack entries).
      for (int c0 = 0; c0 < mine[i].bdy; c0 += kClPad) {
        int x[kClPad], y[kClPad];
#pragma unroll
        for (int u = 0; u < kClPad; ++u) {
          x[u] = wv[c0 + u];
          y[u] = dv[c0 + u];
        }
#pragma unroll
        for (int u = 0; u < kClPad; ++u)
          if (c0 + u < mine[i].bdy) best = min(best, x[u] + y[u]);
      }
      if (best < dist[v]) {
        dist[v] = best;
        changed = true;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int u = tid + i * kClThreads;
      if (u >= nb_total) continue;
      const int du = dist[u];
      if (du >= kInf) continue;
#pragma unroll
      for (int c = 0; c < kBgCutW; ++c) {
        const int t = mine[i].cut_t[c];
        const int nd = du + mine[i].cut_w[c];
        if (nd < dist[t] && nd < atomicMin(&dist[t], nd)) changed = true;
      }
      if (mine[i].ncut > kBgCutW) {  // more cut arcs than the record holds: read the CSR
        const BgPart p = parts[find_part_by_bs(parts, np, u)];
        const int vtx = st2ed[p.offset + (u - p.bs)];
        for (int e = row_off[vtx]; e < row_off[vtx + 1]; ++e) {
          const int t = cut_target(e, p.part_id, parts, pidx_by_pid, npid, col, graph_id, ed2st);
          if (t < 0) continue;
          const int nd = du + (w ? w[e] : 1);
          if (nd < dist[t] && nd < atomicMin(&dist[t], nd)) changed = true;
        }
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  int *dst = out + (int64_t)s * Sb;
  for (int j = tid * 4; j < Sb; j += kClThreads * 4)
    *reinterpret_cast<v4i *>(dst + j) = *reinterpret_cast<const v4i *>(&dist[j]);
}

// The same closure for unit-weight graphs with every G_B vertex's clique weights in
// LDS (16 bits, 0xFFFF = unreachable inside the part; in-part distances stay below
// s_max < 0xFFFF), NS sources per workgroup sharing them, and work lists: a part is
// pulled only when one of its boundary distances fell since its last pull (a byte
// flag per part at the part's first G_B index, double-buffered: the push of a sweep
// sets the next sweep's flags while this sweep's are cleared), and a vertex pushes
// its cut arcs (weight 1) only when its distance fell since its last push.  The final
// sweep changes nothing, so every constraint holds: a part not pulled kept its
// boundary values since a pull that satisfied its clique constraints, and a vertex
// not pushed kept its value since a push that satisfied its arcs.  On MI355X, for
// 1,297 G_B vertices (15-19 sweeps per source), re-pulling every part in every sweep
// took 251 us.  LDS: the weights (wlen u16), NS distance rows of stride Sb + kClPad,
// then 2 x NS flag rows of the same stride in bytes and the part start of every G_B
// vertex (u16, Sb entries).
constexpr int kClwThreads = 512;
template <int VPT, int NS>
__global__ __launch_bounds__(kClwThreads) void bg_closure_lds_kernel(
    const BgVert *__restrict__ verts, const int *__restrict__ wt, int wlen, int nb_total, int Sb,
    const BgPart *__restrict__ parts, int np, const int *__restrict__ pidx_by_pid, int npid,
    const int *__restrict__ row_off, const int *__restrict__ col,
    const int *__restrict__ graph_id, const int *__restrict__ st2ed,
    const int *__restrict__ ed2st, int cb_total, int *__restrict__ out) {
  extern __shared__ int sh[];
  unsigned short *w16 = reinterpret_cast<unsigned short *>(sh);
  const int ld = Sb + kClPad;
  int *dist0 = sh + wlen / 2;
  unsigned char *flags = reinterpret_cast<unsigned char *>(dist0 + NS * ld);  // [2][NS][ld]
  unsigned short *pbs = reinterpret_cast<unsigned short *>(flags + 2 * NS * ld);
  const int s0 = blockIdx.x * NS;
  const int tid = threadIdx.x;
  int bs[VPT], bdy[VPT], woff[VPT], ncut[VPT], cut_t[VPT][kBgCutW];
  int pushed[VPT][NS];  // distance at the vertex's last push (FAP_INF: never pushed)
#pragma unroll
  for (int i = 0; i < VPT; ++i) {
    const int v = tid + i * kClwThreads;
    bs[i] = bdy[i] = woff[i] = ncut[i] = 0;
#pragma unroll
    for (int c = 0; c < kBgCutW; ++c) cut_t[i][c] = v;  // a self arc never improves
    if (v < nb_total) {
      const BgVert bv = verts[v];
      bs[i] = bv.bs;
      bdy[i] = bv.bdy;
      woff[i] = bv.woff;
      ncut[i] = bv.ncut;
#pragma unroll
      for (int c = 0; c < kBgCutW; ++c) cut_t[i][c] = bv.cut_t[c];
      pbs[v] = (unsigned short)bv.bs;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) pushed[i][s] = kInf;
  }
  for (int j = tid; j < wlen; j += kClwThreads)
    w16[j] = j < cb_total ? (unsigned short)min(wt[j], 0xFFFF) : (unsigned short)0xFFFF;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    for (int j = tid; j < ld; j += kClwThreads) dist0[s * ld + j] = j == s0 + s ? 0 : kInf;
  for (int j = tid; j < 2 * NS * ld; j += kClwThreads) flags[j] = 0;
  __syncthreads();
  // The source's part is pulled in the first sweep.
  if (tid < NS && s0 + tid < nb_total) flags[tid * ld + pbs[s0 + tid]] = 1;
  __syncthreads();
  for (int sweep = 0; sweep <= nb_total; ++sweep) {
    unsigned char *cur = flags + (sweep & 1) * NS * ld;
    unsigned char *nxt = flags + ((sweep & 1) ^ 1) * NS * ld;
    bool changed = false;
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int v = tid + i * kClwThreads;
      if (v >= nb_total) continue;
      bool pull[NS], anyp = false;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        pull[s] = cur[s * ld + bs[i]] != 0;
        anyp |= pull[s];
      }
      if (!anyp) continue;
      const unsigned short *wv = w16 + woff[i];
      int best[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) best[s] = dist0[s * ld + v];
      for (int c0 = 0; c0 < bdy[i]; c0 += kClPad) {
        int x[kClPad];
#pragma unroll
        for (int u = 0; u < kClPad; ++u) {
          const int t = wv[c0 + u];
          x[u] = (t == 0xFFFF || c0 + u >= bdy[i]) ? kInf : t;
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (!pull[s]) continue;
          const int *dv = dist0 + s * ld + bs[i] + c0;
          int y[kClPad];
#pragma unroll
          for (int u = 0; u < kClPad; ++u) y[u] = dv[u];
#pragma unroll
          for (int u = 0; u < kClPad; u += 2)
            best[s] = min3_i32(best[s], x[u] + y[u], x[u + 1] + y[u + 1]);
        }
      }
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (pull[s] && best[s] < dist0[s * ld + v]) {
          dist0[s * ld + v] = best[s];
          changed = true;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < VPT; ++i) {
      const int u = tid + i * kClwThreads;
      if (u >= nb_total) continue;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        cur[s * ld + bs[i]] = 0;  // this sweep's flags; the push sets nxt
        int *dist = dist0 + s * ld;
        const int du = dist[u];
        if (du >= pushed[i][s]) continue;  // unchanged since its last push (or FAP_INF)
        pushed[i][s] = du;
        const int nd = du + 1;
#pragma unroll
        for (int c = 0; c < kBgCutW; ++c) {
          const int t = cut_t[i][c];
          if (nd < dist[t] && nd < atomicMin(&dist[t], nd)) {
            changed = true;
            nxt[s * ld + pbs[t]] = 1;
          }
        }
        if (ncut[i] > kBgCutW) {  // more cut arcs than the record holds: read the CSR
          const BgPart p = parts[find_part_by_bs(parts, np, u)];
          const int vtx = st2ed[p.offset + (u - p.bs)];
          for (int e = row_off[vtx]; e < row_off[vtx + 1]; ++e) {
            const int t = cut_target(e, p.part_id, parts, pidx_by_pid, npid, col, graph_id, ed2st);
            if (t < 0) continue;
            if (nd < dist[t] && nd < atomicMin(&dist[t], nd)) {
              changed = true;
              nxt[s * ld + pbs[t]] = 1;
            }
          }
        }
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (s0 + s >= nb_total) break;
    int *dst = out + (int64_t)(s0 + s) * Sb;
    const int *src = dist0 + s * ld;
    for (int j = tid * 4; j < Sb; j += kClwThreads * 4)
      *reinterpret_cast<v4i *>(dst + j) = *reinterpret_cast<const v4i *>(src + j);
  }
}

constexpr int kRowsCols = 1024;  // output columns per workgroup (256 threads x 4)
constexpr int kRowsMax = 16;     // G_B rows staged in LDS at a time

// Stages rows g .. g + rows - 1 of the closed G_B matrix into LDS (stride ldl, INF
// tail past Sb) and their output row offsets (orow[g] * ld_out) into ooff.  Sb and
// ldl are multiples of 4; 16-byte loads, eight in flight per thread before the LDS
// writes.
__device__ __forceinline__ void stage_rows(int *lds, int64_t *ooff, const int *__restrict__ mat,
                                           const int64_t *__restrict__ orow, int64_t ld_out,
                                           int Sb, int ldl, int g, int rows) {
  if ((int)threadIdx.x < rows) ooff[threadIdx.x] = orow[g + threadIdx.x] * ld_out;
  const int q4 = ldl / 4, total = rows * q4;
  const v4i inf4 = {kInf, kInf, kInf, kInf};
  for (int base = threadIdx.x; base < total; base += 8 * 256) {
    v4i t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = base + u * 256;
      const int r = idx / q4, c = (idx - r * q4) * 4;
      t[u] = (idx < total && c < Sb)
                 ? *reinterpret_cast<const v4i *>(mat + (int64_t)(g + r) * Sb + c)
                 : inf4;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = base + u * 256;
      if (idx < total) *reinterpret_cast<v4i *>(lds + idx * 4) = t[u];
    }
  }
}

struct ColInfo {  // one output column: its part's G_B offset and in-part column
  int bs, bdy, S, l;
  const int *fw;
};

// A column past n gets part 0's matrix with no boundary (every fin entry FAP_INF),
// so its loads stay in bounds.
__device__ __forceinline__ ColInfo column_info(const BgPart *__restrict__ parts,
                                               const int *__restrict__ pidx_by_pid, int npid,
                                               const int *__restrict__ graph_id,
                                               const int *__restrict__ ed2st, int j, int n) {
  // A label outside the table, or a sorted position outside its part's padded matrix
  // (inconsistent layouts from a C-ABI caller), reads like a column past n: in bounds.
  const int pid = j < n ? graph_id[j] : -1;
  if (pid >= 0 && pid < npid) {
    const BgPart q = parts[pidx_by_pid[pid]];
    const int l = ed2st[j] - q.offset;
    if (l >= 0 && l < q.S) return ColInfo{q.bs, q.bdy, q.S, l, q.fw};
  }
  return ColInfo{0, 0, parts[0].S, 0, parts[0].fw};
}

// fin[c] = d_q(c, l) for c < bdy, FAP_INF beyond: unconditional loads (rows c < KP <=
// 64 <= S of the part's padded matrix are always readable), then the mask.
template <int KP>
__device__ __forceinline__ void load_fin(int (&fin)[KP], const ColInfo &ci, int c0) {
#pragma unroll
  for (int c = 0; c < KP; ++c) fin[c] = gptr(ci.fw)[(int64_t)(c0 + c) * ci.S + ci.l];
#pragma unroll
  for (int c = 0; c < KP; ++c)
    if (c0 + c >= ci.bdy) fin[c] = kInf;
}

// The same from a symmetric in-part matrix: d_q(c, l) = d_q(l, c), so row l gives the
// KP values contiguously (16-byte loads; KP <= 32 <= S columns are readable).  The
// matrix pointer comes from the part table (a generic pointer): gptr makes the loads
// global ones, which do not hold up the LDS counter as flat loads would.
template <int KP>
__device__ __forceinline__ void load_fin_sym(int (&fin)[KP], const ColInfo &ci) {
  const FAP_GLOBAL v4i *row = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(ci.fw + (int64_t)ci.l * ci.S));
#pragma unroll
  for (int c = 0; c < KP; c += 4) {
    const v4i v = row[c / 4];
    fin[c] = v.x;
    fin[c + 1] = v.y;
    fin[c + 2] = v.z;
    fin[c + 3] = v.w;
  }
#pragma unroll
  for (int c = 0; c < KP; ++c)
    if (c >= ci.bdy) fin[c] = kInf;
}

// Chunk c0 .. c0 + KC - 1 of the same from a symmetric matrix: row l, columns c0 + c
// (16-byte loads; c0 is a multiple of KC, a multiple of 4, and the chunk ends within
// the first ceil(bdy / 64) * 64 <= S columns, which the in-part BFS writes before the
// boundary stage).
template <int KC>
__device__ __forceinline__ void load_fin_sym_chunk(int (&fin)[KC], const ColInfo &ci, int c0) {
  const FAP_GLOBAL v4i *row =
      reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(ci.fw + (int64_t)ci.l * ci.S + c0));
#pragma unroll
  for (int c = 0; c < KC; c += 4) {
    const v4i v = row[c / 4];
    fin[c] = v.x;
    fin[c + 1] = v.y;
    fin[c + 2] = v.z;
    fin[c + 3] = v.w;
  }
#pragma unroll
  for (int c = 0; c < KC; ++c)
    if (c0 + c >= ci.bdy) fin[c] = kInf;
}

// *max_b = max(*max_b, the wave's largest value of mx): one atomic per wave.
__device__ __forceinline__ void publish_max(int *max_b, int mx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
  if ((threadIdx.x & 63) == 0) atomicMax(max_b, mx);
}

// Rows r < rows of the staged chunk for one column j of each lane, W boundary steps:
// W >= the largest boundary of the wave's columns (uniform), so the entries past a
// lane's own boundary meet fin = FAP_INF and never win.
template <int W, int KP>
__device__ __forceinline__ void bg_rows_col(const int *lds, int ldl, const int64_t *ooff,
                                            int rows, int bs, const int (&fin)[KP],
                                            int *__restrict__ out, int j, int n, int &mx) {
  for (int r = 0; r < rows; ++r) {
    const int *a = lds + r * ldl + bs;
    int acc = kInf;
#pragma unroll
    for (int c = 0; c < W; c += 2) acc = min3_i32(acc, a[c] + fin[c], a[c + 1] + fin[c + 1]);
    if (j < n) {
      __builtin_nontemporal_store(acc, out + ooff[r] + j);
      if (acc < kInf) mx = max(mx, acc);
    }
  }
}

// The smallest W in {4, 8, ..., KP} with W >= wb (wb uniform, so is the branch).
template <int KP, int W = 4>
__device__ __forceinline__ void bg_rows_pick(int wb, const int *lds, int ldl,
                                             const int64_t *ooff, int rows, int bs,
                                             const int (&fin)[KP], int *__restrict__ out, int j,
                                             int n, int &mx) {
  if constexpr (W >= KP) {
    bg_rows_col<KP, KP>(lds, ldl, ooff, rows, bs, fin, out, j, n, mx);
  } else {
    if (wb <= W) bg_rows_col<W, KP>(lds, ldl, ooff, rows, bs, fin, out, j, n, mx);
    else bg_rows_pick<KP, W + 4>(wb, lds, ldl, ooff, rows, bs, fin, out, j, n, mx);
  }
}

// Every part has at most KP boundary vertices.  Thread t handles columns
// blockIdx.x * 1024 + e * 256 + t (e < 4): the lanes of a wave cover 64 consecutive
// vertices, so they share a few parts (LDS broadcast reads) and every store is one
// coalesced 256-byte run.  Each lane loads d_q(c, l) of its columns once (registers;
// from row l when the in-part matrices are symmetric), then walks its workgroup's
// share of G_B rows (blockIdx.y of gridDim.y splits) through LDS, R rows at a time.
// A wave's column e takes as many boundary steps as the largest boundary among its
// 64 columns, rounded up to 4 (on the 114,599-vertex road-like graph, K = 128:
// 17 steps per column on average against 24 for the largest part).  LDS stride
// ldl = Sb + KP (INF tail).
template <int KP, bool kSym>
__global__ __launch_bounds__(256) void bg_rows_kernel(
    const BgPart *__restrict__ parts, const int64_t *__restrict__ orow, int nb_total,
    const int *__restrict__ mat, int Sb, int R, const int *__restrict__ pidx_by_pid, int npid,
    const int *__restrict__ graph_id, const int *__restrict__ ed2st, int n,
    int *__restrict__ out, int64_t ld_out, int *__restrict__ max_b) {
  extern __shared__ int lds[];
  __shared__ int64_t ooff[kRowsMax];
  const int ldl = Sb + KP;
  const int per = (nb_total + gridDim.y - 1) / gridDim.y;
  const int gbeg = blockIdx.y * per, gend = min(nb_total, gbeg + per);
  const int jb = blockIdx.x * kRowsCols + threadIdx.x;
  int bs[4], wb[4];
  int fin[4][KP];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const ColInfo ci = column_info(parts, pidx_by_pid, npid, graph_id, ed2st, jb + e * 256, n);
    bs[e] = ci.bs;
    if (kSym) load_fin_sym<KP>(fin[e], ci);
    else load_fin<KP>(fin[e], ci, 0);
    int m = ci.bdy;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    wb[e] = __builtin_amdgcn_readfirstlane(m);
  }
  int mx = 0;  // largest finite distance written by this thread
  for (int g0 = gbeg; g0 < gend; g0 += R) {
    const int rows = min(R, gend - g0);
    __syncthreads();  // the previous chunk's reads are done
    stage_rows(lds, ooff, mat, orow, ld_out, Sb, ldl, g0, rows);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e)
      bg_rows_pick<KP>(wb[e], lds, ldl, ooff, rows, bs[e], fin[e], out, jb + e * 256, n, mx);
  }
  publish_max(max_b, mx);
}

// Parts with more than 32 boundary vertices: the boundary is walked in chunks of KC
// vertices; per chunk each lane loads d_q(c, l) of its four columns (FAP_INF past
// b_q) and folds the chunk into all staged rows' accumulators.  A chunk past the
// lane's own boundary reads the INF tail (base clamped to Sb), so every LDS read
// stays inside its row whatever the boundary sizes.  Columns as in bg_rows_kernel
// (kSym: d_q(c, l) from row l, which the boundary source groups of the in-part BFS
// write; column l of rows c may come from the other groups, still running beside the
// boundary stage); LDS stride ldl = Sb + KC.
template <int KC, bool kSym>
__global__ __launch_bounds__(256) void bg_rows_wide_kernel(
    const BgPart *__restrict__ parts, const int64_t *__restrict__ orow, int nb_total,
    const int *__restrict__ mat, int Sb, int R, int nchunks, const int *__restrict__ pidx_by_pid, int npid,
    const int *__restrict__ graph_id, const int *__restrict__ ed2st, int n,
    int *__restrict__ out, int64_t ld_out, int *__restrict__ max_b) {
  extern __shared__ int lds[];
  __shared__ int64_t ooff[kRowsMax];
  const int ldl = Sb + KC;
  const int per = (nb_total + gridDim.y - 1) / gridDim.y;
  const int gbeg = blockIdx.y * per, gend = min(nb_total, gbeg + per);
  const int jb = blockIdx.x * kRowsCols + threadIdx.x;
  ColInfo ci[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    ci[e] = column_info(parts, pidx_by_pid, npid, graph_id, ed2st, jb + e * 256, n);
  int mx = 0;  // largest finite distance written by this thread
  for (int g0 = gbeg; g0 < gend; g0 += R) {
    const int rows = min(R, gend - g0);
    __syncthreads();
    stage_rows(lds, ooff, mat, orow, ld_out, Sb, ldl, g0, rows);
    __syncthreads();
    int acc[kRowsMax][4];
#pragma unroll
    for (int r = 0; r < kRowsMax; ++r)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[r][e] = kInf;
    for (int ch = 0; ch < nchunks; ++ch) {
      const int c0 = ch * KC;
      int fin[4][KC];
      int base[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        base[e] = min(ci[e].bs + c0, Sb);
        // Chunks past this column's boundary need no loads (rows c0 + c may pass S).
        if (c0 < ci[e].bdy) {
          if (kSym) load_fin_sym_chunk<KC>(fin[e], ci[e], c0);
          else load_fin<KC>(fin[e], ci[e], c0);
        } else {
#pragma unroll
          for (int c = 0; c < KC; ++c) fin[e][c] = kInf;
        }
      }
#pragma unroll
      for (int r = 0; r < kRowsMax; ++r) {
        if (r >= rows) continue;  // (continue, not break: keeps the loop unrolled)
        const int *lr = lds + r * ldl;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int c = 0; c < KC; c += 2)
            acc[r][e] = min3_i32(acc[r][e], lr[base[e] + c] + fin[e][c],
                                 lr[base[e] + c + 1] + fin[e][c + 1]);
      }
    }
#pragma unroll
    for (int r = 0; r < kRowsMax; ++r) {
      if (r >= rows) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = jb + e * 256;
        if (j < n) {
          __builtin_nontemporal_store(acc[r][e], out + ooff[r] + j);
          if (acc[r][e] < kInf) mx = max(mx, acc[r][e]);
        }
      }
    }
  }
  publish_max(max_b, mx);
}

}  // namespace

int launch_bg_vert(const BgPart *d_parts, int np, int nb_total, const int *d_pidx_by_pid, int npid,
                   const int *row_off, const int *col, const int *w, const int *graph_id,
                   const int *st2ed, const int *ed2st, BgVert *d_verts, int *d_wt,
                   int64_t *d_orow, hipStream_t stream) {
  if (nb_total <= 0) return FAP_OK;
  hipLaunchKernelGGL(bg_vert_kernel, dim3((nb_total + 255) / 256), dim3(256), 0, stream, d_parts,
                     np, nb_total, d_pidx_by_pid, npid, row_off, col, w, graph_id, st2ed, ed2st,
                     d_verts, d_wt, d_orow);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_bg_closure(const BgVert *d_verts, const int *d_wt, int cb_total, bool unit,
                      int nb_total, int Sb, const BgPart *d_parts, int np,
                      const int *d_pidx_by_pid, int npid, const int *row_off, const int *col, const int *w,
                      const int *graph_id, const int *st2ed, const int *ed2st, int *d_mat,
                      hipStream_t stream) {
  if (nb_total <= 0) return FAP_OK;
  const int vpt = (nb_total + kClThreads - 1) / kClThreads;
  // LDS form (unit weights, cut arcs of weight 1): weights as u16 (slack of kClPad, a
  // multiple of 8 so the rows after them are 16-byte aligned).  NS sources per
  // workgroup: the fewest that fit every workgroup on the device at once (two per CU:
  // the registers of the 512-thread kernel with VPT = 3 allow two), within 64 KB.
  const int wlen = (cb_total + kClPad + 7) / 8 * 8;
  const int ld = Sb + kClPad;
  const int vptw = (nb_total + kClwThreads - 1) / kClwThreads;
  auto lds_of = [&](int ns) {
    return ((size_t)wlen / 2 + (size_t)ns * ld) * 4 + (size_t)2 * ns * ld + (size_t)Sb * 2;
  };
  int ns = 1;
  const int slots = 2 * std::max(1, device_cus());
  while (ns < 4 && (nb_total + ns - 1) / ns > slots && lds_of(ns + 1) <= 65536) ++ns;
  if (unit && w == nullptr && lds_of(ns) <= 65536 && vptw <= 4) {
    const int grid = (nb_total + ns - 1) / ns;
#define FAP_CLW_CASE(V, NS)                                                                   \
    if (vptw <= V && ns == NS) {                                                              \
      hipLaunchKernelGGL((bg_closure_lds_kernel<V, NS>), dim3(grid), dim3(kClwThreads),        \
                         lds_of(NS), stream, d_verts, d_wt, wlen, nb_total, Sb, d_parts, np,  \
                         d_pidx_by_pid, npid, row_off, col, graph_id, st2ed, ed2st, cb_total,       \
                         d_mat);                                                              \
      FAP_CHECK_LAUNCH();                                                                     \
      return FAP_OK;                                                                          \
    }
#define FAP_CLW_NS(V) FAP_CLW_CASE(V, 1) FAP_CLW_CASE(V, 2) FAP_CLW_CASE(V, 3) FAP_CLW_CASE(V, 4)
    FAP_CLW_NS(1)
    FAP_CLW_NS(2)
    FAP_CLW_NS(3)
    FAP_CLW_NS(4)
#undef FAP_CLW_NS
#undef FAP_CLW_CASE
  }
#define FAP_CL_CASE(V)                                                                        \
  if (vpt <= V) {                                                                             \
    hipLaunchKernelGGL(bg_closure_kernel<V>, dim3(nb_total), dim3(kClThreads),                 \
                       (size_t)(Sb + kClPad) * 4,                                              \
                       stream, d_verts, d_wt, nb_total, Sb, d_parts, np, d_pidx_by_pid, npid,  \
                       row_off, col, w, graph_id, st2ed, ed2st, d_mat);                        \
    FAP_CHECK_LAUNCH();                                                                       \
    return FAP_OK;                                                                            \
  }
  FAP_CL_CASE(2)
  FAP_CL_CASE(4)
  FAP_CL_CASE(8)
  FAP_CL_CASE(16)
#undef FAP_CL_CASE
  set_error("launch_bg_closure: %d boundary vertices exceed the closure kernel", nb_total);
  return FAP_ERR_INVALID;
}

// Rows of G_B staged at a time: they fit in 52 KB of LDS, so three workgroups can
// share a CU (the registers of bg_rows_kernel with KP <= 24 allow three, the wider
// forms two).
int bg_rows_chunk(int ldl) { return std::max(1, std::min(kRowsMax, 53248 / (ldl * 4))); }

int launch_bg_rows(const BgPart *d_parts, const int64_t *d_orow, int nb_total, const int *d_mat,
                   int Sb, int max_bdy, const int *d_pidx_by_pid, int npid, const int *graph_id,
                   const int *ed2st, int n, int *d_out, int64_t ld_out, int *d_max_b,
                   bool sym, hipStream_t stream) {
  if (nb_total <= 0 || n <= 0) return FAP_OK;
  const int cols = (n + kRowsCols - 1) / kRowsCols;
  // Row splits: as many workgroups as fit on the device at once (a second round of a
  // few workgroups would double the time), each split at least one chunk.
  const bool narrow = max_bdy <= 32;
  const int KP = narrow ? std::max(8, (max_bdy + 7) / 8 * 8) : 0;
  const int KC = narrow ? 0 : 16;
  const int ldl = Sb + (narrow ? KP : KC);
  const int R = bg_rows_chunk(ldl);
  const int slots = std::max(1, device_cus()) * (narrow && KP <= 24 ? 3 : 2);
  const int splits = std::max(1, std::min((nb_total + R - 1) / R, slots / cols));
  const dim3 grid(cols, splits);
  const size_t lds = (size_t)R * ldl * 4;
#define FAP_BG_CASE(KPV)                                                                      \
  case KPV:                                                                                   \
    if (sym)                                                                                  \
      hipLaunchKernelGGL((bg_rows_kernel<KPV, true>), grid, dim3(256), lds, stream, d_parts,   \
                         d_orow, nb_total, d_mat, Sb, R, d_pidx_by_pid, npid, graph_id, ed2st, n,    \
                         d_out, ld_out, d_max_b);                                              \
    else                                                                                      \
      hipLaunchKernelGGL((bg_rows_kernel<KPV, false>), grid, dim3(256), lds, stream, d_parts,  \
                         d_orow, nb_total, d_mat, Sb, R, d_pidx_by_pid, npid, graph_id, ed2st, n,    \
                         d_out, ld_out, d_max_b);                                              \
    break;
  if (narrow) {
    switch (KP) {
      FAP_BG_CASE(8)
      FAP_BG_CASE(16)
      FAP_BG_CASE(24)
      FAP_BG_CASE(32)
      default:
        set_error("launch_bg_rows: unsupported boundary width %d", KP);
        return FAP_ERR_INVALID;
    }
  } else {
    const int nchunks = (max_bdy + KC - 1) / KC;
    if (sym)
      hipLaunchKernelGGL((bg_rows_wide_kernel<16, true>), grid, dim3(256), lds, stream, d_parts,
                         d_orow, nb_total, d_mat, Sb, R, nchunks, d_pidx_by_pid, npid, graph_id, ed2st,
                         n, d_out, ld_out, d_max_b);
    else
      hipLaunchKernelGGL((bg_rows_wide_kernel<16, false>), grid, dim3(256), lds, stream, d_parts,
                         d_orow, nb_total, d_mat, Sb, R, nchunks, d_pidx_by_pid, npid, graph_id, ed2st,
                         n, d_out, ld_out, d_max_b);
  }
#undef FAP_BG_CASE
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_inpart_bfs(const InpartDesc *d_descs, int nd, int total_groups, int s_max,
                      hipStream_t stream);

namespace {
Workspace g_ub_meta;   // part tables, G_B vertices, output rows, clique weights, maximum
Workspace g_ub_mat;    // closed G_B matrix
}  // namespace

// Boundary rows of parts that cover a unit-weight graph with paired arcs, for callers
// other than the distance pipeline (whose boundary stage interleaves with its other
// launches): in-part distances from every part's boundary sources by the bit-parallel
// BFS into fw[p] (S[p] x S[p], the source groups that hold the boundary, so columns
// 0 .. 64 ceil(b_p / 64) - 1 of every row), the boundary graph, its closure, and the
// b_p rows of every part into d_out at rows out_row ..  Returns FAP_ERR_INVALID with
// nothing launched when the parts do not qualify (weights, no component labels, parts
// not covering the graph, more than 4096 boundary vertices or a part above the BFS
// size limit).  Every entry is an exact integer min/plus expression, so the rows equal
// one SSSP per boundary vertex.
int boundary_rows_unit(const fap_graph_dev *g, const fap_part_desc *act, int P, int *const *fw,
                       const int *S, int32_t *d_out, int64_t ld_out, hipStream_t st) {
  if (g->d_w || !g->d_comp || P <= 0) return FAP_ERR_INVALID;
  std::vector<std::pair<int, int>> span_of;
  int nb_total = 0, max_bdy = 0, max_pid = 0, s_max = 0;
  for (int p = 0; p < P; ++p) {
    span_of.push_back({act[p].offset, act[p].num});
    nb_total += act[p].bdy;
    max_bdy = std::max(max_bdy, act[p].bdy);
    max_pid = std::max(max_pid, act[p].part_id);
    s_max = std::max(s_max, act[p].num);
    if (S[p] % 64 != 0 || S[p] < act[p].num) return FAP_ERR_INVALID;
  }
  std::sort(span_of.begin(), span_of.end());
  int next = 0;
  for (const auto &x : span_of) {
    if (x.first != next) return FAP_ERR_INVALID;
    next = x.first + x.second;
  }
  if (next != g->n || nb_total == 0 || nb_total > 4096 || s_max > kInpartBfsMaxS)
    return FAP_ERR_INVALID;
  std::vector<InpartDesc> ipa;
  std::vector<BgPart> bgp(P);
  std::vector<int> pidx(max_pid + 1, 0);
  int groups = 0, bs = 0, cb = 0;
  for (int p = 0; p < P; ++p) {
    const fap_part_desc &q = act[p];
    const int ga = std::min(S[p] / 64, (q.bdy + 63) / 64);
    if (ga > 0)
      ipa.push_back(InpartDesc{fw[p], S[p], q.num, q.offset, q.part_id, groups, 0,
                               g->d_row_offset, g->d_col, g->d_graph_id, g->d_st2ed, g->d_ed2st});
    groups += ga;
    bgp[p] = BgPart{fw[p], S[p], q.bdy, q.offset, q.part_id, bs, cb, q.out_row};
    pidx[q.part_id] = p;
    bs += q.bdy;
    cb += q.bdy * q.bdy;
  }
  const int Sb = (nb_total + 63) / 64 * 64;
  auto al = [](int64_t b) { return (b + 255) / 256 * 256; };
  const int64_t o_ip = 0, o_bgp = o_ip + al(sizeof(InpartDesc) * ipa.size());
  const int64_t o_pidx = o_bgp + al(sizeof(BgPart) * P);
  const int64_t upload = o_pidx + al(4LL * (max_pid + 1));
  const int64_t o_verts = upload, o_orow = o_verts + al(sizeof(BgVert) * nb_total);
  const int64_t o_wt = o_orow + al(8LL * nb_total), o_max = o_wt + al(4LL * (cb + 64));
  FAP_TRY(g_ub_meta.ensure(o_max + 256));
  FAP_TRY(g_ub_mat.ensure((int64_t)Sb * Sb * 4));
  char *meta = static_cast<char *>(g_ub_meta.ptr);
  std::vector<char> host(upload);
  std::memcpy(host.data() + o_ip, ipa.data(), sizeof(InpartDesc) * ipa.size());
  std::memcpy(host.data() + o_bgp, bgp.data(), sizeof(BgPart) * P);
  std::memcpy(host.data() + o_pidx, pidx.data(), 4LL * (max_pid + 1));
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host.data(), upload, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  const BgPart *d_bgp = reinterpret_cast<const BgPart *>(meta + o_bgp);
  const int *d_pidx = reinterpret_cast<const int *>(meta + o_pidx);
  BgVert *d_verts = reinterpret_cast<BgVert *>(meta + o_verts);
  int64_t *d_orow = reinterpret_cast<int64_t *>(meta + o_orow);
  int *d_wt = reinterpret_cast<int *>(meta + o_wt);
  int *d_max = reinterpret_cast<int *>(meta + o_max);
  int *d_mat = static_cast<int *>(g_ub_mat.ptr);
  FAP_TRY(launch_inpart_bfs(reinterpret_cast<const InpartDesc *>(meta + o_ip), (int)ipa.size(),
                            groups, s_max, st));
  FAP_TRY(fap_hip_fill_i32(d_max, 0, 1, st));
  FAP_TRY(launch_bg_vert(d_bgp, P, nb_total, d_pidx, max_pid + 1, g->d_row_offset, g->d_col, nullptr,
                         g->d_graph_id, g->d_st2ed, g->d_ed2st, d_verts, d_wt, d_orow, st));
  FAP_TRY(launch_bg_closure(d_verts, d_wt, cb, s_max < 0xFFFF, nb_total, Sb, d_bgp, P, d_pidx, max_pid + 1,
                            g->d_row_offset, g->d_col, nullptr, g->d_graph_id, g->d_st2ed,
                            g->d_ed2st, d_mat, st));
  return launch_bg_rows(d_bgp, d_orow, nb_total, d_mat, Sb, max_bdy, d_pidx, max_pid + 1, g->d_graph_id,
                        g->d_ed2st, g->n, d_out, ld_out, d_max, true, st);
}

}  // namespace fap_hip
