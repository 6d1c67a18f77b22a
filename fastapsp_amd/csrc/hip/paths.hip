// Predecessor ("path") matrices beside the distances: fap_hip_solve_parts_path.
//
// The reference fills path[s][v] = the vertex before v on the shortest s -> v path it
// picked, -1 when unreachable, s on the diagonal (fapGraph::run with is_path_needed,
// fap/fap.cpp:195-503).  Which path is picked depends on the order of its loops, so
// every stage here reproduces that order exactly:
//
//  * boundary rows (dijkstra_path, fap/kernel/batch_sssp.h:79-114): a min-heap pops
//    (dist, id) pairs and a strictly shorter distance sets path[v] = the popped
//    vertex, so path[v] is the first-settled in-neighbour u with
//    d(u) + w(u, v) == d(v): the smallest (d(u), u) among them, counting only u
//    settled before v ((d(u), u) < (d(v), v), which zero-weight arcs make a real
//    condition).  pred_rows_kernel evaluates that rule over the reverse CSR once the
//    distance row is known;
//  * sub-matrix build (MysubMatBuild_path, fap/utils/util_Myalgorithm.h:44-84): the
//    part's arcs with path = the arc's tail (a later arc overwrites), the boundary
//    rows' distances and paths gathered from the rows above, the diagonal 0 / itself;
//  * Floyd-Warshall (floyd_CPU_path, fap/kernel/floyd.h:38-49): k-i-j with a strict
//    '>' and path[i][j] = path[k][j].  A blocked form changes which k wins a tie,
//    because a panel cell can improve through a later pivot of the same block before
//    it is used.  Here every phase uses the values the k-i-j loop would see: phase 1
//    records, for each pivot k of the block, row k (distances and paths) and column k
//    of the diagonal block as they stand at step k; phase 2 updates the panels from
//    those records and records each panel's column k (column panel) and row k with
//    paths (row panel) at step k; phase 3 updates every other tile from the panel
//    records.  Row k and column k do not change during step k, so each record is the
//    value the sequential loop reads, and the result equals it bit for bit (dist and
//    path);
//  * min-plus of the inner rows (min_plus_path_CPU_advanced, fap/kernel/minplus.h:
//    51-64): init MAXVALUE / -1, strict '>' over k ascending, so path = Bpath[k][j]
//    for the smallest k reaching the minimum (-1 when no k gives a finite sum);
//  * decode (MysubMatDecode_path, :236-247): every row of the part takes the
//    Floyd-Warshall distance and path in its own columns.
// Integer distances only (FAP_INF = unreachable).
#include <algorithm>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"

namespace fap_hip {

int launch_sssp(const int *d_row_off, const int *d_col, const int *d_w, int n, const int *d_src,
                int nsrc, const int64_t *d_rows, int *d_out, int64_t ld, bool vec_fill,
                hipStream_t stream);
int launch_ell_build(const int *d_row_off, const int *d_col, int n, int4 *d_ell,
                     hipStream_t stream);
int launch_bfs_lds(const int4 *d_ell, const int *d_row_off, const int *d_col, int n,
                   const int *d_src, int nsrc, const int64_t *d_rows, int *d_out, int64_t ld,
                   bool vec_fill, hipStream_t stream);
int launch_gather_sources(const int *offsets, const int *src_begin, const int64_t *out_rows,
                          int nparts, int total, const int *st2ed, int64_t ld, int *src,
                          int64_t *rows, hipStream_t stream);

namespace {

constexpr int kT = 64;          // Floyd-Warshall tile edge
constexpr int kTT = kT * kT;    // entries per tile
constexpr int kTP = kT + 1;     // LDS row stride of a tile: a column of a wave's rows
                                // then spreads over all banks
constexpr int kBuildChunkP = 8192;  // LDS row buffers of build_path_kernel: 64 KB

typedef int v4i __attribute__((ext_vector_type(4)));

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// One part of the path pipeline.
struct PathPart {
  int *fw, *fwp;      // S x S distances and paths (ld = S)
  int *snap;          // per-round records: diag (3 tiles), column panel (nb), row panel (2 nb)
  int S, nb, s, bdy, offset, part_id;
  int row_begin;      // first global build row
  int64_t out_row;    // first output row of the part
};

__device__ __forceinline__ int find_part(const PathPart *__restrict__ p, int np, int key,
                                         int PathPart::*field) {
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p[mid].*field <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Record areas of a part, each a 64 x 64 tile of ints.
__device__ __forceinline__ int *rec_diag_row(const PathPart &p) { return p.snap; }
__device__ __forceinline__ int *rec_diag_rowp(const PathPart &p) { return p.snap + kTT; }
__device__ __forceinline__ int *rec_diag_col(const PathPart &p) { return p.snap + 2 * kTT; }
__device__ __forceinline__ int *rec_colpanel(const PathPart &p, int t) {
  return p.snap + (int64_t)(3 + t) * kTT;
}
__device__ __forceinline__ int *rec_rowpanel(const PathPart &p, int t) {
  return p.snap + (int64_t)(3 + p.nb + t) * kTT;
}
__device__ __forceinline__ int *rec_rowpanelp(const PathPart &p, int t) {
  return p.snap + (int64_t)(3 + 2 * p.nb + t) * kTT;
}

// ---- boundary rows: predecessors from the distance rows ------------------------
// The largest finite boundary-row entry is kept in kMaxSlots words kSlotStride ints
// apart (each wave's maximum goes to one of them, so the atomics spread over
// channels); -1 when there is none.
constexpr int kMaxSlots = 64;
constexpr int kSlotStride = 16;

// Boundary rows k0 .. k0 + nk - 1 (source src[k], row offset rows[k] in d_out and
// d_path), kPredRows of them per thread: vertex v's in-arcs (up to kPredDeg of them)
// are read once into registers and checked against every row (one row per thread
// re-read the reverse CSR for each of the 1,297 rows of the 114,599-vertex graph).
constexpr int kPredRows = 8;
constexpr int kPredDeg = 8;
__global__ __launch_bounds__(256) void pred_rows_kernel(
    const int *__restrict__ rro, const int *__restrict__ rcol, const int *__restrict__ rw, int n,
    const int *__restrict__ src, const int64_t *__restrict__ rows, int64_t ld_rows, int k0,
    int nk, const int *__restrict__ out, int *__restrict__ path, int *__restrict__ bmax_slots) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const int kb = k0 + blockIdx.y * kPredRows, ke = min(k0 + nk, kb + kPredRows);
  int e0 = 0, deg = 0;
  int us[kPredDeg], ws[kPredDeg];
  if (v < n) {
    e0 = rro[v];
    deg = rro[v + 1] - e0;
  }
#pragma unroll
  for (int a = 0; a < kPredDeg; ++a) {
    us[a] = 0;
    ws[a] = 1;
    if (a < deg) {
      us[a] = rcol[e0 + a];
      if (rw) ws[a] = rw[e0 + a];
    }
  }
  for (int k = kb; k < ke; ++k) {
    const int64_t row = rows ? rows[k] : (int64_t)k * ld_rows;  // element offset of row k
    const int *d = out + row;
    const int dv = v < n ? d[v] : kInf;
    int mx = dv < kInf ? dv : -1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0 && mx >= 0) {
      const int slot = (blockIdx.x * 4 + (threadIdx.x >> 6) + k) % kMaxSlots;
      atomicMax(bmax_slots + slot * kSlotStride, mx);
    }
    if (v >= n) continue;
    const int s = src[k];
    int p = -1;
    if (v == s) {
      p = s;
    } else if (dv < kInf) {
      int bd = kInf, bu = 0x7FFFFFFF;
      auto consider = [&](int u, int wt) {
        const int du = d[u];
        if (du >= kInf || du + wt != dv) return;
        if (du > dv || (du == dv && u >= v)) return;  // settled after v
        if (du < bd || (du == bd && u < bu)) {
          bd = du;
          bu = u;
        }
      };
#pragma unroll
      for (int a = 0; a < kPredDeg; ++a)
        if (a < deg) consider(us[a], ws[a]);
      for (int e = e0 + kPredDeg; e < e0 + deg; ++e) consider(rcol[e], rw ? rw[e] : 1);
      p = bu == 0x7FFFFFFF ? -1 : bu;
    }
    path[row + v] = p;
  }
}

// ---- build ----------------------------------------------------------------------
// One workgroup per padded row of every part (rows of PathPart::row_begin prefix).
// pmax[part * kSlotStride] = max(itself, largest finite entry of the built rows).
__global__ __launch_bounds__(256) void build_path_kernel(
    const PathPart *__restrict__ parts, int np, int total_rows, const int *__restrict__ row_off,
    const int *__restrict__ col, const int *__restrict__ w, const int *__restrict__ graph_id,
    const int *__restrict__ st2ed, const int *__restrict__ ed2st, const int *__restrict__ out,
    const int *__restrict__ pout, int64_t ld_out, int *__restrict__ pmax) {
  const int grow = blockIdx.x;
  if (grow >= total_rows) return;
  const int pi = find_part(parts, np, grow, &PathPart::row_begin);
  const PathPart d = parts[pi];
  const int i = grow - d.row_begin;
  __shared__ int rd[kBuildChunkP];
  __shared__ int rp[kBuildChunkP];
  const int tid = threadIdx.x;
  const int ver = i < d.s ? st2ed[d.offset + i] : 0;
  int mx = -1;
  for (int c0 = 0; c0 < d.S; c0 += kBuildChunkP) {  // rows in chunks of LDS size
    const int cw = min(kBuildChunkP, d.S - c0);
    if (c0 > 0) __syncthreads();
    if (i < d.s && i < d.bdy) {
      const int64_t base = (d.out_row + i) * ld_out;
      for (int j = tid; j < cw; j += 256) {
        const bool in = c0 + j < d.s;
        const int v = in ? st2ed[d.offset + c0 + j] : 0;
        rd[j] = in ? out[base + v] : kInf;
        rp[j] = in ? pout[base + v] : -1;
      }
    } else {
      for (int j = tid; j < cw; j += 256) {
        rd[j] = kInf;
        rp[j] = -1;
      }
    }
    __syncthreads();
    if (tid == 0 && i < d.s) {
      if (i >= d.bdy) {  // boundary rows: the SSSP row replaces the arcs
        for (int e = row_off[ver]; e < row_off[ver + 1]; ++e) {
          const int nb = col[e];
          if (graph_id[nb] != d.part_id) continue;
          const int idx = ed2st[nb] - d.offset - c0;
          if (idx >= 0 && idx < cw && c0 + idx < d.s) {
            rd[idx] = w ? w[e] : 1;
            rp[idx] = ver;
          }
        }
      }
      if (i >= c0 && i < c0 + cw) {
        rd[i - c0] = 0;
        rp[i - c0] = ver;
      }
    }
    __syncthreads();
    FAP_GLOBAL int *dd = gptr(d.fw + (int64_t)i * d.S + c0);
    FAP_GLOBAL int *dp = gptr(d.fwp + (int64_t)i * d.S + c0);
    for (int j = tid; j < cw; j += 256) {
      dd[j] = rd[j];
      dp[j] = rp[j];
      if (rd[j] < kInf) mx = max(mx, rd[j]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
  __shared__ int red[4];
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  if (tid == 0) {
    mx = max(max(red[0], red[1]), max(red[2], red[3]));
    // Most rows hold nothing above what their part's word already has: read first,
    // so that only a new maximum takes the (serialised, same-line) atomic.
    int *slot = pmax + (int64_t)pi * kSlotStride;
    if (mx > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, mx);
  }
}

// ---- Floyd-Warshall with k-i-j path semantics -----------------------------------
// 256 threads per 64 x 64 tile: thread t owns row t / 4, columns (t % 4) * 16 + 0..15.
__device__ __forceinline__ int *tile(int *m, const PathPart &p, int ti, int tj) {
  return m + (int64_t)(ti * kT) * p.S + tj * kT;
}

// The part tables hold generic pointers: every global access of the Floyd-Warshall
// kernels goes through gptr (global_load / global_store).  A flat access also counts
// against the LDS counter, so each LDS write after a flat load, and each barrier after
// a flat store, waited for the memory round trip.  Row r, 16 columns c0.. as 16-byte
// loads (tiles start on 64-int boundaries of rows of S = 64 k ints).
__device__ __forceinline__ void tile_to_lds(int (*D)[kTP], int (*P)[kTP], const int *gd,
                                            const int *gp, int64_t ld) {
  const int r = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  const FAP_GLOBAL v4i *d4 = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(gd + (int64_t)r * ld + c0));
  const FAP_GLOBAL v4i *p4 = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(gp + (int64_t)r * ld + c0));
  v4i dv[4], pv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    dv[q] = d4[q];
    pv[q] = p4[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      D[r][c0 + 4 * q + c] = dv[q][c];
      P[r][c0 + 4 * q + c] = pv[q][c];
    }
}

__device__ __forceinline__ void lds_to_tile(int (*D)[kTP], int (*P)[kTP], int *gd, int *gp,
                                            int64_t ld) {
  const int r = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  FAP_GLOBAL v4i *d4 = reinterpret_cast<FAP_GLOBAL v4i *>(gptr(gd + (int64_t)r * ld + c0));
  FAP_GLOBAL v4i *p4 = reinterpret_cast<FAP_GLOBAL v4i *>(gptr(gp + (int64_t)r * ld + c0));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    d4[q] = v4i{D[r][c0 + 4 * q], D[r][c0 + 4 * q + 1], D[r][c0 + 4 * q + 2], D[r][c0 + 4 * q + 3]};
    p4[q] = v4i{P[r][c0 + 4 * q], P[r][c0 + 4 * q + 1], P[r][c0 + 4 * q + 2], P[r][c0 + 4 * q + 3]};
  }
}

// Keyed form (a part whose finite Floyd-Warshall entries stay below kFkLim, flag
// fwkey[part]): an entry d with its path source packed as (min(d, kFkLim) << 7) | tag,
// tag 0 = the path the cell had when the phase started, tag k + 1 = set at pivot k.
// A candidate through pivot k is (a << 7) + ((b << 7) | (k + 1)) (a, b clamped), so
// one unsigned minimum keeps the smallest distance, the old value on a tie and the
// first pivot among equal candidates: the strict '>' of the k-i-j loop.  Sums that
// reach kFkLim are never a minimum (every finite entry, and every value the loop
// keeps, is a simple-path length <= (s - 1) * max entry < kFkLim, checked on the host;
// weights are non-negative).  Distance records of such a part are keys: diag row /
// column records (min(d, L) << 7) | (k + 1) for phase 2, column panel records
// min(d, L) << 7 and row panel records (min(d, L) << 7) | (k + 1) for phase 3.
constexpr uint32_t kFkLim = (1u << 24) - 1;

__device__ __forceinline__ uint32_t fkey(int d) { return min((uint32_t)d, kFkLim) << 7; }
__device__ __forceinline__ int fkey_dist(uint32_t key) {
  const uint32_t d = key >> 7;
  return d >= kFkLim ? kInf : (int)d;
}

// Phase 1: the diagonal tile of round r, sequentially over its 64 pivots, recording
// row k (dist, path) and column k as they stand at step k.
__global__ __launch_bounds__(256) void fwp_phase1_kernel(const PathPart *__restrict__ parts,
                                                         const int *__restrict__ fwkey, int r) {
  const PathPart p = parts[blockIdx.x];
  if (r >= p.nb) return;
  const bool keyed = fwkey[blockIdx.x] != 0;
  __shared__ int D[kT][kTP];
  __shared__ int P[kT][kTP];
  tile_to_lds(D, P, tile(p.fw, p, r, r), tile(p.fwp, p, r, r), p.S);
  FAP_GLOBAL int *rrow = gptr(rec_diag_row(p)), *rrowp = gptr(rec_diag_rowp(p)),
                 *rcol = gptr(rec_diag_col(p));
  const int i = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  __syncthreads();
  for (int k = 0; k < kT; ++k) {
    // The records go to global memory and are read by later launches only, so the
    // per-step barrier orders the LDS tile only (__syncthreads would also wait for
    // the step's record stores).
    if (k > 0) lds_barrier();
    if (i == k) {
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int v = D[k][c0 + c];
        rrow[k * kT + c0 + c] = keyed ? (int)(fkey(v) | (k + 1)) : v;
        rrowp[k * kT + c0 + c] = P[k][c0 + c];
      }
    }
    if (k >= c0 && k < c0 + 16) {  // stored k-major
      const int v = D[i][k];
      rcol[k * kT + i] = keyed ? (int)(fkey(v) | (k + 1)) : v;
    }
    const int dik = D[i][k];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int cand = dik + D[k][c0 + c];
      if (D[i][c0 + c] > cand) {
        D[i][c0 + c] = cand;
        P[i][c0 + c] = P[k][c0 + c];
      }
    }
  }
  __syncthreads();
  lds_to_tile(D, P, tile(p.fw, p, r, r), tile(p.fwp, p, r, r), p.S);
}

// Phase 2: blockIdx.x < nb - 1: column panel tile (t, r); else row panel tile (r, t).
__global__ __launch_bounds__(256) void fwp_phase2_kernel(const PathPart *__restrict__ parts,
                                                         const int *__restrict__ fwkey, int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb || fwkey[blockIdx.y]) return;
  const int others = p.nb - 1;
  int t = blockIdx.x;
  if (t >= 2 * others) return;
  const bool colp = t < others;
  if (!colp) t -= others;
  const int o = t < r ? t : t + 1;  // tile index outside the pivot
  __shared__ int D[kT][kTP];
  __shared__ int P[kT][kTP];
  __shared__ int R0[kT][kTP];  // column panel: diag row records; row panel: diag column records
  __shared__ int R1[kT][kTP];  // column panel: diag row path records
  int *gd = colp ? tile(p.fw, p, o, r) : tile(p.fw, p, r, o);
  int *gp = colp ? tile(p.fwp, p, o, r) : tile(p.fwp, p, r, o);
  tile_to_lds(D, P, gd, gp, p.S);
  {
    const FAP_GLOBAL int *a = gptr(colp ? rec_diag_row(p) : rec_diag_col(p));
    const FAP_GLOBAL int *b = gptr(rec_diag_rowp(p));
    for (int e = threadIdx.x; e < kTT; e += 256) {
      R0[e / kT][e % kT] = a[e];
      if (colp) R1[e / kT][e % kT] = b[e];
    }
  }
  FAP_GLOBAL int *rec = gptr(colp ? rec_colpanel(p, o) : rec_rowpanel(p, o));
  FAP_GLOBAL int *recp = gptr(colp ? nullptr : rec_rowpanelp(p, o));
  const int i = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  __syncthreads();
  for (int k = 0; k < kT; ++k) {
    if (k > 0) lds_barrier();  // LDS only: the records are read by later launches
    if (colp) {
      // cell (i, c) of tile (o, r): D[i][k] (this tile) + diag row k at step k.
      if (k >= c0 && k < c0 + 16) rec[k * kT + i] = D[i][k];  // k-major: phase 3 A operand
      const int dik = D[i][k];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int cand = dik + R0[k][c0 + c];
        if (D[i][c0 + c] > cand) {
          D[i][c0 + c] = cand;
          P[i][c0 + c] = R1[k][c0 + c];
        }
      }
    } else {
      // cell (i, c) of tile (r, o): diag column k at step k (row i) + D[k][c] (this tile).
      if (i == k) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          rec[k * kT + c0 + c] = D[k][c0 + c];
          recp[k * kT + c0 + c] = P[k][c0 + c];
        }
      }
      const int cik = R0[k][i];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int cand = cik + D[k][c0 + c];
        if (D[i][c0 + c] > cand) {
          D[i][c0 + c] = cand;
          P[i][c0 + c] = P[k][c0 + c];
        }
      }
    }
  }
  __syncthreads();
  lds_to_tile(D, P, gd, gp, p.S);
}

// Phase 3: tile (a, b) outside the pivot row and column, from the panel records.
// Thread owns a 4 x 4 block: rows ty*4.., columns tx*4.. (ty, tx < 16).
__global__ __launch_bounds__(256) void fwp_phase3_kernel(const PathPart *__restrict__ parts,
                                                         const int *__restrict__ fwkey, int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb || fwkey[blockIdx.y]) return;
  const int others = p.nb - 1;
  const int t = blockIdx.x;
  if (t >= others * others) return;
  const int a0 = t / others, b0 = t - a0 * others;
  const int a = a0 < r ? a0 : a0 + 1, b = b0 < r ? b0 : b0 + 1;
  __shared__ int As[kT][kT];  // As[k][i]: column panel record of tile (a, r)
  __shared__ int Bs[kT][kT];  // Bs[k][j]: row panel record of tile (r, b)
  __shared__ int Bp[kT][kT];
  {
    const FAP_GLOBAL int *ra = gptr(rec_colpanel(p, a)), *rb = gptr(rec_rowpanel(p, b)),
                         *rbp = gptr(rec_rowpanelp(p, b));
    for (int e = threadIdx.x; e < kTT; e += 256) {
      As[e / kT][e % kT] = ra[e];
      Bs[e / kT][e % kT] = rb[e];
      Bp[e / kT][e % kT] = rbp[e];
    }
  }
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  FAP_GLOBAL int *gd = gptr(tile(p.fw, p, a, b)), *gp = gptr(tile(p.fwp, p, a, b));
  int dv[4][4], pv[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      dv[q][c] = gd[(int64_t)(ty * 4 + q) * p.S + tx * 4 + c];
      pv[q][c] = gp[(int64_t)(ty * 4 + q) * p.S + tx * 4 + c];
    }
  __syncthreads();
  for (int k = 0; k < kT; ++k) {
    int av[4], bv[4], bpv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      av[q] = As[k][ty * 4 + q];
      bv[q] = Bs[k][tx * 4 + q];
      bpv[q] = Bp[k][tx * 4 + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cand = av[q] + bv[c];
        if (dv[q][c] > cand) {
          dv[q][c] = cand;
          pv[q][c] = bpv[c];
        }
      }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      gd[(int64_t)(ty * 4 + q) * p.S + tx * 4 + c] = dv[q][c];
      gp[(int64_t)(ty * 4 + q) * p.S + tx * 4 + c] = pv[q][c];
    }
}

// Phase 2, keyed form: one wave per panel tile, each lane holding one line of 64 keys
// in registers (no barriers: every line's update depends on that line and the diag
// records only).  The 64 pivots are expanded at compile time (PanelSteps) so that key
// k is a fixed register.  Column panel: the diag records (64 x 64 keys) are staged in
// LDS once and read as broadcasts (69 against 107 us per launch on MI355X: with scalar
// loads each pivot waited for its own 256 bytes, SMEM returning out of order).  Row
// panel (256 VGPRs, one wave per SIMD): scalar loads from rkc (83 against 135 us with
// the LDS records, whose latency one wave cannot hide).
//
// Column panel tile (o, r), blockIdx.x = o' (o' < nb - 1, skipping r): lane i = row i;
// pivot k uses the lane's own key k and the diag row record k; the path of a cell
// whose tag is k + 1 is the diag row path record k.
// Row panel tile (r, o): lane j = column j; pivot k uses the lane's own key k and the
// diag column record k; a cell whose tag is k + 1 takes the path of cell (k, j) at
// step k.  The tag of key k at step k is kept (8 bits each in tagw); after the pivots,
// the row paths are resolved in pivot order: tab[k][j] holds the tile's original path
// until it is replaced by the resolved one (tab[tag - 1][j], or itself for tag 0).
template <int K, bool kRowPanel>
struct PanelSteps {
  static __device__ __forceinline__ void run(uint32_t (&key)[kT], uint32_t (&tagw)[kT / 4],
                                             const uint32_t *rk, const FAP_CONST uint32_t *rkc,
                                             FAP_GLOBAL uint32_t *recd, int lane) {
    if constexpr (K < kT) {
      const uint32_t kk = key[K];
      const uint32_t dk = kk & ~0x7Fu;
      if constexpr (kRowPanel) {
        recd[K * kT + lane] = dk | (uint32_t)(K + 1);
        tagw[K / 4] |= (kk & 0x7F) << (8 * (K % 4));
      } else {
        recd[K * kT + lane] = dk;  // phase 3 A operand
      }
      if constexpr (kRowPanel) {
#pragma unroll
        for (int c = 0; c < kT; ++c) key[c] = min(key[c], dk + rkc[K * kT + c]);
      } else {
        const uint4 *row = reinterpret_cast<const uint4 *>(rk + K * kT);
#pragma unroll
        for (int c = 0; c < kT; c += 4) {
          const uint4 v = row[c / 4];
          key[c] = min(key[c], dk + v.x);
          key[c + 1] = min(key[c + 1], dk + v.y);
          key[c + 2] = min(key[c + 2], dk + v.z);
          key[c + 3] = min(key[c + 3], dk + v.w);
        }
      }
      PanelSteps<K + 1, kRowPanel>::run(key, tagw, rk, rkc, recd, lane);
    }
  }
};

// The 64 x 64 diag records into LDS (one wave: 16 B per lane per load).
__device__ __forceinline__ void stage_records(uint32_t *rks, const int *rec, int lane) {
  const FAP_GLOBAL v4i *src = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(rec));
  v4i *dst = reinterpret_cast<v4i *>(rks);
#pragma unroll
  for (int q = 0; q < kTT / 4 / 64; ++q) dst[q * 64 + lane] = src[q * 64 + lane];
}

__global__ __launch_bounds__(64) void fwp_colpanel_key_kernel(const PathPart *__restrict__ parts,
                                                              const int *__restrict__ fwkey,
                                                              int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb || !fwkey[blockIdx.y]) return;
  const int t = blockIdx.x;
  if (t >= p.nb - 1) return;
  const int o = t < r ? t : t + 1;
  const int lane = threadIdx.x;
  // One LDS buffer: the diag row records during the pivots, then the diag row path
  // records (tab).
  __shared__ uint4 buf4[(kT * kTP + 3) / 4];
  uint32_t *rks = reinterpret_cast<uint32_t *>(buf4);
  int(*tab)[kTP] = reinterpret_cast<int(*)[kTP]>(buf4);
  FAP_GLOBAL int *gd = gptr(tile(p.fw, p, o, r) + (int64_t)lane * p.S);
  FAP_GLOBAL int *gp = gptr(tile(p.fwp, p, o, r) + (int64_t)lane * p.S);
  uint32_t key[kT];
#pragma unroll
  for (int c = 0; c < kT; c += 4) {
    const v4i v = *reinterpret_cast<const FAP_GLOBAL v4i *>(gd + c);
    key[c] = fkey(v.x);
    key[c + 1] = fkey(v.y);
    key[c + 2] = fkey(v.z);
    key[c + 3] = fkey(v.w);
  }
  stage_records(rks, rec_diag_row(p), lane);
  __syncthreads();
  uint32_t tagw[kT / 4];  // unused by the column panel
  PanelSteps<0, false>::run(key, tagw, rks, nullptr,
                            gptr(reinterpret_cast<uint32_t *>(rec_colpanel(p, o))), lane);
  __syncthreads();
  const FAP_GLOBAL int *rp = gptr(rec_diag_rowp(p));
  for (int e = lane; e < kTT; e += 64) tab[e >> 6][e & 63] = rp[e];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kT; c += 4) {
    const v4i v = {fkey_dist(key[c]), fkey_dist(key[c + 1]), fkey_dist(key[c + 2]),
                   fkey_dist(key[c + 3])};
    *reinterpret_cast<FAP_GLOBAL v4i *>(gd + c) = v;
  }
#pragma unroll
  for (int c = 0; c < kT; ++c) {
    const uint32_t tg = key[c] & 0x7F;
    if (tg) gp[c] = tab[tg - 1][c];
  }
}

__global__ __launch_bounds__(64) void fwp_rowpanel_key_kernel(const PathPart *__restrict__ parts,
                                                              const int *__restrict__ fwkey,
                                                              int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb || !fwkey[blockIdx.y]) return;
  const int t = blockIdx.x;
  if (t >= p.nb - 1) return;
  const int o = t < r ? t : t + 1;
  const unsigned lane = threadIdx.x;
  __shared__ int tab[kT][kTP];  // row paths of this tile, resolved as each step records them
  // The tile's rows are staged through tab (16 B per lane, 4 rows per load), the lane
  // then reads its column: first the distances, then the paths.  The results go back
  // the same way (on MI355X 83 us per launch, against 98 us writing each lane's column
  // with 128 dword stores, although holding 64 paths besides the 64 keys spills).
  FAP_GLOBAL int *gd = gptr(tile(p.fw, p, r, o));
  FAP_GLOBAL int *gp = gptr(tile(p.fwp, p, r, o));
  const int sr = lane >> 4, sc = (lane & 15) * 4;  // staging: row sr + 4 q, columns sc..
  auto stage = [&](const FAP_GLOBAL int *g) {
#pragma unroll
    for (int q = 0; q < kT / 4; ++q) {
      const v4i v = *reinterpret_cast<const FAP_GLOBAL v4i *>(g + (int64_t)(sr + 4 * q) * p.S + sc);
      tab[sr + 4 * q][sc] = v.x;
      tab[sr + 4 * q][sc + 1] = v.y;
      tab[sr + 4 * q][sc + 2] = v.z;
      tab[sr + 4 * q][sc + 3] = v.w;
    }
  };
  auto unstage = [&](FAP_GLOBAL int *g) {
#pragma unroll
    for (int q = 0; q < kT / 4; ++q) {
      const v4i v = {tab[sr + 4 * q][sc], tab[sr + 4 * q][sc + 1], tab[sr + 4 * q][sc + 2],
                     tab[sr + 4 * q][sc + 3]};
      *reinterpret_cast<FAP_GLOBAL v4i *>(g + (int64_t)(sr + 4 * q) * p.S + sc) = v;
    }
  };
  uint32_t key[kT];
  stage(gd);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kT; ++i) key[i] = fkey(tab[i][lane]);
  __syncthreads();
  stage(gp);
  __syncthreads();
  uint32_t tagw[kT / 4];
#pragma unroll
  for (int q = 0; q < kT / 4; ++q) tagw[q] = 0;
  PanelSteps<0, true>::run(key, tagw, nullptr,
                           uniform_cptr(reinterpret_cast<const uint32_t *>(rec_diag_col(p))),
                           gptr(reinterpret_cast<uint32_t *>(rec_rowpanel(p, o))), lane);
  {
    FAP_GLOBAL int *recp = gptr(rec_rowpanelp(p, o));
#pragma unroll
    for (int k = 0; k < kT; ++k) {  // tag <= k: row tag - 1 is resolved
      const uint32_t tg = (tagw[k / 4] >> (8 * (k % 4))) & 0x7F;
      const int pk = tab[tg ? tg - 1 : k][lane];
      tab[k][lane] = pk;  // this lane's column only: no barrier
      recp[k * kT + lane] = pk;
    }
  }
  // Final paths: tag t -> row path t - 1; tag 0 -> tab[i][j], which is the path the
  // cell had at step i, i.e. its original one when no pivot ever replaced it.
  int pv[kT];
#pragma unroll
  for (int i = 0; i < kT; ++i) {
    const uint32_t tg = key[i] & 0x7F;
    pv[i] = tab[tg ? tg - 1 : i][lane];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kT; ++i) tab[i][lane] = fkey_dist(key[i]);
  __syncthreads();
  unstage(gd);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kT; ++i) tab[i][lane] = pv[i];
  __syncthreads();
  unstage(gp);
}

// Phase 3, keyed form: tile (a, b) from the keyed panel records; thread owns a 4 x 4
// block (rows ty * 4.., columns tx * 4..), two pivots per v_min3.  Only cells whose
// tag is set get a new path (the row panel path record of that pivot).
__global__ __launch_bounds__(256) void fwp_phase3_key_kernel(const PathPart *__restrict__ parts,
                                                             const int *__restrict__ fwkey,
                                                             int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb || !fwkey[blockIdx.y]) return;
  const int others = p.nb - 1;
  const int t = blockIdx.x;
  if (t >= others * others) return;
  const int a0 = t / others, b0 = t - a0 * others;
  const int a = a0 < r ? a0 : a0 + 1, b = b0 < r ? b0 : b0 + 1;
  __shared__ uint4 As[kT][kT / 4];  // As[k][i / 4]: column panel keys of tile (a, r)
  __shared__ uint4 Bs[kT][kT / 4];  // Bs[k][j / 4]: row panel keys of tile (r, b)
  __shared__ int Bp[kT][kTP];       // row panel path records
  {
    const FAP_GLOBAL v4i *ra = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(rec_colpanel(p, a)));
    const FAP_GLOBAL v4i *rb = reinterpret_cast<const FAP_GLOBAL v4i *>(gptr(rec_rowpanel(p, b)));
    const FAP_GLOBAL int *rbp = gptr(rec_rowpanelp(p, b));
    v4i x[kTT / 4 / 256], y[kTT / 4 / 256];
    int z[kTT / 256];
#pragma unroll
    for (int u = 0; u < kTT / 4 / 256; ++u) {
      x[u] = ra[threadIdx.x + 256 * u];
      y[u] = rb[threadIdx.x + 256 * u];
    }
#pragma unroll
    for (int u = 0; u < kTT / 256; ++u) z[u] = rbp[threadIdx.x + 256 * u];
#pragma unroll
    for (int u = 0; u < kTT / 4 / 256; ++u) {
      const int e = threadIdx.x + 256 * u;
      As[e >> 4][e & 15] = make_uint4(x[u].x, x[u].y, x[u].z, x[u].w);
      Bs[e >> 4][e & 15] = make_uint4(y[u].x, y[u].y, y[u].z, y[u].w);
    }
#pragma unroll
    for (int u = 0; u < kTT / 256; ++u) {
      const int e = threadIdx.x + 256 * u;
      Bp[e >> 6][e & 63] = z[u];
    }
  }
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  FAP_GLOBAL int *gd = gptr(tile(p.fw, p, a, b) + (int64_t)(ty * 4) * p.S + tx * 4);
  FAP_GLOBAL int *gp = gptr(tile(p.fwp, p, a, b) + (int64_t)(ty * 4) * p.S + tx * 4);
  uint32_t key[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const v4i v = *reinterpret_cast<const FAP_GLOBAL v4i *>(gd + (int64_t)q * p.S);
    key[q][0] = fkey(v.x);
    key[q][1] = fkey(v.y);
    key[q][2] = fkey(v.z);
    key[q][3] = fkey(v.w);
  }
  __syncthreads();
#pragma unroll 4
  for (int k = 0; k < kT; k += 2) {
    const uint4 x0 = As[k][ty], y0 = Bs[k][tx], x1 = As[k + 1][ty], y1 = Bs[k + 1][tx];
    const uint32_t av0[4] = {x0.x, x0.y, x0.z, x0.w}, bv0[4] = {y0.x, y0.y, y0.z, y0.w};
    const uint32_t av1[4] = {x1.x, x1.y, x1.z, x1.w}, bv1[4] = {y1.x, y1.y, y1.z, y1.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        key[q][c] = min(key[q][c], min(av0[q] + bv0[c], av1[q] + bv1[c]));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const v4i v = {fkey_dist(key[q][0]), fkey_dist(key[q][1]), fkey_dist(key[q][2]),
                   fkey_dist(key[q][3])};
    *reinterpret_cast<FAP_GLOBAL v4i *>(gd + (int64_t)q * p.S) = v;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t tg = key[q][c] & 0x7F;
      if (tg) gp[(int64_t)q * p.S + c] = Bp[tg - 1][tx * 4 + c];
    }
  }
}

// ---- min-plus of the inner rows with paths, and the decode of own columns ----------
// Work unit: 16 inner rows x 128 columns of one part (a wave; 2 columns per lane).  k is
// walked in chunks of 8 with the B (boundary rows) and Bpath values of the chunk in
// registers; the running minimum and its path persist per row and column, so the
// first k reaching the minimum is kept (strict '<').  A values (the part's FW matrix,
// boundary columns of its inner rows) are wave-uniform.  Columns of the part's own
// vertices then take the Floyd-Warshall distance and path (decode).
constexpr int kMpRows = 16;
constexpr int kMpCols = 128;
constexpr int kMpKc = 8;

struct PathMergeDesc {
  int part;           // index into the part table
  int unit_begin;     // first unit of this part
  int row_blocks;     // ceil(inner / 16)
};

__global__ __launch_bounds__(64) void merge_path_kernel(
    const PathPart *__restrict__ parts, const PathMergeDesc *__restrict__ md, int nmd,
    int total_units, int n, const int *__restrict__ graph_id, const int *__restrict__ ed2st,
    int *__restrict__ out, int *__restrict__ pout, int64_t ld_out) {
  const int u = blockIdx.x;
  if (u >= total_units) return;
  int lo = 0, hi = nmd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (md[mid].unit_begin <= u) lo = mid;
    else hi = mid - 1;
  }
  const PathMergeDesc m = md[lo];
  const PathPart p = parts[m.part];
  const int local = u - m.unit_begin;
  const int cb = local / m.row_blocks, rb = local - cb * m.row_blocks;
  const int inner = p.s - p.bdy;
  const int r0 = rb * kMpRows, r1 = min(inner, r0 + kMpRows);
  const int j0 = cb * kMpCols + threadIdx.x * 2;
  const int K = p.bdy;
  const int64_t brow0 = p.out_row * ld_out;  // boundary row 0 of the part
  int best[kMpRows][2], bpath[kMpRows][2];
#pragma unroll
  for (int q = 0; q < kMpRows; ++q)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      best[q][e] = kInf;
      bpath[q][e] = -1;
    }
  for (int k0 = 0; k0 < K; k0 += kMpKc) {
    int bv[kMpKc][2], bp[kMpKc][2];
#pragma unroll
    for (int kk = 0; kk < kMpKc; ++kk)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int j = j0 + e;
        const bool ok = k0 + kk < K && j < n;
        bv[kk][e] = ok ? out[brow0 + (int64_t)(k0 + kk) * ld_out + j] : kInf;
        bp[kk][e] = ok ? pout[brow0 + (int64_t)(k0 + kk) * ld_out + j] : -1;
      }
#pragma unroll
    for (int q = 0; q < kMpRows; ++q) {
      const int i = r0 + q;
      if (i >= r1) continue;
      const FAP_GLOBAL int *arow = gptr(p.fw + (int64_t)(p.bdy + i) * p.S + k0);  // A = FW[bdy + i][k]
#pragma unroll
      for (int kk = 0; kk < kMpKc; ++kk) {
        const int av = k0 + kk < K ? arow[kk] : kInf;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int cand = av + bv[kk][e];
          if (cand < best[q][e]) {  // cand < FAP_INF: a finite sum
            best[q][e] = cand;
            bpath[q][e] = bp[kk][e];
          }
        }
      }
    }
  }
  // Decode: own columns take the FW distance and path.
  int own[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int j = j0 + e;
    own[e] = (j < n && graph_id[j] == p.part_id) ? ed2st[j] - p.offset : -1;
  }
#pragma unroll
  for (int q = 0; q < kMpRows; ++q) {
    const int i = r0 + q;
    if (i >= r1) continue;
    const int64_t orow = (p.out_row + p.bdy + i) * ld_out;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int j = j0 + e;
      if (j >= n) continue;
      int dv = best[q][e], pv = bpath[q][e];
      if (own[e] >= 0) {
        dv = gptr(p.fw)[(int64_t)(p.bdy + i) * p.S + own[e]];
        pv = gptr(p.fwp)[(int64_t)(p.bdy + i) * p.S + own[e]];
      }
      out[orow + j] = dv;
      pout[orow + j] = pv;
    }
  }
}

// Largest finite entry of a part's inner rows in its boundary columns (the A operand
// of the min-plus) after Floyd-Warshall, -1 when there is none; one workgroup per part.
// Parts with akey_kp[part] = kp > 0 (8, 16 or 32, with kb = 3, 4 or 5 tag bits) also
// get their A keys for the keyed min-plus: akey[akey_off + i * kp + k] =
// min(A[i][k], L) << kb for k < bdy, 2^31 for bdy <= k < kp (paired with an infinite
// B key, never the minimum).
__global__ __launch_bounds__(256) void path_amax_kernel(const PathPart *__restrict__ parts,
                                                        int *__restrict__ amax,
                                                        uint32_t *__restrict__ akey,
                                                        const int64_t *__restrict__ akey_off,
                                                        const int *__restrict__ akey_kp) {
  const PathPart p = parts[blockIdx.x];
  const int inner = p.s - p.bdy;
  const int64_t ko = akey_off[blockIdx.x];
  const int kp = akey_kp[blockIdx.x];
  const int kb = kp == 8 ? 3 : kp == 16 ? 4 : 5;
  const uint32_t lim = (1u << (31 - kb)) - 1;
  int mx = -1;
  for (int64_t e = threadIdx.x; e < (int64_t)inner * p.bdy; e += 256) {
    const int i = (int)(e / p.bdy), k = (int)(e - (int64_t)i * p.bdy);
    const int v = gptr(p.fw)[(int64_t)(p.bdy + i) * p.S + k];
    if (v < kInf) mx = max(mx, v);
  }
  if (kp > 0) {
    for (int64_t e = threadIdx.x; e < (int64_t)inner * kp; e += 256) {
      const int i = (int)(e / kp), k = (int)(e - (int64_t)i * kp);
      uint32_t key = 0x80000000u;
      if (k < p.bdy) key = min((uint32_t)gptr(p.fw)[(int64_t)(p.bdy + i) * p.S + k], lim) << kb;
      akey[ko + e] = key;
    }
  }
  __shared__ int red;
  if (threadIdx.x == 0) red = -1;
  __syncthreads();
  atomicMax(&red, mx);
  __syncthreads();
  if (threadIdx.x == 0) amax[blockIdx.x] = red;
}

// ---- min-plus with paths, keyed form ---------------------------------------------
// When every finite A entry plus every finite boundary-row entry stays below
// L = 2^(31 - KB) - 1 and the part has at most 2^KB boundary vertices, a candidate
// A[i][k] + B[k][j] and its k fit one 32-bit key: (min(a, L) << KB) +
// ((min(b, L) << KB) | k).  The unsigned minimum over k is then the smallest sum with
// ties to the smallest k, which is what the strict '<' over ascending k keeps, and
// the path is Bpath[key & (2^KB - 1)][j]; a key whose sum reaches L is an infinite
// candidate (finite sums stay below L), so the initial key L << KB means FAP_INF / -1.
// One add and half a v_min3 per candidate against an add, a compare and two selects.
//
// Unit: a workgroup of 4 waves, kMkCols columns (lane l holds cb * 256 + c * 64 + l,
// c < 4, so every store writes 256 contiguous bytes) x up to kMkRows inner rows of one
// part, wave w taking rows w, w + 4, ...  The B keys of the unit's columns stay in
// registers for all its rows, the boundary paths sit in LDS once per workgroup as
// [k][c][lane] (a gather of any k per lane is bank-conflict free), the A keys
// (path_amax_kernel, KP per row) come in with scalar loads, the next row's while the
// current one is reduced.  The part's own columns are written here too and then
// overwritten by decode_path_kernel (kept out of this loop: a load between the stores
// would make every store wait).
constexpr int kMkCols = 256;
constexpr int kMkRows = 1024;
constexpr int kMkWaves = 4;

struct PathKeyDesc {
  int part;              // index into the part table
  int unit_begin;        // first unit of this part
  int col_blocks;        // ceil(n / 256)
  int pad;
  const uint32_t *akey;  // A keys of the part's inner rows, KP per row
};

template <int KP, int KB>
__global__ __launch_bounds__(256) void merge_path_key_kernel(
    const PathPart *__restrict__ parts, const PathKeyDesc *__restrict__ md, int nmd,
    int total_units, int n, int *__restrict__ out, int *__restrict__ pout, int64_t ld_out) {
  static_assert((1 << KB) >= KP, "tag field too narrow");
  constexpr uint32_t L = (1u << (31 - KB)) - 1;
  constexpr uint32_t kInit = L << KB;
  constexpr uint32_t kTag = (1u << KB) - 1;
  extern __shared__ int bpl[];  // K x 4 x 64 boundary paths
  const int u = blockIdx.x;
  if (u >= total_units) return;
  int lo = 0, hi = nmd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (md[mid].unit_begin <= u) lo = mid;
    else hi = mid - 1;
  }
  const PathKeyDesc m = md[lo];
  const PathPart p = parts[m.part];
  const int local = u - m.unit_begin;
  const int rc = local / m.col_blocks, cb = local - rc * m.col_blocks;
  const int inner = p.s - p.bdy, K = p.bdy;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = rc * kMkRows + wave, r1 = min(inner, rc * kMkRows + kMkRows);
  const int jb = cb * kMkCols + lane;
  const int64_t brow0 = p.out_row * ld_out;
  for (int k = wave; k < K; k += kMkWaves)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = jb + c * 64;
      bpl[(k * 4 + c) * 64 + lane] = j < n ? pout[brow0 + (int64_t)k * ld_out + j] : -1;
    }
  uint32_t bk[KP][4];
#pragma unroll
  for (int k = 0; k < KP; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = jb + c * 64;
      bk[k][c] = kInit;  // k >= K: paired with an A key of 2^31, never the minimum
      if (k < K) {
        const uint32_t b = j < n ? (uint32_t)out[brow0 + (int64_t)k * ld_out + j] : L;
        bk[k][c] = (min(b, L) << KB) | (uint32_t)k;
      }
    }
  __syncthreads();
  // The key array holds kMkWaves rows of slack past the last part: reading the row
  // after a wave's last one is safe.
  const FAP_CONST uint32_t *arow = uniform_cptr(m.akey + (int64_t)r0 * KP);
  uint32_t a[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) a[k] = arow[k];
  for (int i = r0; i < r1; i += kMkWaves) {
    arow += kMkWaves * KP;
    uint32_t an[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) an[k] = arow[k];
    uint32_t best[4] = {kInit, kInit, kInit, kInit};
#pragma unroll
    for (int k = 0; k < KP; k += 2) {
      if (k < K) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          best[c] = min(best[c], min(a[k] + bk[k][c], a[k + 1] + bk[k + 1][c]));
      }
    }
    const int64_t orow = (p.out_row + p.bdy + i) * ld_out;
    // The minimum's tag is below K (the initial key has tag 0): the reads stay in the
    // K rows of bpl.
    int bp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bp[c] = bpl[((best[c] & kTag) * 4 + c) * 64 + lane];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = jb + c * 64;
      const uint32_t sum = best[c] >> KB;
      if (j < n) {
        __builtin_nontemporal_store(sum < L ? (int)sum : kInf, out + orow + j);
        __builtin_nontemporal_store(sum < L ? bp[c] : -1, pout + orow + j);
      }
    }
#pragma unroll
    for (int k = 0; k < KP; ++k) a[k] = an[k];
  }
}

// Own columns: out[row][st2ed[offset + c]] = FW[row][c] (dist, path) for rows
// 0 .. rows_p - 1 of each part: its boundary rows, and all its rows when the keyed
// min-plus wrote the inner rows.  One workgroup per row (row_begin: prefix of rows_p).
__global__ __launch_bounds__(256) void decode_path_kernel(const PathPart *__restrict__ parts,
                                                          int np,
                                                          const int64_t *__restrict__ row_begin,
                                                          int64_t total,
                                                          const int *__restrict__ st2ed,
                                                          int *__restrict__ out,
                                                          int *__restrict__ pout, int64_t ld_out) {
  const int64_t row = blockIdx.x;
  if (row >= total) return;
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (row_begin[mid] <= row) lo = mid;
    else hi = mid - 1;
  }
  const PathPart p = parts[lo];
  const int i = (int)(row - row_begin[lo]);
  const int64_t orow = (p.out_row + i) * ld_out;
  const FAP_GLOBAL int *fr = gptr(p.fw + (int64_t)i * p.S), *fpr = gptr(p.fwp + (int64_t)i * p.S);
  for (int c = threadIdx.x; c < p.s; c += 256) {
    const int v = st2ed[p.offset + c];
    out[orow + v] = fr[c];
    pout[orow + v] = fpr[c];
  }
}

Workspace g_ws_pfw;    // FW distance and path matrices
Workspace g_ws_psnap;  // per-part records
Workspace g_ws_pmeta;  // descriptors, sources, row offsets
Workspace g_ws_pell;   // ELL adjacency (unit weights)
Workspace g_ws_pakey;  // A keys of the keyed min-plus

}  // namespace
}  // namespace fap_hip

using namespace fap_hip;

static int solve_parts_path(const fap_graph_dev *g, const int32_t *d_rro, const int32_t *d_rcol,
                            const int32_t *d_rw, const fap_part_desc *parts, int32_t nparts,
                            int32_t *d_out, int32_t *d_path, int64_t ld_out, hipStream_t st) {
  const int n = g->n;
  FAP_REQUIRE(n > 0 && g->m >= 0 && ld_out >= n && g->d_row_offset && g->d_graph_id &&
                  g->d_st2ed && g->d_ed2st && (g->m == 0 || g->d_col) && d_rro &&
                  (g->m == 0 || d_rcol) && (d_rw != nullptr) == (g->d_w != nullptr),
              "fap_hip_solve_parts_path: bad graph view");
  FAP_REQUIRE(g->d_wf == nullptr, "fap_hip_solve_parts_path: float weights have no int32 paths");
  std::vector<fap_part_desc> act;
  for (int k = 0; k < nparts; ++k) {
    const fap_part_desc &q = parts[k];
    FAP_REQUIRE(q.part_id >= 1 && q.num >= 0 && q.bdy >= 0 && q.bdy <= q.num && q.offset >= 0 &&
                    (int64_t)q.offset + q.num <= n && q.out_row >= 0 && q.num <= kMaxPart,
                "fap_hip_solve_parts_path: bad part %d", k);
    if (q.num > 0) act.push_back(q);
  }
  const int P = (int)act.size();
  FAP_TRY(timing_start(st));
  if (P == 0) return timing_finish(st);

  std::vector<int> S(P), nb(P);
  std::vector<int64_t> fw_off(P), snap_off(P);
  int64_t fw_total = 0, snap_total = 0;
  int max_nb = 0, rows_total = 0, src_total = 0;
  for (int k = 0; k < P; ++k) {
    S[k] = (int)align_up(act[k].num, kT);
    nb[k] = S[k] / kT;
    max_nb = std::max(max_nb, nb[k]);
    fw_off[k] = fw_total;
    fw_total += (int64_t)S[k] * S[k];
    snap_off[k] = snap_total;
    snap_total += (int64_t)(3 + 3 * nb[k]) * kTT;
  }
  FAP_TRY(g_ws_pfw.ensure(fw_total * 8));
  FAP_TRY(g_ws_psnap.ensure(snap_total * 4));
  int *fw_base = static_cast<int *>(g_ws_pfw.ptr);
  int *fwp_base = fw_base + fw_total;
  int *snap_base = static_cast<int *>(g_ws_psnap.ptr);

  std::vector<PathPart> pp(P);
  std::vector<int> offsets(P), src_begin(P);
  std::vector<int64_t> out_rows(P);
  int merge_parts = 0;
  for (int k = 0; k < P; ++k) {
    const fap_part_desc &q = act[k];
    pp[k] = PathPart{fw_base + fw_off[k], fwp_base + fw_off[k], snap_base + snap_off[k], S[k],
                     nb[k], q.num, q.bdy, q.offset, q.part_id, rows_total, q.out_row};
    rows_total += S[k];
    offsets[k] = q.offset;
    src_begin[k] = src_total;
    out_rows[k] = q.out_row;
    src_total += q.bdy;
    merge_parts += q.num > q.bdy;
  }

  int64_t off = 0;
  auto carve = [&](int64_t bytes) {
    const int64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const int64_t o_pp = carve(sizeof(PathPart) * P);
  const int64_t o_off = carve(4LL * P), o_sb = carve(4LL * P), o_or = carve(8LL * P);
  const int64_t upload = off;
  const int64_t o_db = carve(8LL * P);
  const int64_t o_md = carve(sizeof(PathMergeDesc) * P);
  const int64_t o_kd = carve(sizeof(PathKeyDesc) * P);
  const int nstat = kMaxSlots * kSlotStride + P;  // boundary-row maximum slots, then A maxima
  const int64_t o_stat = carve(4LL * nstat);
  const int64_t o_akoff = carve(12LL * P);  // A key offsets (int64), then classes (int)
  const int64_t o_pmax = carve(4LL * P * kSlotStride);  // largest finite built entry per part
  const int64_t o_fwkey = carve(4LL * P);    // 1: keyed Floyd-Warshall
  const int64_t o_src = carve(4LL * std::max(src_total, 1));
  const int64_t o_rows = carve(8LL * std::max(src_total, 1));
  FAP_TRY(g_ws_pmeta.ensure(off));
  char *meta = static_cast<char *>(g_ws_pmeta.ptr);
  std::vector<char> host(upload);
  std::memcpy(host.data() + o_pp, pp.data(), sizeof(PathPart) * P);
  std::memcpy(host.data() + o_off, offsets.data(), 4LL * P);
  std::memcpy(host.data() + o_sb, src_begin.data(), 4LL * P);
  std::memcpy(host.data() + o_or, out_rows.data(), 8LL * P);
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host.data(), upload, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  const PathPart *d_pp = reinterpret_cast<const PathPart *>(meta + o_pp);
  PathMergeDesc *d_md = reinterpret_cast<PathMergeDesc *>(meta + o_md);
  PathKeyDesc *d_kd = reinterpret_cast<PathKeyDesc *>(meta + o_kd);
  int *d_stat = reinterpret_cast<int *>(meta + o_stat);
  int64_t *d_akoff = reinterpret_cast<int64_t *>(meta + o_akoff);
  int *d_pmax = reinterpret_cast<int *>(meta + o_pmax);
  int *d_fwkey = reinterpret_cast<int *>(meta + o_fwkey);
  const int *d_offsets = reinterpret_cast<const int *>(meta + o_off);
  const int *d_sb = reinterpret_cast<const int *>(meta + o_sb);
  const int64_t *d_or = reinterpret_cast<const int64_t *>(meta + o_or);
  int64_t *d_db = reinterpret_cast<int64_t *>(meta + o_db);
  int *d_src = reinterpret_cast<int *>(meta + o_src);
  int64_t *d_rows = reinterpret_cast<int64_t *>(meta + o_rows);

  auto span = [&](int stage, auto &&fn) -> int {
    hipEvent_t a, b;
    FAP_TRY(timing_record(st, &a));
    FAP_TRY(fn());
    FAP_TRY(timing_record(st, &b));
    timing_add_span(stage, a, b);
    return FAP_OK;
  };
  const bool vec = (ld_out % 4 == 0) && ((uintptr_t)d_out % 16 == 0);
  bool bgraph = false;
  // 1. boundary rows: distances from the boundary graph (parts covering a unit-weight
  // graph with paired arcs; 0.8 against 3.0 ms of BFS on the 114,599-vertex road-like
  // graph, K = 128) or by SSSP, then their predecessors, which depend on the
  // distances only.
  FAP_TRY(span(0, [&]() -> int {
    FAP_TRY(fap_hip_fill_i32(d_stat, -1, kMaxSlots * kSlotStride, st));
    if (src_total == 0) return FAP_OK;
    FAP_TRY(launch_gather_sources(d_offsets, d_sb, d_or, P, src_total, g->d_st2ed, ld_out, d_src,
                                  d_rows, st));
    int s = FAP_ERR_INVALID;
    if (!g->d_w && g->d_comp) {
      std::vector<int *> fwm(P);
      for (int k = 0; k < P; ++k) fwm[k] = fw_base + fw_off[k];
      s = boundary_rows_unit(g, act.data(), P, fwm.data(), S.data(), d_out, ld_out, st);
      if (s != FAP_OK && s != FAP_ERR_INVALID) return s;
      bgraph = s == FAP_OK;
    }
    if (s != FAP_OK && !g->d_w) {
      FAP_TRY(g_ws_pell.ensure((int64_t)n * sizeof(int4)));
      int4 *ell = static_cast<int4 *>(g_ws_pell.ptr);
      FAP_TRY(launch_ell_build(g->d_row_offset, g->d_col, n, ell, st));
      s = launch_bfs_lds(ell, g->d_row_offset, g->d_col, n, d_src, src_total, d_rows, d_out, 1,
                         vec, st);
      if (s != FAP_OK && s != FAP_ERR_INVALID) return s;
    }
    if (s != FAP_OK)
      FAP_TRY(launch_sssp(g->d_row_offset, g->d_col, g->d_w, n, d_src, src_total, d_rows, d_out,
                          1, vec, st));
    for (int k0 = 0; k0 < src_total; k0 += 65535 * kPredRows) {
      const int nk = std::min(65535 * kPredRows, src_total - k0);
      hipLaunchKernelGGL(pred_rows_kernel, dim3((n + 255) / 256, (nk + kPredRows - 1) / kPredRows),
                         dim3(256), 0, st, d_rro, d_rcol, d_rw, n, d_src, d_rows, (int64_t)0, k0,
                         nk, d_out, d_path, d_stat);
      FAP_CHECK_LAUNCH();
    }
    return FAP_OK;
  }));
  // 2. sub-matrix build with paths.
  // The Floyd-Warshall of a part is keyed when (s - 1) * (largest finite built entry)
  // stays below kFkLim (read back here): no simple path in its matrix is longer.
  int fw_keyed = 0;
  FAP_TRY(span(1, [&]() -> int {
    FAP_TRY(fap_hip_fill_i32(d_pmax, -1, (int64_t)P * kSlotStride, st));
    hipLaunchKernelGGL(build_path_kernel, dim3(rows_total), dim3(256), 0, st, d_pp, P, rows_total,
                       g->d_row_offset, g->d_col, g->d_w, g->d_graph_id, g->d_st2ed, g->d_ed2st,
                       d_out, d_path, ld_out, d_pmax);
    FAP_CHECK_LAUNCH();
    std::vector<int> pm((size_t)P * kSlotStride), fk(P);
    FAP_CHECK_HIP(hipMemcpyAsync(pm.data(), d_pmax, 4LL * P * kSlotStride, hipMemcpyDeviceToHost,
                                 st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < P; ++k) {
      const int64_t mk = std::max(pm[(size_t)k * kSlotStride], 0);
      fk[k] = (int64_t)std::max(act[k].num - 1, 0) * mk < (int64_t)kFkLim;
      fw_keyed += fk[k];
    }
    FAP_CHECK_HIP(hipMemcpyAsync(d_fwkey, fk.data(), 4LL * P, hipMemcpyHostToDevice, st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    return FAP_OK;
  }));
  // 3. Floyd-Warshall with k-i-j path semantics, all parts per launch.
  FAP_TRY(span(2, [&]() -> int {
    for (int r = 0; r < max_nb; ++r) {
      hipLaunchKernelGGL(fwp_phase1_kernel, dim3(P), dim3(256), 0, st, d_pp, d_fwkey, r);
      FAP_CHECK_LAUNCH();
      if (max_nb == 1) continue;
      const dim3 g2(2 * (max_nb - 1), P), g3((max_nb - 1) * (max_nb - 1), P);
      if (fw_keyed > 0) {
        const dim3 gp(max_nb - 1, P);
        hipLaunchKernelGGL(fwp_colpanel_key_kernel, gp, dim3(64), 0, st, d_pp, d_fwkey, r);
        FAP_CHECK_LAUNCH();
        hipLaunchKernelGGL(fwp_rowpanel_key_kernel, gp, dim3(64), 0, st, d_pp, d_fwkey, r);
        FAP_CHECK_LAUNCH();
      }
      if (fw_keyed < P) {
        hipLaunchKernelGGL(fwp_phase2_kernel, g2, dim3(256), 0, st, d_pp, d_fwkey, r);
        FAP_CHECK_LAUNCH();
      }
      if (fw_keyed > 0) {
        hipLaunchKernelGGL(fwp_phase3_key_kernel, g3, dim3(256), 0, st, d_pp, d_fwkey, r);
        FAP_CHECK_LAUNCH();
      }
      if (fw_keyed < P) {
        hipLaunchKernelGGL(fwp_phase3_kernel, g3, dim3(256), 0, st, d_pp, d_fwkey, r);
        FAP_CHECK_LAUNCH();
      }
    }
    return FAP_OK;
  }));
  // 4. min-plus of the inner rows with paths.  A part takes the keyed form when its
  // boundary is 1..32 vertices and its largest finite A entry plus the largest finite
  // boundary-row entry stays below the key limit (both read back here), the
  // compare-and-select form (own columns decoded in its epilogue) otherwise.
  std::vector<char> keyed(P, 0);
  int merge_keyed = 0;
  FAP_TRY(span(3, [&]() -> int {
    if (merge_parts == 0) return FAP_OK;
    // Key classes by boundary size: (KP, KB) = (8, 3), (16, 4), (32, 5).
    auto kp_of = [&](int k) {
      const int K = act[k].bdy;
      if (act[k].num <= K || K < 1 || K > 32) return 0;
      return K <= 8 ? 8 : K <= 16 ? 16 : 32;
    };
    std::vector<int64_t> akey_off(P, -1);
    std::vector<int> akey_kp(P);
    int64_t akey_total = 0;
    for (int k = 0; k < P; ++k) {
      akey_kp[k] = kp_of(k);
      if (akey_kp[k] == 0) continue;
      akey_off[k] = akey_total;
      akey_total += (int64_t)(act[k].num - act[k].bdy) * akey_kp[k];
    }
    FAP_TRY(g_ws_pakey.ensure(4 * (akey_total + kMkWaves * 32)));  // + rows of slack
    uint32_t *d_akey = static_cast<uint32_t *>(g_ws_pakey.ptr);
    int *d_amax = d_stat + kMaxSlots * kSlotStride;
    std::vector<char> up(12LL * P);
    std::memcpy(up.data(), akey_off.data(), 8LL * P);
    std::memcpy(up.data() + 8LL * P, akey_kp.data(), 4LL * P);
    FAP_CHECK_HIP(hipMemcpyAsync(d_akoff, up.data(), 12LL * P, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(path_amax_kernel, dim3(P), dim3(256), 0, st, d_pp, d_amax, d_akey, d_akoff,
                       reinterpret_cast<const int *>(d_akoff + P));
    FAP_CHECK_LAUNCH();
    std::vector<int> stat(nstat);
    FAP_CHECK_HIP(hipMemcpyAsync(stat.data(), d_stat, 4LL * nstat, hipMemcpyDeviceToHost, st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    int bmax = -1;
    for (int q = 0; q < kMaxSlots; ++q) bmax = std::max(bmax, stat[q * kSlotStride]);
    const int *amax = stat.data() + kMaxSlots * kSlotStride;
    std::vector<PathMergeDesc> md;
    std::vector<PathKeyDesc> kd[3];
    int units = 0, kunits[3] = {0, 0, 0}, kmax[3] = {0, 0, 0};
    int64_t rows[4] = {0, 0, 0, 0};  // inner rows per key class, then the compare form
    const int cblocks = (n + kMkCols - 1) / kMkCols;
    for (int k = 0; k < P; ++k) {
      const int inner = act[k].num - act[k].bdy, K = act[k].bdy;
      if (inner <= 0) continue;
      const int kp = akey_kp[k], cls = kp == 8 ? 0 : kp == 16 ? 1 : 2;
      const int64_t lim = (1LL << (31 - (cls + 3))) - 1;
      keyed[k] = kp > 0 && (int64_t)std::max(amax[k], 0) + std::max(bmax, 0) < lim;
      if (keyed[k]) {
        kd[cls].push_back(PathKeyDesc{k, kunits[cls], cblocks, 0, d_akey + akey_off[k]});
        kunits[cls] += (inner + kMkRows - 1) / kMkRows * cblocks;
        kmax[cls] = std::max(kmax[cls], K);
        rows[cls] += inner;
      } else {
        rows[3] += inner;
        const int rbk = (inner + kMpRows - 1) / kMpRows;
        md.push_back(PathMergeDesc{k, units, rbk});
        units += rbk * ((n + kMpCols - 1) / kMpCols);
      }
      FAP_REQUIRE(units < (1 << 30) && kunits[cls] < (1 << 30),
                  "fap_hip_solve_parts_path: too many merge units");
    }
    // Descriptor lists: compare form at d_md, key classes after one another at d_kd.
    const size_t nk = kd[0].size() + kd[1].size() + kd[2].size();
    std::vector<char> desc(sizeof(PathMergeDesc) * md.size() + sizeof(PathKeyDesc) * nk);
    if (!md.empty()) std::memcpy(desc.data(), md.data(), sizeof(PathMergeDesc) * md.size());
    char *kdst = desc.data() + sizeof(PathMergeDesc) * md.size();
    for (int c = 0; c < 3; ++c) {
      if (kd[c].empty()) continue;
      std::memcpy(kdst, kd[c].data(), sizeof(PathKeyDesc) * kd[c].size());
      kdst += sizeof(PathKeyDesc) * kd[c].size();
    }
    if (!md.empty())
      FAP_CHECK_HIP(hipMemcpyAsync(d_md, desc.data(), sizeof(PathMergeDesc) * md.size(),
                                   hipMemcpyHostToDevice, st));
    if (nk > 0)
      FAP_CHECK_HIP(hipMemcpyAsync(d_kd, desc.data() + sizeof(PathMergeDesc) * md.size(),
                                   sizeof(PathKeyDesc) * nk, hipMemcpyHostToDevice, st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    merge_keyed = (int)nk;
    // The launches (key classes, compare form) write disjoint rows: they run side by
    // side on up to three streams, largest first, so that one's tail overlaps another.
    struct Launch {
      int cls;  // 0..2 key class, 3 compare form
      const PathKeyDesc *kd;
      int64_t work;
    };
    std::vector<Launch> ls;
    const PathKeyDesc *kbase = d_kd;
    for (int c = 0; c < 3; ++c) {
      if (kunits[c] > 0) ls.push_back(Launch{c, kbase, rows[c]});
      kbase += kd[c].size();
    }
    if (units > 0) ls.push_back(Launch{3, nullptr, rows[3] * 2});  // slower per row
    std::sort(ls.begin(), ls.end(), [](const Launch &x, const Launch &y) { return x.work > y.work; });
    const hipStream_t lanes[3] = {st, aux_stream(0), aux_stream(1)};
    const int nl = std::min<int>(3, (int)ls.size());
    if (nl > 1) {
      hipEvent_t fork;
      FAP_TRY(timing_record(st, &fork));
      for (int l = 1; l < nl; ++l) FAP_CHECK_HIP(hipStreamWaitEvent(lanes[l], fork, 0));
    }
    for (size_t q = 0; q < ls.size(); ++q) {
      const hipStream_t on = lanes[q % nl];
      const int c = ls[q].cls;
      if (c < 3) {
        const size_t lds = (size_t)kmax[c] * 4 * 64 * sizeof(int);
        auto kern = c == 0 ? merge_path_key_kernel<8, 3>
                    : c == 1 ? merge_path_key_kernel<16, 4> : merge_path_key_kernel<32, 5>;
        hipLaunchKernelGGL(kern, dim3(kunits[c]), dim3(64 * kMkWaves), lds, on, d_pp, ls[q].kd,
                           (int)kd[c].size(), kunits[c], n, d_out, d_path, ld_out);
      } else {
        hipLaunchKernelGGL(merge_path_kernel, dim3(units), dim3(64), 0, on, d_pp, d_md,
                           (int)md.size(), units, n, g->d_graph_id, g->d_ed2st, d_out, d_path,
                           ld_out);
      }
      FAP_CHECK_LAUNCH();
    }
    for (int l = 1; l < nl; ++l) {
      hipEvent_t joined;
      FAP_TRY(timing_record(lanes[l], &joined));
      FAP_CHECK_HIP(hipStreamWaitEvent(st, joined, 0));
    }
    return FAP_OK;
  }));
  // 5. own columns of the boundary rows, and of every row of a keyed part.
  std::vector<int64_t> dec_begin(P);
  int64_t dec_total = 0;
  for (int k = 0; k < P; ++k) {
    dec_begin[k] = dec_total;
    dec_total += keyed[k] ? act[k].num : act[k].bdy;
  }
  FAP_TRY(span(4, [&]() -> int {
    if (dec_total == 0) return FAP_OK;
    FAP_CHECK_HIP(hipMemcpyAsync(d_db, dec_begin.data(), 8LL * P, hipMemcpyHostToDevice, st));
    FAP_CHECK_HIP(hipStreamSynchronize(st));
    hipLaunchKernelGGL(decode_path_kernel, dim3((unsigned)dec_total), dim3(256), 0, st, d_pp, P,
                       d_db, dec_total, g->d_st2ed, d_out, d_path, ld_out);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }));
  timing_note_keyed(merge_keyed, fw_keyed);
  if (bgraph) timing_note_bgraph();
  return timing_finish(st);
}

extern "C" int fap_hip_solve_parts_path(const fap_graph_dev *g, const int32_t *d_rrow_offset,
                                        const int32_t *d_rcol, const int32_t *d_rw,
                                        const fap_part_desc *parts, int32_t nparts,
                                        int32_t *d_out, int32_t *d_path, int64_t ld_out,
                                        void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && nparts >= 0 && (nparts == 0 || parts) && d_out && d_path,
              "fap_hip_solve_parts_path: null argument");
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  const int s = solve_parts_path(g, d_rrow_offset, d_rcol, d_rw, parts, nparts, d_out, d_path,
                                 ld_out, st);
  const int f = ws_fence_record(st);
  return s != FAP_OK ? s : f;
}

// ---- kernel-level seam with predecessors ---------------------------------------
// The reference's seam functions carry path matrices beside the distances:
// floyd_path (fap/kernel/floyd.h:162-170), min_plus_path_advanced
// (fap/kernel/minplus.h:93-105) and handle_boundry_path (fap/kernel/batch_sssp.h:
// 280-303).  These entry points give the same (distance, path) results as their CPU
// branches, using the kernels of the path pipeline above.
namespace fap_hip {
namespace {

Workspace g_ws_seam_fw;    // padded distance + path matrices of fap_hip_floyd_path
Workspace g_ws_seam_snap;  // its per-round records
Workspace g_ws_seam_meta;  // part record + flag; predecessor statistics

__global__ void pad_in_path_kernel(int *__restrict__ dd, int *__restrict__ dp, int S,
                                   const int *__restrict__ sd, const int *__restrict__ sp,
                                   int64_t ld, int s) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)S * S) return;
  const int i = (int)(idx / S), j = (int)(idx - (int64_t)i * S);
  const bool in = i < s && j < s;
  dd[idx] = in ? sd[(int64_t)i * ld + j] : kInf;
  dp[idx] = in ? sp[(int64_t)i * ld + j] : -1;
}

__global__ void pad_out_path_kernel(int *__restrict__ sd, int *__restrict__ sp, int64_t ld, int s,
                                    const int *__restrict__ dd, const int *__restrict__ dp, int S) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)s * s) return;
  const int i = (int)(idx / s), j = (int)(idx - (int64_t)i * s);
  sd[(int64_t)i * ld + j] = dd[(int64_t)i * S + j];
  sp[(int64_t)i * ld + j] = dp[(int64_t)i * S + j];
}

// C = A (x) B with paths: one 64 x 64 output tile per 256-thread workgroup, K in LDS
// chunks of 16 taken in ascending order, so each thread's strict '<' keeps the
// smallest k reaching the minimum (min_plus_path_CPU_advanced, minplus.h:51-64).
// Thread (ty, tx) owns rows ty*4 .. ty*4+3 and columns tx, tx+16, tx+32, tx+48.
constexpr int kSpT = 64, kSpKc = 16;
__global__ __launch_bounds__(256) void minplus_path_seam_kernel(
    const int *__restrict__ A, int64_t lda, const int *__restrict__ B, int64_t ldb,
    const int *__restrict__ Bp, int64_t ldbp, int *__restrict__ C, int64_t ldc,
    int *__restrict__ Cp, int64_t ldcp, int M, int N, int K) {
  __shared__ int As[kSpKc][kSpT];
  __shared__ int Bs[kSpKc][kSpT];
  __shared__ int Bps[kSpKc][kSpT];
  const int i0 = blockIdx.y * kSpT, j0 = blockIdx.x * kSpT;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  int best[4][4], bp[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      best[q][c] = kInf;
      bp[q][c] = -1;
    }
  for (int k0 = 0; k0 < K; k0 += kSpKc) {
    for (int e = tid; e < kSpKc * kSpT; e += 256) {
      const int r = e / kSpKc, kk = e - r * kSpKc;  // A: row-major reads
      As[kk][r] = (i0 + r < M && k0 + kk < K) ? A[(int64_t)(i0 + r) * lda + k0 + kk] : kInf;
      const int kb = e / kSpT, c = e - kb * kSpT;  // B: coalesced rows
      const bool ok = k0 + kb < K && j0 + c < N;
      Bs[kb][c] = ok ? B[(int64_t)(k0 + kb) * ldb + j0 + c] : kInf;
      Bps[kb][c] = ok ? Bp[(int64_t)(k0 + kb) * ldbp + j0 + c] : -1;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < kSpKc; ++kk) {
      int a[4], b[4], pb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) a[q] = As[kk][ty * 4 + q];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        b[c] = Bs[kk][tx + 16 * c];
        pb[c] = Bps[kk][tx + 16 * c];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int cand = a[q] + b[c];
          if (cand < best[q][c]) {
            best[q][c] = cand;
            bp[q][c] = pb[c];
          }
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = i0 + ty * 4 + q;
    if (i >= M) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + tx + 16 * c;
      if (j >= N) continue;
      C[(int64_t)i * ldc + j] = best[q][c];
      Cp[(int64_t)i * ldcp + j] = bp[q][c];
    }
  }
}

}  // namespace
}  // namespace fap_hip

extern "C" int fap_hip_floyd_path(int32_t *d_mat, int32_t *d_path, int32_t s, int64_t ld,
                                  void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(s >= 0 && ld >= s, "fap_hip_floyd_path: bad sizes s=%d ld=%lld", s, (long long)ld);
  if (s == 0) return FAP_OK;
  FAP_REQUIRE(d_mat && d_path, "fap_hip_floyd_path: null matrix");
  FAP_REQUIRE(s <= kMaxPart, "fap_hip_floyd_path: s=%d above the limit %d", s, kMaxPart);
  hipStream_t st = resolve(stream);
  const int S = (int)align_up(s, kT), nb = S / kT;
  const int64_t tot = (int64_t)S * S;
  FAP_TRY(g_ws_seam_fw.ensure(tot * 8));
  FAP_TRY(g_ws_seam_snap.ensure((int64_t)(3 + 3 * nb) * kTT * 4));
  FAP_TRY(g_ws_seam_meta.ensure(512));
  int *dd = static_cast<int *>(g_ws_seam_fw.ptr), *dp = dd + tot;
  char *meta = static_cast<char *>(g_ws_seam_meta.ptr);
  PathPart *d_pp = reinterpret_cast<PathPart *>(meta);
  int *d_fwkey = reinterpret_cast<int *>(meta + 256);  // 0: compare-and-select form
  const PathPart h{dd, dp, static_cast<int *>(g_ws_seam_snap.ptr), S, nb, s, 0, 0, 1, 0, 0};
  FAP_CHECK_HIP(hipMemcpyAsync(d_pp, &h, sizeof(h), hipMemcpyHostToDevice, st));
  FAP_TRY(fap_hip_fill_i32(d_fwkey, 0, 1, st));
  hipLaunchKernelGGL(pad_in_path_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     dd, dp, S, d_mat, d_path, ld, s);
  FAP_CHECK_LAUNCH();
  for (int r = 0; r < nb; ++r) {
    hipLaunchKernelGGL(fwp_phase1_kernel, dim3(1), dim3(256), 0, st, d_pp, d_fwkey, r);
    FAP_CHECK_LAUNCH();
    if (nb == 1) continue;
    hipLaunchKernelGGL(fwp_phase2_kernel, dim3(2 * (nb - 1), 1), dim3(256), 0, st, d_pp, d_fwkey,
                       r);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(fwp_phase3_kernel, dim3((nb - 1) * (nb - 1), 1), dim3(256), 0, st, d_pp,
                       d_fwkey, r);
    FAP_CHECK_LAUNCH();
  }
  const int64_t out = (int64_t)s * s;
  hipLaunchKernelGGL(pad_out_path_kernel, dim3((unsigned)((out + 255) / 256)), dim3(256), 0, st,
                     d_mat, d_path, ld, s, dd, dp, S);
  FAP_CHECK_LAUNCH();
  // The workspaces are shared between calls: finish before they can be reused.
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}

extern "C" int fap_hip_minplus_path(const int32_t *d_A, int64_t lda, const int32_t *d_B,
                                    int64_t ldb, const int32_t *d_Bpath, int64_t ldbp,
                                    int32_t *d_C, int64_t ldc, int32_t *d_Cpath, int64_t ldcp,
                                    int32_t M, int32_t N, int32_t K, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(M >= 0 && N >= 0 && K >= 0, "fap_hip_minplus_path: negative size");
  if (M == 0 || N == 0) return FAP_OK;
  FAP_REQUIRE(d_C && d_Cpath && ldc >= N && ldcp >= N, "fap_hip_minplus_path: bad C");
  FAP_REQUIRE(K == 0 || (d_A && d_B && d_Bpath && lda >= K && ldb >= N && ldbp >= N),
              "fap_hip_minplus_path: bad A/B");
  const int64_t gy = (M + kSpT - 1) / kSpT, gx = (N + kSpT - 1) / kSpT;
  FAP_REQUIRE(gy <= 65535 && gx < (1LL << 31), "fap_hip_minplus_path: product too large");
  hipStream_t st = resolve(stream);
  hipLaunchKernelGGL(minplus_path_seam_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st,
                     d_A, lda, d_B, ldb, d_Bpath, ldbp, d_C, ldc, d_Cpath, ldcp, M, N, K);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

extern "C" int fap_hip_sssp_batch_path(const int32_t *d_row_offset, const int32_t *d_col,
                                       const int32_t *d_w, const int32_t *d_rrow_offset,
                                       const int32_t *d_rcol, const int32_t *d_rw, int32_t n,
                                       int64_t m, const int32_t *d_src, int32_t nsrc,
                                       int32_t *d_out, int32_t *d_path, int64_t ld_out,
                                       void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(n >= 0 && m >= 0 && nsrc >= 0 && ld_out >= n,
              "fap_hip_sssp_batch_path: bad sizes n=%d m=%lld nsrc=%d ld=%lld", n, (long long)m,
              nsrc, (long long)ld_out);
  if (nsrc == 0 || n == 0) return FAP_OK;
  FAP_REQUIRE(d_path && d_rrow_offset && (m == 0 || d_rcol) && (d_rw != nullptr) == (d_w != nullptr),
              "fap_hip_sssp_batch_path: null pointer (reverse CSR, or d_rw without d_w)");
  hipStream_t st = resolve(stream);
  FAP_TRY(fap_hip_sssp_batch(d_row_offset, d_col, d_w, n, m, d_src, nsrc, d_out, ld_out, stream));
  FAP_TRY(g_ws_seam_meta.ensure(512 + 4LL * kMaxSlots * kSlotStride));
  int *stat = reinterpret_cast<int *>(static_cast<char *>(g_ws_seam_meta.ptr) + 512);
  for (int k0 = 0; k0 < nsrc; k0 += 65535 * kPredRows) {
    const int nk = std::min(65535 * kPredRows, nsrc - k0);
    hipLaunchKernelGGL(pred_rows_kernel, dim3((n + 255) / 256, (nk + kPredRows - 1) / kPredRows),
                       dim3(256), 0, st, d_rrow_offset, d_rcol, d_rw, n, d_src,
                       (const int64_t *)nullptr, ld_out, k0, nk, d_out, d_path, stat);
    FAP_CHECK_LAUNCH();
  }
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}
