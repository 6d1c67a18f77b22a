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
constexpr int kMaxS = 8192;

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
// Boundary row k = k0 + blockIdx.y (source src[k], row offset rows[k] in d_out and
// d_path).
__global__ void pred_rows_kernel(const int *__restrict__ rro, const int *__restrict__ rcol,
                                 const int *__restrict__ rw, int n, const int *__restrict__ src,
                                 const int64_t *__restrict__ rows, int k0,
                                 const int *__restrict__ out, int *__restrict__ path) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int k = k0 + blockIdx.y;
  const int s = src[k];
  const int *d = out + rows[k];
  const int dv = d[v];
  int p = -1;
  if (v == s) {
    p = s;
  } else if (dv < kInf) {
    int bd = kInf, bu = 0x7FFFFFFF;
    for (int e = rro[v]; e < rro[v + 1]; ++e) {
      const int u = rcol[e];
      const int du = d[u];
      if (du >= kInf || du + (rw ? rw[e] : 1) != dv) continue;
      if (du > dv || (du == dv && u >= v)) continue;  // settled after v
      if (du < bd || (du == bd && u < bu)) {
        bd = du;
        bu = u;
      }
    }
    p = bu == 0x7FFFFFFF ? -1 : bu;
  }
  path[rows[k] + v] = p;
}

// ---- build ----------------------------------------------------------------------
// One workgroup per padded row of every part (rows of PathPart::row_begin prefix).
__global__ __launch_bounds__(256) void build_path_kernel(
    const PathPart *__restrict__ parts, int np, int total_rows, const int *__restrict__ row_off,
    const int *__restrict__ col, const int *__restrict__ w, const int *__restrict__ graph_id,
    const int *__restrict__ st2ed, const int *__restrict__ ed2st, const int *__restrict__ out,
    const int *__restrict__ pout, int64_t ld_out) {
  const int grow = blockIdx.x;
  if (grow >= total_rows) return;
  const PathPart d = parts[find_part(parts, np, grow, &PathPart::row_begin)];
  const int i = grow - d.row_begin;
  __shared__ int rd[kMaxS];
  __shared__ int rp[kMaxS];
  const int tid = threadIdx.x;
  if (i < d.s && i < d.bdy) {
    const int64_t base = (d.out_row + i) * ld_out;
    for (int j = tid; j < d.S; j += 256) {
      const int v = j < d.s ? st2ed[d.offset + j] : 0;
      rd[j] = j < d.s ? out[base + v] : kInf;
      rp[j] = j < d.s ? pout[base + v] : -1;
    }
  } else {
    for (int j = tid; j < d.S; j += 256) {
      rd[j] = kInf;
      rp[j] = -1;
    }
  }
  __syncthreads();
  if (tid == 0 && i < d.s) {
    const int ver = st2ed[d.offset + i];
    if (i >= d.bdy) {  // boundary rows: the SSSP row replaces the arcs
      for (int e = row_off[ver]; e < row_off[ver + 1]; ++e) {
        const int nb = col[e];
        if (graph_id[nb] != d.part_id) continue;
        const int idx = ed2st[nb] - d.offset;
        if (idx >= 0 && idx < d.s) {
          rd[idx] = w ? w[e] : 1;
          rp[idx] = ver;
        }
      }
    }
    rd[i] = 0;
    rp[i] = ver;
  }
  __syncthreads();
  int *dd = d.fw + (int64_t)i * d.S;
  int *dp = d.fwp + (int64_t)i * d.S;
  for (int j = tid; j < d.S; j += 256) {
    dd[j] = rd[j];
    dp[j] = rp[j];
  }
}

// ---- Floyd-Warshall with k-i-j path semantics -----------------------------------
// 256 threads per 64 x 64 tile: thread t owns row t / 4, columns (t % 4) * 16 + 0..15.
__device__ __forceinline__ int *tile(int *m, const PathPart &p, int ti, int tj) {
  return m + (int64_t)(ti * kT) * p.S + tj * kT;
}

__device__ __forceinline__ void tile_to_lds(int (*D)[kT], int (*P)[kT], const int *gd,
                                            const int *gp, int64_t ld) {
  const int r = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    D[r][c0 + c] = gd[(int64_t)r * ld + c0 + c];
    P[r][c0 + c] = gp[(int64_t)r * ld + c0 + c];
  }
}

__device__ __forceinline__ void lds_to_tile(int (*D)[kT], int (*P)[kT], int *gd, int *gp,
                                            int64_t ld) {
  const int r = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    gd[(int64_t)r * ld + c0 + c] = D[r][c0 + c];
    gp[(int64_t)r * ld + c0 + c] = P[r][c0 + c];
  }
}

// Phase 1: the diagonal tile of round r, sequentially over its 64 pivots, recording
// row k (dist, path) and column k as they stand at step k.
__global__ __launch_bounds__(256) void fwp_phase1_kernel(const PathPart *__restrict__ parts,
                                                         int r) {
  const PathPart p = parts[blockIdx.x];
  if (r >= p.nb) return;
  __shared__ int D[kT][kT];
  __shared__ int P[kT][kT];
  tile_to_lds(D, P, tile(p.fw, p, r, r), tile(p.fwp, p, r, r), p.S);
  int *rrow = rec_diag_row(p), *rrowp = rec_diag_rowp(p), *rcol = rec_diag_col(p);
  const int i = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  for (int k = 0; k < kT; ++k) {
    __syncthreads();
    if (i == k) {
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        rrow[k * kT + c0 + c] = D[k][c0 + c];
        rrowp[k * kT + c0 + c] = P[k][c0 + c];
      }
    }
    if (k >= c0 && k < c0 + 16) rcol[k * kT + i] = D[i][k];  // stored k-major
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
                                                         int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb) return;
  const int others = p.nb - 1;
  int t = blockIdx.x;
  if (t >= 2 * others) return;
  const bool colp = t < others;
  if (!colp) t -= others;
  const int o = t < r ? t : t + 1;  // tile index outside the pivot
  __shared__ int D[kT][kT];
  __shared__ int P[kT][kT];
  __shared__ int R0[kT][kT];  // column panel: diag row records; row panel: diag column records
  __shared__ int R1[kT][kT];  // column panel: diag row path records
  int *gd = colp ? tile(p.fw, p, o, r) : tile(p.fw, p, r, o);
  int *gp = colp ? tile(p.fwp, p, o, r) : tile(p.fwp, p, r, o);
  tile_to_lds(D, P, gd, gp, p.S);
  {
    const int *a = colp ? rec_diag_row(p) : rec_diag_col(p);
    const int *b = rec_diag_rowp(p);
    for (int e = threadIdx.x; e < kTT; e += 256) {
      R0[e / kT][e % kT] = a[e];
      if (colp) R1[e / kT][e % kT] = b[e];
    }
  }
  int *rec = colp ? rec_colpanel(p, o) : rec_rowpanel(p, o);
  int *recp = colp ? nullptr : rec_rowpanelp(p, o);
  const int i = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
  for (int k = 0; k < kT; ++k) {
    __syncthreads();
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
                                                         int r) {
  const PathPart p = parts[blockIdx.y];
  if (r >= p.nb) return;
  const int others = p.nb - 1;
  const int t = blockIdx.x;
  if (t >= others * others) return;
  const int a0 = t / others, b0 = t - a0 * others;
  const int a = a0 < r ? a0 : a0 + 1, b = b0 < r ? b0 : b0 + 1;
  __shared__ int As[kT][kT];  // As[k][i]: column panel record of tile (a, r)
  __shared__ int Bs[kT][kT];  // Bs[k][j]: row panel record of tile (r, b)
  __shared__ int Bp[kT][kT];
  {
    const int *ra = rec_colpanel(p, a), *rb = rec_rowpanel(p, b), *rbp = rec_rowpanelp(p, b);
    for (int e = threadIdx.x; e < kTT; e += 256) {
      As[e / kT][e % kT] = ra[e];
      Bs[e / kT][e % kT] = rb[e];
      Bp[e / kT][e % kT] = rbp[e];
    }
  }
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  int *gd = tile(p.fw, p, a, b), *gp = tile(p.fwp, p, a, b);
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
      const int *arow = p.fw + (int64_t)(p.bdy + i) * p.S + k0;  // A = FW[bdy + i][k]
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
        dv = p.fw[(int64_t)(p.bdy + i) * p.S + own[e]];
        pv = p.fwp[(int64_t)(p.bdy + i) * p.S + own[e]];
      }
      out[orow + j] = dv;
      pout[orow + j] = pv;
    }
  }
}

// Boundary rows' own columns: out[row][st2ed[offset + c]] = FW[row][c] (dist, path).
__global__ void decode_path_kernel(const PathPart *__restrict__ parts, int np,
                                   const int64_t *__restrict__ item_begin, int64_t total,
                                   const int *__restrict__ st2ed, int *__restrict__ out,
                                   int *__restrict__ pout, int64_t ld_out) {
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= total) return;
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item_begin[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  const PathPart p = parts[lo];
  const int64_t local = item - item_begin[lo];
  const int i = (int)(local / p.s), c = (int)(local - (int64_t)i * p.s);
  const int64_t dst = (p.out_row + i) * ld_out + st2ed[p.offset + c];
  out[dst] = p.fw[(int64_t)i * p.S + c];
  pout[dst] = p.fwp[(int64_t)i * p.S + c];
}

Workspace g_ws_pfw;    // FW distance and path matrices
Workspace g_ws_psnap;  // per-part records
Workspace g_ws_pmeta;  // descriptors, sources, row offsets
Workspace g_ws_pell;   // ELL adjacency (unit weights)

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
                    (int64_t)q.offset + q.num <= n && q.out_row >= 0 && q.num <= kMaxS,
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
  std::vector<int64_t> out_rows(P), dec_begin(P);
  std::vector<PathMergeDesc> md;
  int units = 0;
  int64_t dec_total = 0;
  for (int k = 0; k < P; ++k) {
    const fap_part_desc &q = act[k];
    pp[k] = PathPart{fw_base + fw_off[k], fwp_base + fw_off[k], snap_base + snap_off[k], S[k],
                     nb[k], q.num, q.bdy, q.offset, q.part_id, rows_total, q.out_row};
    rows_total += S[k];
    offsets[k] = q.offset;
    src_begin[k] = src_total;
    out_rows[k] = q.out_row;
    src_total += q.bdy;
    const int inner = q.num - q.bdy;
    if (inner > 0) {
      const int rbk = (inner + kMpRows - 1) / kMpRows;
      md.push_back(PathMergeDesc{k, units, rbk});
      units += rbk * ((n + kMpCols - 1) / kMpCols);
      FAP_REQUIRE(units < (1 << 30), "fap_hip_solve_parts_path: too many merge units");
    }
    dec_begin[k] = dec_total;
    dec_total += (int64_t)q.bdy * q.num;
  }

  int64_t off = 0;
  auto carve = [&](int64_t bytes) {
    const int64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const int64_t o_pp = carve(sizeof(PathPart) * P);
  const int64_t o_md = carve(sizeof(PathMergeDesc) * std::max<size_t>(md.size(), 1));
  const int64_t o_off = carve(4LL * P), o_sb = carve(4LL * P), o_or = carve(8LL * P);
  const int64_t o_db = carve(8LL * P);
  const int64_t upload = off;
  const int64_t o_src = carve(4LL * std::max(src_total, 1));
  const int64_t o_rows = carve(8LL * std::max(src_total, 1));
  FAP_TRY(g_ws_pmeta.ensure(off));
  char *meta = static_cast<char *>(g_ws_pmeta.ptr);
  std::vector<char> host(upload);
  std::memcpy(host.data() + o_pp, pp.data(), sizeof(PathPart) * P);
  if (!md.empty()) std::memcpy(host.data() + o_md, md.data(), sizeof(PathMergeDesc) * md.size());
  std::memcpy(host.data() + o_off, offsets.data(), 4LL * P);
  std::memcpy(host.data() + o_sb, src_begin.data(), 4LL * P);
  std::memcpy(host.data() + o_or, out_rows.data(), 8LL * P);
  std::memcpy(host.data() + o_db, dec_begin.data(), 8LL * P);
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host.data(), upload, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  const PathPart *d_pp = reinterpret_cast<const PathPart *>(meta + o_pp);
  const PathMergeDesc *d_md = reinterpret_cast<const PathMergeDesc *>(meta + o_md);
  const int *d_offsets = reinterpret_cast<const int *>(meta + o_off);
  const int *d_sb = reinterpret_cast<const int *>(meta + o_sb);
  const int64_t *d_or = reinterpret_cast<const int64_t *>(meta + o_or);
  const int64_t *d_db = reinterpret_cast<const int64_t *>(meta + o_db);
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
  // 1. boundary rows: distances by SSSP, then their predecessors.
  FAP_TRY(span(0, [&]() -> int {
    if (src_total == 0) return FAP_OK;
    FAP_TRY(launch_gather_sources(d_offsets, d_sb, d_or, P, src_total, g->d_st2ed, ld_out, d_src,
                                  d_rows, st));
    int s = FAP_ERR_INVALID;
    if (!g->d_w) {
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
    for (int k0 = 0; k0 < src_total; k0 += 65535) {
      const int rows = std::min(65535, src_total - k0);
      hipLaunchKernelGGL(pred_rows_kernel, dim3((n + 255) / 256, rows), dim3(256), 0, st, d_rro,
                         d_rcol, d_rw, n, d_src, d_rows, k0, d_out, d_path);
      FAP_CHECK_LAUNCH();
    }
    return FAP_OK;
  }));
  // 2. sub-matrix build with paths.
  FAP_TRY(span(1, [&]() -> int {
    hipLaunchKernelGGL(build_path_kernel, dim3(rows_total), dim3(256), 0, st, d_pp, P, rows_total,
                       g->d_row_offset, g->d_col, g->d_w, g->d_graph_id, g->d_st2ed, g->d_ed2st,
                       d_out, d_path, ld_out);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }));
  // 3. Floyd-Warshall with k-i-j path semantics, all parts per launch.
  FAP_TRY(span(2, [&]() -> int {
    for (int r = 0; r < max_nb; ++r) {
      hipLaunchKernelGGL(fwp_phase1_kernel, dim3(P), dim3(256), 0, st, d_pp, r);
      FAP_CHECK_LAUNCH();
      if (max_nb > 1) {
        hipLaunchKernelGGL(fwp_phase2_kernel, dim3(2 * (max_nb - 1), P), dim3(256), 0, st, d_pp,
                           r);
        FAP_CHECK_LAUNCH();
        hipLaunchKernelGGL(fwp_phase3_kernel, dim3((max_nb - 1) * (max_nb - 1), P), dim3(256), 0,
                           st, d_pp, r);
        FAP_CHECK_LAUNCH();
      }
    }
    return FAP_OK;
  }));
  // 4. min-plus of the inner rows with paths, own columns decoded.
  FAP_TRY(span(3, [&]() -> int {
    if (units == 0) return FAP_OK;
    hipLaunchKernelGGL(merge_path_kernel, dim3(units), dim3(64), 0, st, d_pp, d_md,
                       (int)md.size(), units, n, g->d_graph_id, g->d_ed2st, d_out, d_path, ld_out);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }));
  // 5. boundary rows' own columns.
  FAP_TRY(span(4, [&]() -> int {
    if (dec_total == 0) return FAP_OK;
    hipLaunchKernelGGL(decode_path_kernel, dim3((unsigned)((dec_total + 255) / 256)), dim3(256),
                       0, st, d_pp, P, d_db, dec_total, g->d_st2ed, d_out, d_path, ld_out);
    FAP_CHECK_LAUNCH();
    return FAP_OK;
  }));
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
