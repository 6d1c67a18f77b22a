// In-part all-pairs distances of unit-weight symmetric graphs by bit-parallel BFS.
//
// The boundary-graph form of the part pipeline needs, for every part, the distances
// inside the part's induced subgraph (fap/fap.cpp:195-503 gets them from
// floyd_CPU, fap/kernel/floyd.h, on the part's sub-matrix).  On unit weights with
// every arc paired with its reverse, these are BFS distances, and 64 BFS runs of one
// part advance together as the bits of one 64-bit word per vertex: a vertex's next
// frontier word is the OR of its in-part neighbours' frontier words minus the sources
// that already reached it.  One level costs about deg(v) LDS reads per vertex for 64
// sources, against s * 64 min-plus steps per vertex for the same 64 matrix columns
// in Floyd-Warshall.
//
// Workgroup (256 threads): one part and one group of 64 consecutive sorted sources
// [g * 64, g * 64 + 64).  Thread t owns vertices t + 256 q (q < Q): their visited
// words and in-part neighbour lists (up to kDeg local ids, the rest re-read from the
// CSR each level) stay in registers; the frontier words (s x 8 B) and the levels
// (s x 64 bytes, 0xFF = not reached, 0xFE = level above 253 stored directly) sit in
// LDS.  By symmetry d(src, v) = d(v, src), so the group fills columns g*64 .. g*64+63
// of every row of the part's S x S int32 matrix: rows are written as 256 contiguous
// bytes at the end (FAP_INF where unreached, for source columns >= s and for the
// padding rows >= s).
#include <cstdint>

#include "fap_internal.h"
#include "fw_desc.h"

namespace fap_hip {
namespace {

constexpr int kDeg = 8;            // in-part neighbours kept in registers per vertex
constexpr int kMaxByteLevel = 253;  // levels above are written to global directly
constexpr uint8_t kUnreached = 0xFF, kDirect = 0xFE;

template <int Q>
__global__ __launch_bounds__(256) void inpart_bfs_kernel(const InpartDesc *__restrict__ descs,
                                                         int nd, int total_groups,
                                                         const int *__restrict__ row_off,
                                                         const int *__restrict__ col,
                                                         const int *__restrict__ graph_id,
                                                         const int *__restrict__ st2ed,
                                                         const int *__restrict__ ed2st) {
  extern __shared__ unsigned long long lds_words[];
  const int blk = blockIdx.x;
  if (blk >= total_groups) return;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].group_begin <= blk) lo = mid;
    else hi = mid - 1;
  }
  const InpartDesc d = descs[lo];
  const int src0 = (blk - d.group_begin) * 64;
  const int s = d.s, tid = threadIdx.x;
  unsigned long long *F = lds_words;                        // frontier words, s
  uint8_t *lev = reinterpret_cast<uint8_t *>(lds_words + s);  // levels, s x 64

  // Neighbour lists, visited words, first frontier.
  int nbr[Q][kDeg], deg[Q], ovf[Q];  // ovf: CSR index of the first neighbour not kept
  unsigned long long vis[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int v = tid + 256 * q;
    deg[q] = 0;
    ovf[q] = -1;
    vis[q] = 0;
    if (v < s) {
      const int ver = st2ed[d.offset + v];
      const int e1 = row_off[ver + 1];
      for (int e = row_off[ver]; e < e1; ++e) {
        const int nb = col[e];
        if (graph_id[nb] != d.part_id) continue;
        if (deg[q] == kDeg) {
          ovf[q] = e;
          break;
        }
        // A fixed register index per unrolled slot: no dynamic register indexing.
#pragma unroll
        for (int k = 0; k < kDeg; ++k)
          if (k == deg[q]) nbr[q][k] = ed2st[nb] - d.offset;
        ++deg[q];
      }
      const int b = v - src0;
      vis[q] = (b >= 0 && b < 64) ? 1ull << b : 0ull;
      F[v] = vis[q];
      uint32_t *row = reinterpret_cast<uint32_t *>(lev + (int64_t)v * 64);
#pragma unroll
      for (int w = 0; w < 16; ++w) row[w] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int v = tid + 256 * q, b = v - src0;
    if (v < s && b >= 0 && b < 64) lev[(int64_t)v * 64 + b] = 0;
  }

  for (int level = 1;; ++level) {
    unsigned long long nx[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      unsigned long long acc = 0;
#pragma unroll
      for (int k = 0; k < kDeg; ++k)
        if (k < deg[q]) acc |= F[nbr[q][k]];
      if (ovf[q] >= 0) {  // more than kDeg in-part neighbours: the rest from the CSR
        const int v = tid + 256 * q;
        const int ver = st2ed[d.offset + v];
        for (int e = ovf[q]; e < row_off[ver + 1]; ++e) {
          const int nb = col[e];
          if (graph_id[nb] == d.part_id) acc |= F[ed2st[nb] - d.offset];
        }
      }
      nx[q] = acc & ~vis[q];
    }
    __syncthreads();  // every read of this level's frontier is done
    int any = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int v = tid + 256 * q;
      if (v >= s) continue;
      F[v] = nx[q];
      vis[q] |= nx[q];
      unsigned long long m = nx[q];
      any |= m != 0;
      while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        if (level <= kMaxByteLevel) {
          lev[(int64_t)v * 64 + b] = (uint8_t)level;
        } else {
          lev[(int64_t)v * 64 + b] = kDirect;
          d.fw[(int64_t)v * d.S + src0 + b] = level;
        }
      }
    }
    if (!__syncthreads_or(any)) break;
  }

  // Write columns src0 .. src0 + 63 of every row: 16 lanes per row, 4 columns each.
  for (int r = tid >> 4; r < d.S; r += 16) {
    const int c = (tid & 15) * 4;
    int out[4];
    bool direct = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint8_t x = r < s ? lev[(int64_t)r * 64 + c + e] : kUnreached;
      direct |= x == kDirect;
      out[e] = (x == kUnreached || src0 + c + e >= s) ? kInf : (int)x;
    }
    int *dst = d.fw + (int64_t)r * d.S + src0 + c;
    if (!direct) {
      *reinterpret_cast<int4 *>(dst) = make_int4(out[0], out[1], out[2], out[3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (lev[(int64_t)r * 64 + c + e] != kDirect) dst[e] = out[e];
    }
  }
}

}  // namespace

int inpart_bfs_lds_bytes(int s_max) { return s_max * 8 + s_max * 64; }

int launch_inpart_bfs(const InpartDesc *d_descs, int nd, int total_groups, int s_max,
                      const int *row_off, const int *col, const int *graph_id, const int *st2ed,
                      const int *ed2st, hipStream_t stream) {
  if (nd <= 0 || total_groups <= 0) return FAP_OK;
  FAP_REQUIRE(s_max <= kInpartBfsMaxS, "launch_inpart_bfs: part of %d vertices", s_max);
  const size_t lds = (size_t)inpart_bfs_lds_bytes(s_max);
  if (s_max <= 1024)
    hipLaunchKernelGGL(inpart_bfs_kernel<4>, dim3(total_groups), dim3(256), lds, stream, d_descs,
                       nd, total_groups, row_off, col, graph_id, st2ed, ed2st);
  else
    hipLaunchKernelGGL(inpart_bfs_kernel<8>, dim3(total_groups), dim3(256), lds, stream, d_descs,
                       nd, total_groups, row_off, col, graph_id, st2ed, ed2st);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

}  // namespace fap_hip
