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
// Workgroup (kThreads = 512): one part and one group of 64 consecutive sorted sources
// [g * 64, g * 64 + 64).  Thread t owns vertices t + 512 q (q < Q): their visited
// words, in-part neighbour lists (up to kDeg local ids, the rest re-read from the CSR
// each level) and level bit-planes stay in registers; only the frontier words
// (two buffers of 512 Q + 1 words at a fixed distance, each with a zero word at
// index s that pads the neighbour lists) sit in LDS.
// Each level reads the padded lists up to the wave's largest degree with no per-lane
// branch, so all the frontier reads of a thread are in flight together.  A level L
// ORs the new bits into the planes j with bit j of L set (L is uniform, so is the
// branch): bit b of plane j is bit j of d(v, source b), and no lane loops over the
// bits it found (a branch per neighbour read and a loop over new bits took 3 us per
// level).  By symmetry d(src, v) = d(v, src), so the group fills columns
// g*64 .. g*64+63 of every row of the part's S x S int32 matrix: each owned row as
// 256 contiguous bytes at the end (FAP_INF where unreached and for source columns
// >= s), the padding rows >= s with FAP_INF.
#include <cstdint>
#include <type_traits>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"

namespace fap_hip {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int kDeg = 8;       // in-part neighbours kept in registers per vertex
constexpr int kPlanes = 11;   // levels below 2048 > kInpartBfsMaxS
constexpr int kThreads = 512;  // 8 waves: Q = 2 registers' worth of vertices up to 1024

// Frontier buffer stride in words for Q vertex registers per thread: two buffers at
// a fixed distance, so with the level parity known at compile time every frontier
// access is one LDS instruction whose constant part sits in its offset field.
template <int Q>
constexpr int bfs_stride() { return kThreads * Q + 1; }

// OR of the frontier words of the first W neighbour slots of each owned vertex: W
// unconditional loads per vertex (slots past a vertex's degree hold the zero word),
// all in flight before the first OR.
template <int W, int Q>
__device__ __forceinline__ void bfs_gather(unsigned long long (&acc)[Q],
                                           const unsigned long long *cur,
                                           const unsigned (&nbr)[Q][kDeg]) {
  unsigned long long fr[Q][W > 0 ? W : 1];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int k = 0; k < W; ++k)
      fr[q][k] = *reinterpret_cast<const unsigned long long *>(
          reinterpret_cast<const char *>(cur) + nbr[q][k]);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    unsigned long long x = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) x |= fr[q][k];
    acc[q] = x;
  }
}

template <int Q>
__global__ __launch_bounds__(kThreads) void inpart_bfs_kernel(const InpartDesc *__restrict__ descs,
                                                         int nd, int total_groups) {
  // Frontier words, two buffers of bfs_stride<Q>() words (level L reads buffer
  // (L - 1) & 1 and writes the other: one barrier per level), then the three
  // level flags.
  // A static array: its address is a constant, which the LDS instructions fold.
  constexpr int kStride = bfs_stride<Q>();
  __shared__ unsigned long long F[2 * kStride + 2];
  int *found = reinterpret_cast<int *>(F + 2 * kStride);
  const int blk = blockIdx.x;
  if (blk >= total_groups) return;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].group_begin <= blk) lo = mid;
    else hi = mid - 1;
  }
  const InpartDesc d = descs[lo];
  const int *__restrict__ row_off = d.row_off;
  const int *__restrict__ col = d.col;
  const int *__restrict__ graph_id = d.graph_id;
  const int *__restrict__ st2ed = d.st2ed;
  const int *__restrict__ ed2st = d.ed2st;
  const int src0 = (d.g0 + blk - d.group_begin) * 64;
  const int s = d.s, tid = threadIdx.x;

  // "Some vertex gained a bit at level L": written during level L by any wave that
  // found one, read after that level's barrier, and cleared during level L - 1 (three
  // slots: the slot cleared during L was last read after the barrier of L - 2, which
  // every wave has passed).  One barrier per level; __syncthreads_or cost three
  // barriers and a reload of the workgroup size per level.
  if (tid == 0) {
    F[s] = 0;
    F[kStride + s] = 0;
  }
  if (tid < 3) found[tid] = 0;
  // Neighbour lists (byte offsets into a frontier buffer), visited words, first
  // frontier.
  unsigned nbr[Q][kDeg];
  int ovf[Q];  // ovf: CSR index of the first neighbour not kept
  unsigned long long vis[Q], pl[Q][kPlanes];
  int wmax = 0;  // largest in-part degree (capped at kDeg) of the wave's vertices
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int v = tid + kThreads * q;
    int deg = 0;
    ovf[q] = -1;
    vis[q] = 0;
#pragma unroll
    for (int j = 0; j < kPlanes; ++j) pl[q][j] = 0;
#pragma unroll
    for (int k = 0; k < kDeg; ++k) nbr[q][k] = (unsigned)s * 8u;  // the zero word
    if (v < s) {
      const int ver = st2ed[d.offset + v];
      const int e1 = row_off[ver + 1];
      for (int e = row_off[ver]; e < e1; ++e) {
        const int nb = col[e];
        if (graph_id[nb] != d.part_id) continue;
        if (deg == kDeg) {
          ovf[q] = e;
          break;
        }
        // A fixed register index per unrolled slot: no dynamic register indexing.
#pragma unroll
        for (int k = 0; k < kDeg; ++k)
          if (k == deg) nbr[q][k] = (unsigned)(ed2st[nb] - d.offset) * 8u;
        ++deg;
      }
      const int b = v - src0;
      vis[q] = (b >= 0 && b < 64) ? 1ull << b : 0ull;  // d(v, v) = 0: no plane bit
      F[v] = vis[q];
    }
    wmax = max(wmax, deg);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
  wmax = __builtin_amdgcn_readfirstlane(wmax);
  __syncthreads();

  // One level: PAR = level & 1 (reads buffer 1 - PAR, writes buffer PAR).  Returns
  // whether any vertex of the part gained a bit.
  auto step = [&](auto par_c, int level) -> bool {
    constexpr int PAR = decltype(par_c)::value;
    const unsigned long long *cur = F + (1 - PAR) * kStride;
    unsigned long long *nxt = F + PAR * kStride;
    unsigned long long nx[Q];
    switch (wmax) {  // uniform
      case 0: bfs_gather<0, Q>(nx, cur, nbr); break;
      case 1: bfs_gather<1, Q>(nx, cur, nbr); break;
      case 2: bfs_gather<2, Q>(nx, cur, nbr); break;
      case 3: bfs_gather<3, Q>(nx, cur, nbr); break;
      case 4: bfs_gather<4, Q>(nx, cur, nbr); break;
      case 5: bfs_gather<5, Q>(nx, cur, nbr); break;
      case 6: bfs_gather<6, Q>(nx, cur, nbr); break;
      case 7: bfs_gather<7, Q>(nx, cur, nbr); break;
      default: bfs_gather<8, Q>(nx, cur, nbr); break;
    }
    int any = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (ovf[q] >= 0) {  // more than kDeg in-part neighbours: the rest from the CSR
        const int v = tid + kThreads * q;
        const int ver = st2ed[d.offset + v];
        for (int e = ovf[q]; e < row_off[ver + 1]; ++e) {
          const int nb = col[e];
          if (graph_id[nb] == d.part_id) nx[q] |= cur[ed2st[nb] - d.offset];
        }
      }
      nx[q] &= ~vis[q];
      const int v = tid + kThreads * q;
      if (v < s) nxt[v] = nx[q];  // rows >= s never get a bit
      vis[q] |= nx[q];
      any |= nx[q] != 0;
    }
    // A uniform branch per plane: the empty asm keeps the compiler from turning it
    // into an AND with a mask and an OR on every plane of every level.
#pragma unroll
    for (int j = 0; j < kPlanes; ++j)
      if (level & (1 << j)) {
        asm volatile("");
#pragma unroll
        for (int q = 0; q < Q; ++q) pl[q][j] |= nx[q];
      }
    if (__builtin_amdgcn_ballot_w64(any) != 0 && (tid & 63) == 0) found[level % 3] = 1;
    if (tid == 0) found[(level + 1) % 3] = 0;
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(found[level % 3]) != 0;
  };
  for (int level = 1;; level += 2) {
    if (!step(std::integral_constant<int, 1>{}, level)) break;
    if (!step(std::integral_constant<int, 0>{}, level + 1)) break;
  }

  // Owned rows: column src0 + b = d(v, source b) from the planes, on 32-bit halves.
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int v = tid + kThreads * q;
    if (v >= s) continue;
    // d.fw is a generic pointer from the descriptor: gptr gives global stores.
    FAP_GLOBAL v4i *dst = reinterpret_cast<FAP_GLOBAL v4i *>(gptr(d.fw + (int64_t)v * d.S + src0));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t p32[kPlanes];
#pragma unroll
      for (int j = 0; j < kPlanes; ++j) p32[j] = (uint32_t)(pl[q][j] >> (32 * h));
      const uint32_t v32 = (uint32_t)(vis[q] >> (32 * h));
      for (int c = 0; c < 32; c += 4) {
        int out[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int b = c + e;
          uint32_t x = 0;
#pragma unroll
          for (int j = 0; j < kPlanes; ++j) x |= ((p32[j] >> b) & 1u) << j;
          out[e] = ((v32 >> b) & 1u) ? (int)x : kInf;
        }
        dst[(32 * h + c) / 4] = v4i{out[0], out[1], out[2], out[3]};
      }
    }
  }
  // Padding rows s .. S - 1: 16 lanes per row, 4 columns each.
  for (int r = s + (tid >> 4); r < d.S; r += kThreads / 16)
    *reinterpret_cast<FAP_GLOBAL v4i *>(gptr(d.fw + (int64_t)r * d.S + src0 + (tid & 15) * 4)) =
        v4i{kInf, kInf, kInf, kInf};
}

}  // namespace


int launch_inpart_bfs(const InpartDesc *d_descs, int nd, int total_groups, int s_max,
                      hipStream_t stream) {
  if (nd <= 0 || total_groups <= 0) return FAP_OK;
  FAP_REQUIRE(s_max <= kInpartBfsMaxS, "launch_inpart_bfs: part of %d vertices", s_max);
  if (s_max <= 2 * kThreads)
    hipLaunchKernelGGL(inpart_bfs_kernel<2>, dim3(total_groups), dim3(kThreads), 0, stream,
                       d_descs, nd, total_groups);
  else
    hipLaunchKernelGGL(inpart_bfs_kernel<4>, dim3(total_groups), dim3(kThreads), 0, stream,
                       d_descs, nd, total_groups);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

}  // namespace fap_hip
