// Blocked three-phase Floyd-Warshall on capped 16-bit distances, batched over parts.
//
// Same rounds and phases as the int32 form in floyd.hip (floyd_CPU_path,
// fap/kernel/floyd.h:38-49), on matrices whose entries are min(d, kCap16) as uint16.
// Capping keeps the closure exact below the cap: if every stored entry is
// min(true value, C), then
//   min(min(d_ij, C), min(d_ik, C) + min(d_kj, C)) = min(d_ij, d_ik + d_kj, C)
// (if either operand is capped the sum is >= C >= the stored d_ij), so the final
// matrix is min(closure, C) entry by entry: every entry below C is exact, and an
// entry equal to C means "C or more, or unreachable".  The pipeline turns C into
// FAP_INF and redoes a part in int32 when C appears between two vertices of one
// connected component (UnpackDesc).
//
// Phases 2 and 3 run the min-plus steps on packed pairs: two adjacent columns per
// 32-bit register, v_pk_add_u16 for the sums (the A value broadcast to both halves
// through op_sel) and one v_pk_minimum3_f16 folding two k-steps into the running
// minimum.  Every value and sum lies in [0, 0x7C00]: non-negative IEEE halves up to
// +inf, whose bit patterns order as the integers do (no NaN, no negative zero, and
// the kernels keep FP16 denormals), so the half-precision minimum returns exactly
// the integer minimum.  That is 0.75 VALU instructions per (add, min) step and
// entry, against 1.5 for int32 with v_min3_i32, and half the bytes per tile.
#include <algorithm>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"
#include "pk16.h"

namespace fap_hip {

namespace {

constexpr int kT = 64;         // tile edge (entries)
constexpr int kTW = kT / 2;    // 32-bit words per tile row

typedef int v4i __attribute__((ext_vector_type(4)));
typedef FAP_GLOBAL uint16_t g16;
typedef FAP_GLOBAL v4i gv4i;

__device__ __forceinline__ g16 *tile16(const Fw16Desc &d, int ti, int tj) {
  return gptr(d.mat + (int64_t)(ti * kT) * d.ld + tj * kT);
}

// Phase 1: the 64 x 64 diagonal tile closed by one wave: lane i holds row i as 32
// packed words.  Step k broadcasts row k through LDS (lane k writes it, every lane
// reads it back) and updates D[i][j] = min(D[i][j], D[i][k] + D[k][j]) on packed
// pairs.  Row k and column k do not change during step k (D[k][k] >= 0), so the
// row read at step k is final for that step.
__device__ __forceinline__ unsigned pk_add_half(unsigned a, int half, unsigned b) {
  return half ? pk_add_hi(a, b) : pk_add_lo(a, b);
}

__global__ __launch_bounds__(64) void fw16_phase1_kernel(const Fw16Desc *__restrict__ descs,
                                                         int r) {
  const Fw16Desc d = descs[blockIdx.x];
  if (r >= d.nb) return;
  __shared__ unsigned rowk[kTW];
  g16 *g = tile16(d, r, r);
  const int lane = threadIdx.x;
  unsigned row[kTW];
  {
    const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)lane * d.ld);
#pragma unroll
    for (int c = 0; c < kTW / 4; ++c) {
      const v4i t = src[c];
#pragma unroll
      for (int e = 0; e < 4; ++e) row[4 * c + e] = (unsigned)t[e];
    }
  }
#pragma unroll
  for (int k = 0; k < kT; ++k) {
    if (lane == k) {
#pragma unroll
      for (int c = 0; c < kTW / 4; ++c)
        *reinterpret_cast<v4i *>(&rowk[4 * c]) =
            v4i{(int)row[4 * c], (int)row[4 * c + 1], (int)row[4 * c + 2], (int)row[4 * c + 3]};
    }
    __syncthreads();  // one wave: orders lane k's write before the reads
    const unsigned dik = row[k >> 1];
    unsigned sum[kTW];  // all sums first: a minimum right after its sum costs a wait state
#pragma unroll
    for (int c = 0; c < kTW / 4; ++c) {
      const v4i rk = *reinterpret_cast<const v4i *>(&rowk[4 * c]);
#pragma unroll
      for (int e = 0; e < 4; ++e) sum[4 * c + e] = pk_add_half(dik, k & 1, (unsigned)rk[e]);
    }
#pragma unroll
    for (int w = 0; w < kTW; ++w)
      row[w] = __builtin_bit_cast(unsigned, __builtin_elementwise_min(__builtin_bit_cast(us2, row[w]),
                                                                      __builtin_bit_cast(us2, sum[w])));
    __syncthreads();  // every lane has read row k before lane k + 1 overwrites it
  }
  gv4i *dst = reinterpret_cast<gv4i *>(g + (int64_t)lane * d.ld);
#pragma unroll
  for (int c = 0; c < kTW / 4; ++c)
    dst[c] = v4i{(int)row[4 * c], (int)row[4 * c + 1], (int)row[4 * c + 2], (int)row[4 * c + 3]};
}

// Phases 2 and 3: one wave64 per 64 x 64 tile.  Lane (tr, tc) = (lane >> 3, lane & 7)
// owns rows tr*4 + {0..3} and 32 + tr*4 + {0..3}, and the 8 columns 8*tc .. 8*tc+7
// (4 packed words).  LDS: As[kk][i] = the word (A[i][2kk], A[i][2kk+1]) (transposed
// so a lane reads its rows' k-pair as one 16-byte run), Bs[k][w] = B's row k.
__device__ __forceinline__ int own_row(int tr, int q) { return (q < 4 ? 0 : 32) + tr * 4 + (q & 3); }

__device__ __forceinline__ void load_acc(unsigned (&acc)[8][4], const g16 *g, int64_t ld, int tr,
                                         int tc) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const v4i t = *reinterpret_cast<const gv4i *>(g + (int64_t)own_row(tr, q) * ld + tc * 8);
#pragma unroll
    for (int w = 0; w < 4; ++w) acc[q][w] = (unsigned)t[w];
  }
}

__device__ __forceinline__ void store_acc(const unsigned (&acc)[8][4], g16 *g, int64_t ld, int tr,
                                          int tc) {
#pragma unroll
  for (int q = 0; q < 8; ++q)
    *reinterpret_cast<gv4i *>(g + (int64_t)own_row(tr, q) * ld + tc * 8) =
        v4i{(int)acc[q][0], (int)acc[q][1], (int)acc[q][2], (int)acc[q][3]};
}

// acc = min(acc, A (x) B) for one 64 x 64 A tile (at a) and B tile (at b).
__device__ __forceinline__ void tile_product16(unsigned (&acc)[8][4], const g16 *a, const g16 *b,
                                               int64_t ld, unsigned (*As)[kT],
                                               unsigned (*Bs)[kTW], int tr, int tc) {
  const int lane = threadIdx.x;
  {
    // A: lane = row, 8 x 16 bytes of its row, stored transposed.
    const gv4i *src = reinterpret_cast<const gv4i *>(a + (int64_t)lane * ld);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const v4i t = src[c];
#pragma unroll
      for (int e = 0; e < 4; ++e) As[4 * c + e][lane] = (unsigned)t[e];
    }
    // B: 8 rows per instruction, 16 bytes per lane.
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = 8 * c + (lane >> 3), w = (lane & 7) * 4;
      *reinterpret_cast<v4i *>(&Bs[k][w]) =
          *reinterpret_cast<const gv4i *>(b + (int64_t)k * ld + 2 * w);
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int kk = 0; kk < kTW; ++kk) {
    const v4i a0 = *reinterpret_cast<const v4i *>(&As[kk][tr * 4]);
    const v4i a1 = *reinterpret_cast<const v4i *>(&As[kk][32 + tr * 4]);
    const v4i b0 = *reinterpret_cast<const v4i *>(&Bs[2 * kk][tc * 4]);
    const v4i b1 = *reinterpret_cast<const v4i *>(&Bs[2 * kk + 1][tc * 4]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      // The 8 sums of a row first, then its 4 minima: a packed minimum reading a
      // register written by the instruction just before it costs a wait state.
      const unsigned av = (unsigned)(q < 4 ? a0[q] : a1[q - 4]);
      unsigned s0[4], s1[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        s0[w] = pk_add_lo(av, (unsigned)b0[w]);
        s1[w] = pk_add_hi(av, (unsigned)b1[w]);
      }
#pragma unroll
      for (int w = 0; w < 4; ++w) acc[q][w] = pk_min3(acc[q][w], s0[w], s1[w]);
    }
  }
}

// Tile index o in [0, nb - 1) -> tile outside the pivot p.
__device__ __forceinline__ int skip_pivot(int o, int p) { return o < p ? o : o + 1; }

// blockIdx.x -> logical block so that consecutive logical blocks (the tiles of one
// part) run on one XCD; workgroups are dispatched round-robin over the 8 XCDs.
// gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_logical_block() {
  const int per_xcd = gridDim.x >> 3;
  return (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
}

// Phase 2: per part the row tiles (p, o) then the column tiles (o, p), o != p.
// Logical block t = part * 2 * (max_nb - 1) + tile.
__global__ __launch_bounds__(64) void fw16_phase2_kernel(const Fw16Desc *__restrict__ descs,
                                                         int ndesc, int per_part, int r) {
  const int t = xcd_logical_block();
  const int part = t / per_part;
  if (part >= ndesc) return;
  const Fw16Desc d = descs[part];
  const int others = d.nb - 1;
  int o = t - part * per_part;
  if (r >= d.nb || o >= 2 * others) return;
  const bool is_row = o < others;
  if (!is_row) o -= others;
  o = skip_pivot(o, r);
  __shared__ unsigned As[kTW][kT];
  __shared__ unsigned Bs[kT][kTW];
  g16 *target = is_row ? tile16(d, r, o) : tile16(d, o, r);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  unsigned acc[8][4];
  load_acc(acc, target, d.ld, tr, tc);
  if (is_row)
    tile_product16(acc, tile16(d, r, r), tile16(d, r, o), d.ld, As, Bs, tr, tc);
  else
    tile_product16(acc, tile16(d, o, r), tile16(d, r, r), d.ld, As, Bs, tr, tc);
  store_acc(acc, target, d.ld, tr, tc);
}

// Phase 3: the (nb - 1)^2 tiles outside the pivot row and column of every part.
// Logical block t = part * (max_nb - 1)^2 + tile.
__global__ __launch_bounds__(64) void fw16_phase3_kernel(const Fw16Desc *__restrict__ descs,
                                                         int ndesc, int per_part, int r) {
  const int t = xcd_logical_block();
  const int part = t / per_part;
  if (part >= ndesc) return;
  const Fw16Desc d = descs[part];
  const int others = d.nb - 1;
  const int o = t - part * per_part;
  if (r >= d.nb || o >= others * others) return;
  const int i = skip_pivot(o / others, r);
  const int j = skip_pivot(o - (o / others) * others, r);
  __shared__ unsigned As[kTW][kT];
  __shared__ unsigned Bs[kT][kTW];
  g16 *c = tile16(d, i, j);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  unsigned acc[8][4];
  load_acc(acc, c, d.ld, tr, tc);
  tile_product16(acc, tile16(d, i, r), tile16(d, r, j), d.ld, As, Bs, tr, tc);
  store_acc(acc, c, d.ld, tr, tc);
}

// 16-bit result -> int32 FW matrix, 8 entries per thread; cap -> FAP_INF; a capped
// entry between vertices of one component sets flags[part_index].
__global__ __launch_bounds__(256) void fw16_unpack_kernel(const UnpackDesc *__restrict__ descs,
                                                          int nd, int64_t total,
                                                          const int *__restrict__ comp_sorted,
                                                          int *__restrict__ flags) {
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= total) return;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].item_begin <= item) lo = mid;
    else hi = mid - 1;
  }
  const UnpackDesc d = descs[lo];
  const int64_t e0 = (item - d.item_begin) * 8;  // first entry (row-major, ld = S)
  const int i = (int)(e0 / d.S);
  const int j0 = (int)(e0 - (int64_t)i * d.S);
  const v4i packed = *reinterpret_cast<const gv4i *>(gptr(d.src) + e0);
  int out[8];
  bool over = false;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int v = (int)(((unsigned)packed[e >> 1] >> (16 * (e & 1))) & 0xFFFFu);
    const int j = j0 + e;
    if (v >= kCap16) {
      out[e] = kInf;
      if (comp_sorted && i < d.s && j < d.s &&
          comp_sorted[d.offset + i] == comp_sorted[d.offset + j])
        over = true;
    } else {
      out[e] = v;
    }
  }
  gv4i *dst = reinterpret_cast<gv4i *>(gptr(d.dst) + e0);
  dst[0] = v4i{out[0], out[1], out[2], out[3]};
  dst[1] = v4i{out[4], out[5], out[6], out[7]};
  if (over) flags[d.part_index] = 1;
}

__global__ void comp_sorted_kernel(const int *__restrict__ comp, const int *__restrict__ st2ed,
                                   int n, int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = comp[st2ed[i]];
}

int round_up8(int64_t x) { return (int)((x + 7) / 8 * 8); }

}  // namespace

int launch_fw16(const Fw16Desc *d_descs, int ndesc, int max_nb, hipStream_t stream) {
  if (ndesc <= 0) return FAP_OK;
  const int others = max_nb - 1;
  const int per2 = 2 * others, per3 = others * others;
  const int64_t g2 = (int64_t)per2 * ndesc, g3 = (int64_t)per3 * ndesc;
  if (g3 >= (1LL << 30)) {
    set_error("launch_fw16: grid too large (%lld tiles)", (long long)g3);
    return FAP_ERR_INVALID;
  }
  for (int r = 0; r < max_nb; ++r) {
    hipLaunchKernelGGL(fw16_phase1_kernel, dim3(ndesc), dim3(64), 0, stream, d_descs, r);
    FAP_CHECK_LAUNCH();
    if (others > 0) {  // a one-tile matrix has no tile outside the pivot
      hipLaunchKernelGGL(fw16_phase2_kernel, dim3(round_up8(g2)), dim3(64), 0, stream, d_descs,
                         ndesc, per2, r);
      FAP_CHECK_LAUNCH();
      hipLaunchKernelGGL(fw16_phase3_kernel, dim3(round_up8(g3)), dim3(64), 0, stream, d_descs,
                         ndesc, per3, r);
      FAP_CHECK_LAUNCH();
    }
  }
  return FAP_OK;
}

int launch_fw16_unpack(const UnpackDesc *d_descs, int nd, int64_t total_items,
                       const int *d_comp_sorted, int *d_flags, hipStream_t stream) {
  if (nd <= 0 || total_items <= 0) return FAP_OK;
  hipLaunchKernelGGL(fw16_unpack_kernel, dim3((unsigned)((total_items + 255) / 256)), dim3(256), 0,
                     stream, d_descs, nd, total_items, d_comp_sorted, d_flags);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_comp_sorted(const int *d_comp, const int *d_st2ed, int n, int *d_out,
                       hipStream_t stream) {
  if (n <= 0) return FAP_OK;
  hipLaunchKernelGGL(comp_sorted_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_comp,
                     d_st2ed, n, d_out);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

}  // namespace fap_hip
