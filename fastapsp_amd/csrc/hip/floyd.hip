// Blocked three-phase Floyd-Warshall on int32 distances, batched over parts.
//
// Replaces floyd_CPU_path (fap/kernel/floyd.h:38-49, the k-i-j triple loop) and
// its GPU twin floyd_GPU_Nvidia_path_gpu (fap/kernel/cuda/floyd.cu:286-342,
// TILE 32, one launch pair per round per part).
//
// Every part's s x s matrix is stored padded to S = roundup(s, 64) with FAP_INF
// outside the real block, so every tile is a full 64 x 64 tile and no kernel
// needs bounds checks.  A round pivots on a panel P of kPanel 64-tiles (up to two
// 64-tiles; see kPanel), so a matrix with nb tiles per side takes
// ceil(nb / kPanel) rounds and every other tile is read and written once per round.  Round r of every part runs in the same three launches:
//   phase 1  closes the diagonal block (P, P) in LDS (one 1024-thread workgroup);
//   phase 2  updates the pivot row and column panels as ONE min-plus product with
//            the closed block:  T(P,j) = P* (x) T(P,j),  T(i,P) = T(i,P) (x) P*.
//            With non-negative weights P* has a 0 diagonal and satisfies the
//            triangle inequality, so the product equals the sequential k-loop.
//            A workgroup writes one 64-row (or 64-column) half of a panel tile
//            while another reads it as an operand: either value is a path length
//            at least the new minimum and P*(a,k) + P*(k,m) >= P*(a,m), so the
//            result is the same whichever value is read;
//   phase 3  C(i,j) = min(C(i,j), T(i,P) (x) T(P,j)) for every tile outside P.
// Phases 2 and 3 use the register-tiled min-plus micro-kernel (mp_tile.h) with
// one wave64 per 64 x 64 output tile.  Integer distances make the result
// identical to the k-i-j loop.
#include <algorithm>

#include "fap_internal.h"
#include "fw_desc.h"
#include "gmem.h"
#include "mp_tile.h"

namespace fap_hip {

constexpr int kT = 64;  // tile edge

// Descriptor pointers are generic to the compiler; tiles live in global memory.
typedef int v4i __attribute__((ext_vector_type(4)));
typedef FAP_GLOBAL int gint;
typedef FAP_GLOBAL v4i gv4i;

__device__ __forceinline__ gint *tile_ptr(const FwDesc &d, int ti, int tj) {
  return gptr(d.mat + (int64_t)(ti * kT) * d.ld + tj * kT);
}

// Pivot panel width in 64-tiles.  Two-tile panels halve the rounds (and the C-tile
// traffic) but serialise a 128-step closure in phase 1 (~200 us per round on
// MI355X against ~20 us for 64 steps), which costs more than it saves.
constexpr int kPanel = 1;
constexpr int kP1Threads = kPanel * kT * kPanel * kT / 16;  // 16 entries per thread

__device__ __forceinline__ int panel_tiles(const FwDesc &d, int r) {
  return min(kPanel, d.nb - kPanel * r);
}

// Phase 1: the 64 x 64 diagonal tile closed by 256 threads.  Wave w holds rows w, w + 4,
// ..., w + 60 in registers, lane j their column j.  At step k, column k (the entries
// D[i][k] of a wave's rows) comes from lane k by readlane and row k from LDS, where the
// wave owning row k published it after step k - 1 (two buffers, one barrier per step).
// Row k and column k do not change during step k (D[k][k] = 0, or the padded row and
// column are FAP_INF throughout), so this is the k-i-j loop.  The earlier form (16
// consecutive entries per thread, the tile in LDS, three dependent LDS accesses per step)
// took ~41 us per tile on MI355X, the bound of a single part's Floyd-Warshall (one tile
// per round).
static_assert(kPanel == 1, "phase 1 closes one 64-tile");
__global__ __launch_bounds__(kP1Threads) void fw_phase1_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (r >= d.nb) return;
  __shared__ int Rk[2][kT];
  gint *g = tile_ptr(d, r, r);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = g[(int64_t)(wave + 4 * q) * d.ld + lane];
  if (wave == 0) Rk[0][lane] = v[0];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kT; ++k) {
    const int rk = Rk[k & 1][lane];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = min(v[q], __builtin_amdgcn_readlane(v[q], k) + rk);
    if (k + 1 < kT && wave == ((k + 1) & 3)) Rk[(k + 1) & 1][lane] = v[(k + 1) >> 2];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) g[(int64_t)(wave + 4 * q) * d.ld + lane] = v[q];
}

// Phases 2 and 3: one wave64 per 64x64 tile product, 8 x 8 accumulators per lane
// (rows tr*4 + {0..3} and 32 + tr*4 + {0..3}, columns likewise with tc), k in two
// LDS chunks of 32.  With 16 KB of LDS per workgroup a CU holds ten such waves.
constexpr int kKc = 32;

// A tile rows 0..63, columns k0 .. k0+31 -> As[k][i] (transposed): lane = row.
__device__ __forceinline__ void load_a_chunk(int (*As)[kT], const gint *g, int64_t ld, int k0) {
  const int i = threadIdx.x;
  const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)i * ld + k0);
#pragma unroll
  for (int c = 0; c < kKc / 4; ++c) {
    const v4i t = src[c];
    As[4 * c + 0][i] = t[0];
    As[4 * c + 1][i] = t[1];
    As[4 * c + 2][i] = t[2];
    As[4 * c + 3][i] = t[3];
  }
}

// B tile rows k0 .. k0+31, columns 0..63 -> Bs[k][j]: 16 lanes per row, four rows per
// instruction, so the 16-byte LDS writes of each 8 lanes cover 128 contiguous bytes
// (all banks once) and the global reads 256 contiguous bytes per row.
__device__ __forceinline__ void load_b_chunk(int (*Bs)[kT], const gint *g, int64_t ld, int k0) {
  const int c4 = (threadIdx.x & 15) * 4;
#pragma unroll
  for (int q = 0; q < kKc / 4; ++q) {
    const int k = 4 * q + (threadIdx.x >> 4);
    *reinterpret_cast<v4i *>(&Bs[k][c4]) =
        *reinterpret_cast<const gv4i *>(g + (int64_t)(k0 + k) * ld + c4);
  }
}

__device__ __forceinline__ int own_row(int tr, int q) { return (q < 4 ? 0 : 32) + tr * 4 + (q & 3); }

__device__ __forceinline__ void load_acc8(int (&acc)[8][8], const gint *g, int64_t ld, int tr,
                                          int tc) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const gint *r = g + (int64_t)own_row(tr, q) * ld;
    const v4i a = *reinterpret_cast<const gv4i *>(r + tc * 4);
    const v4i b = *reinterpret_cast<const gv4i *>(r + 32 + tc * 4);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      acc[q][e] = a[e];
      acc[q][4 + e] = b[e];
    }
  }
}

__device__ __forceinline__ void store_acc8(const int (&acc)[8][8], gint *g, int64_t ld, int tr,
                                           int tc) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    gint *r = g + (int64_t)own_row(tr, q) * ld;
    *reinterpret_cast<gv4i *>(r + tc * 4) = v4i{acc[q][0], acc[q][1], acc[q][2], acc[q][3]};
    *reinterpret_cast<gv4i *>(r + 32 + tc * 4) = v4i{acc[q][4], acc[q][5], acc[q][6], acc[q][7]};
  }
}

// acc = min(acc, A (x) B): A is 64 x kdim (row-major at a), B is kdim x 64 (at b),
// kdim a multiple of 32.
__device__ __forceinline__ void tile_product(int (&acc)[8][8], const gint *a, const gint *b,
                                             int64_t ld, int kdim, int (*As)[kT], int (*Bs)[kT],
                                             int tr, int tc) {
  for (int k0 = 0; k0 < kdim; k0 += kKc) {
    if (k0 > 0) __syncthreads();
    load_a_chunk(As, a, ld, k0);
    load_b_chunk(Bs, b, ld, k0);
    __syncthreads();
    mp_accumulate<8, 8, kKc>(acc, &As[0][0], kT, &Bs[0][0], kT, tr * 4, 32, tc * 4, 32);
  }
}

// Tile index o in [0, nb - wt) -> tile outside the pivot panel [p, p + wt).
__device__ __forceinline__ int skip_panel(int o, int p, int wt) { return o < p ? o : o + wt; }

// Phase 2: 2 * wt * (nb - wt) output tiles per part: the row panel tiles (P_a, j)
// first, then the column panel tiles (i, P_a).
__global__ __launch_bounds__(64) void fw_phase2_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (kPanel * r >= d.nb) return;
  const int wt = panel_tiles(d, r), p = kPanel * r, others = d.nb - wt;
  int t = blockIdx.x;
  if (t >= 2 * wt * others) return;
  const bool is_row = t < wt * others;
  if (!is_row) t -= wt * others;
  const int a = t / others;                       // which 64-slice of the panel
  const int o = skip_panel(t - a * others, p, wt);  // the tile outside the panel
  __shared__ int As[kKc][kT];
  __shared__ int Bs[kKc][kT];
  gint *target = is_row ? tile_ptr(d, p + a, o) : tile_ptr(d, o, p + a);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  int acc[8][8];
  load_acc8(acc, target, d.ld, tr, tc);
  // row tile: P*(P_a, P) (x) T(P, o);  column tile: T(o, P) (x) P*(P, P_a).
  if (is_row)
    tile_product(acc, tile_ptr(d, p + a, p), tile_ptr(d, p, o), d.ld, wt * kT, As, Bs, tr, tc);
  else
    tile_product(acc, tile_ptr(d, o, p), tile_ptr(d, p, p + a), d.ld, wt * kT, As, Bs, tr, tc);
  store_acc8(acc, target, d.ld, tr, tc);
}

// Phase 3: (nb - wt)^2 tiles per part outside the pivot row and column.
__global__ __launch_bounds__(64) void fw_phase3_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (kPanel * r >= d.nb) return;
  const int wt = panel_tiles(d, r), p = kPanel * r, others = d.nb - wt;
  const int t = blockIdx.x;
  if (t >= others * others) return;
  const int i = skip_panel(t / others, p, wt);
  const int j = skip_panel(t - (t / others) * others, p, wt);
  __shared__ int As[kKc][kT];
  __shared__ int Bs[kKc][kT];
  gint *c = tile_ptr(d, i, j);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  int acc[8][8];
  load_acc8(acc, c, d.ld, tr, tc);
  tile_product(acc, tile_ptr(d, i, p), tile_ptr(d, p, j), d.ld, wt * kT, As, Bs, tr, tc);
  store_acc8(acc, c, d.ld, tr, tc);
}

// Phases 2 and 3 with four waves per tile (thread t: rows (t / 16) * 4 + {0..3},
// columns (t % 16) * 4 + {0..3}), for launches of few tiles: a single part's
// Floyd-Warshall has (nb - 1)^2 tiles per round, too few to fill the device one wave per
// tile, so each round's latency is that of one wave's 64^3-step product.
constexpr int kWideThreads = 256;
constexpr int64_t kWideTiles = 1024;  // phase-3 tiles per round below which the wide form runs

__device__ __forceinline__ void load_a_chunk_w(int (*As)[kT], const gint *g, int64_t ld, int k0) {
  const int i = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 8;
  const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)i * ld + k0 + kq);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const v4i t = src[c];
#pragma unroll
    for (int e = 0; e < 4; ++e) As[kq + 4 * c + e][i] = t[e];
  }
}

__device__ __forceinline__ void load_b_chunk_w(int (*Bs)[kT], const gint *g, int64_t ld, int k0) {
  const int k = threadIdx.x >> 3, c8 = (threadIdx.x & 7) * 8;
  const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)(k0 + k) * ld + c8);
  *reinterpret_cast<v4i *>(&Bs[k][c8]) = src[0];
  *reinterpret_cast<v4i *>(&Bs[k][c8 + 4]) = src[1];
}

__device__ __forceinline__ void tile_product_w(int (&acc)[4][4], const gint *a, const gint *b,
                                               int64_t ld, int (*As)[kT], int (*Bs)[kT], int tr,
                                               int tc) {
  for (int k0 = 0; k0 < kT; k0 += kKc) {
    if (k0 > 0) __syncthreads();
    load_a_chunk_w(As, a, ld, k0);
    load_b_chunk_w(Bs, b, ld, k0);
    __syncthreads();
    mp_accumulate<4, 4, kKc>(acc, &As[0][0], kT, &Bs[0][0], kT, tr * 4, 0, tc * 4, 0);
  }
}

// kPhase 2: 2 (nb - 1) panel tiles (row panel first); 3: (nb - 1)^2 tiles.  A tile that
// is its own operand (phase 2) is fully staged before any thread stores.
template <int kPhase>
__global__ __launch_bounds__(kWideThreads) void fw_wide_kernel(const FwDesc *__restrict__ descs,
                                                               int r) {
  const FwDesc d = descs[blockIdx.y];
  const int p = r, others = d.nb - 1;
  const int t = blockIdx.x;
  if (r >= d.nb || t >= (kPhase == 2 ? 2 * others : others * others)) return;
  gint *c;
  const gint *a, *b;
  if (kPhase == 2) {
    const bool is_row = t < others;
    const int o = skip_panel(is_row ? t : t - others, p, 1);
    c = is_row ? tile_ptr(d, p, o) : tile_ptr(d, o, p);
    a = is_row ? tile_ptr(d, p, p) : c;
    b = is_row ? c : tile_ptr(d, p, p);
  } else {
    const int i = skip_panel(t / others, p, 1), j = skip_panel(t % others, p, 1);
    c = tile_ptr(d, i, j);
    a = tile_ptr(d, i, p);
    b = tile_ptr(d, p, j);
  }
  __shared__ int As[kKc][kT];
  __shared__ int Bs[kKc][kT];
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  int acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const v4i x = *reinterpret_cast<const gv4i *>(c + (int64_t)(tr * 4 + q) * d.ld + tc * 4);
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[q][e] = x[e];
  }
  tile_product_w(acc, a, b, d.ld, As, Bs, tr, tc);
  __syncthreads();  // (phase 2) every read of the tile precedes the first store
#pragma unroll
  for (int q = 0; q < 4; ++q)
    *reinterpret_cast<gv4i *>(c + (int64_t)(tr * 4 + q) * d.ld + tc * 4) =
        v4i{acc[q][0], acc[q][1], acc[q][2], acc[q][3]};
}

// max_nb: the largest tile count of the descriptors; rounds cover kPanel tiles.
int launch_fw_round(const FwDesc *d_descs, int ndesc, int max_nb, int r, hipStream_t stream) {
  if (ndesc <= 0 || kPanel * r >= max_nb) return FAP_OK;
  hipLaunchKernelGGL(fw_phase1_kernel, dim3(1, ndesc), dim3(kP1Threads), 0, stream, d_descs, r);
  FAP_CHECK_LAUNCH();
  // Grid sized for the largest part; smaller parts leave the extra blocks idle.
  const int wt = std::min(kPanel, max_nb - kPanel * r);
  int others = max_nb - wt;
  // A part with fewer tiles may have a full panel where the largest has a short one.
  if (wt < kPanel) others = std::max(others, max_nb - 1 - kPanel);
  const int p2 = 2 * kPanel * others;
  if (others > 0 && (int64_t)ndesc * others * others < kWideTiles) {  // few tiles
    hipLaunchKernelGGL(fw_wide_kernel<2>, dim3(p2, ndesc), dim3(kWideThreads), 0, stream,
                       d_descs, r);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(fw_wide_kernel<3>, dim3(others * others, ndesc), dim3(kWideThreads), 0,
                       stream, d_descs, r);
    FAP_CHECK_LAUNCH();
  } else if (others > 0) {  // a single-panel matrix has no tile outside the pivot panel
    hipLaunchKernelGGL(fw_phase2_kernel, dim3(p2, ndesc), dim3(64), 0, stream, d_descs, r);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(fw_phase3_kernel, dim3(others * others, ndesc), dim3(64), 0, stream,
                       d_descs, r);
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

int launch_fw(const FwDesc *d_descs, int ndesc, int max_nb, hipStream_t stream) {
  const int rounds = (max_nb + kPanel - 1) / kPanel;
  for (int r = 0; r < rounds; ++r) FAP_TRY(launch_fw_round(d_descs, ndesc, max_nb, r, stream));
  return FAP_OK;
}

// Copies an s x s matrix (stride ld) into a padded S x S buffer (INF outside).
__global__ void pad_in_kernel(int *__restrict__ dst, int S, const int *__restrict__ src, int64_t ld,
                              int s) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)S * S) return;
  const int i = (int)(idx / S), j = (int)(idx - (int64_t)(idx / S) * S);
  dst[idx] = (i < s && j < s) ? src[(int64_t)i * ld + j] : kInf;
}

__global__ void pad_out_kernel(int *__restrict__ dst, int64_t ld, int s, const int *__restrict__ src,
                               int S) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)s * s) return;
  const int i = (int)(idx / s), j = (int)(idx - (int64_t)(idx / s) * s);
  dst[(int64_t)i * ld + j] = src[(int64_t)i * S + j];
}

// ---- allocation-free variant (fap_hip_floyd_noalloc) -------------------------
// One descriptor slot (256 bytes) followed by a padded matrix of at most
// FAP_FLOYD_NOALLOC_MAX_S^2 entries.  In host code `scratch` is only a stand-in for
// the device symbol: its device address on each device comes from
// hipGetSymbolAddress (floyd_scratch_setup) and only that address is ever used.
constexpr int kNoAllocMaxS = FAP_FLOYD_NOALLOC_MAX_S;
constexpr int64_t kScratchDescInts = 64;
static_assert(sizeof(FwDesc) <= 4 * kScratchDescInts, "descriptor slot too small");
static_assert(kNoAllocMaxS % kT == 0, "limit must be a whole number of tiles");
__device__ __align__(256) int scratch[kScratchDescInts + (int64_t)kNoAllocMaxS * kNoAllocMaxS];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int floyd_scratch_setup(int device) {
  FAP_REQUIRE(device >= 0 && device < kMaxDevices, "floyd scratch: device %d out of range",
              device);
  if (g_scratch_ptr[device]) return FAP_OK;
  void *p = nullptr;
  FAP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  FAP_REQUIRE(p != nullptr, "floyd scratch: no device address");
  g_scratch_ptr[device] = p;
  return FAP_OK;
}

__global__ void set_fw_desc_kernel(FwDesc *__restrict__ dst, FwDesc h) { *dst = h; }

}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_floyd_noalloc(int32_t *d_mat, int32_t s, int64_t ld, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(s >= 0 && ld >= s, "fap_hip_floyd_noalloc: bad sizes s=%d ld=%lld", s,
              (long long)ld);
  FAP_REQUIRE(s <= kNoAllocMaxS,
              "fap_hip_floyd_noalloc: s=%d exceeds the scratch limit of %d vertices", s,
              kNoAllocMaxS);
  if (s == 0) return FAP_OK;
  FAP_REQUIRE(d_mat != nullptr, "fap_hip_floyd_noalloc: null matrix");
  const int dev = lib_device();
  FAP_REQUIRE(dev >= 0 && dev < kMaxDevices && g_scratch_ptr[dev],
              "fap_hip_floyd_noalloc: scratch not set up on device %d", dev);
  hipStream_t st = resolve(stream);
  int *base = static_cast<int *>(g_scratch_ptr[dev]);
  FwDesc *desc = reinterpret_cast<FwDesc *>(base);
  int *pad = base + kScratchDescInts;
  const int S = (s + kT - 1) / kT * kT;
  const int64_t tot = (int64_t)S * S;
  // Every entry of the padded S x S block is written, so nothing left by an earlier
  // call (of any size) is read.
  hipLaunchKernelGGL(pad_in_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, pad, S,
                     d_mat, ld, s);
  FAP_CHECK_LAUNCH();
  hipLaunchKernelGGL(set_fw_desc_kernel, dim3(1), dim3(1), 0, st, desc, FwDesc{pad, S, S / kT});
  FAP_CHECK_LAUNCH();
  FAP_TRY(launch_fw(desc, 1, S / kT, st));
  const int64_t out = (int64_t)s * s;
  hipLaunchKernelGGL(pad_out_kernel, dim3((unsigned)((out + 255) / 256)), dim3(256), 0, st, d_mat,
                     ld, s, pad, S);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

extern "C" int fap_hip_floyd(int32_t *d_mat, int32_t s, int64_t ld, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(s >= 0 && ld >= s, "fap_hip_floyd: bad sizes s=%d ld=%lld", s, (long long)ld);
  if (s == 0) return FAP_OK;
  FAP_REQUIRE(d_mat != nullptr, "fap_hip_floyd: null matrix");
  hipStream_t st = resolve(stream);
  static Workspace ws_mat, ws_desc;
  const int S = (s + kT - 1) / kT * kT;
  FAP_TRY(ws_mat.ensure((int64_t)S * S * 4));
  FAP_TRY(ws_desc.ensure(sizeof(FwDesc)));
  int *pad = static_cast<int *>(ws_mat.ptr);
  const int64_t tot = (int64_t)S * S;
  hipLaunchKernelGGL(pad_in_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, pad, S,
                     d_mat, ld, s);
  FAP_CHECK_LAUNCH();
  FwDesc h{pad, S, S / kT};
  FAP_CHECK_HIP(hipMemcpyAsync(ws_desc.ptr, &h, sizeof(h), hipMemcpyHostToDevice, st));
  FAP_TRY(launch_fw(static_cast<FwDesc *>(ws_desc.ptr), 1, S / kT, st));
  const int64_t out = (int64_t)s * s;
  hipLaunchKernelGGL(pad_out_kernel, dim3((unsigned)((out + 255) / 256)), dim3(256), 0, st, d_mat,
                     ld, s, pad, S);
  FAP_CHECK_LAUNCH();
  // The workspace is shared between calls: finish before it can be reused.
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}
