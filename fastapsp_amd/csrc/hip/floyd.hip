// Blocked three-phase Floyd-Warshall on int32 distances, batched over parts.
//
// Replaces floyd_CPU_path (fap/kernel/floyd.h:38-49, the k-i-j triple loop) and
// its GPU twin floyd_GPU_Nvidia_path_gpu (fap/kernel/cuda/floyd.cu:286-342,
// TILE 32, one launch pair per round per part).
//
// Every part's s x s matrix is stored padded to S = roundup(s, 64) with FAP_INF
// outside the real block, so all tiles are full 64 x 64 tiles and no kernel
// needs bounds checks.  Round r of every part runs in the same three launches:
//   phase 1  closes the pivot tile (r, r) in LDS: 64 k-steps, one barrier each;
//   phase 2  updates the pivot row and column tiles as ONE min-plus product with
//            the closed pivot:  T(r,j) = P* (x) T(r,j),  T(i,r) = T(i,r) (x) P*.
//            With non-negative weights P* has a 0 diagonal, so the product equals
//            the in-place sequential k-loop of the textbook blocked algorithm;
//   phase 3  C(i,j) = min(C(i,j), T(i,r) (x) T(r,j)) for every other tile.
// Phases 2 and 3 use the register-tiled min-plus micro-kernel (mp_tile.h).
// Integer distances make the result identical to the k-i-j loop.
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

// Global row-major 64x64 tile -> LDS row-major panel Bs[k][j].
__device__ __forceinline__ void load_tile(int (*Bs)[kT], const gint *g, int64_t ld) {
  const int k = threadIdx.x >> 2;
  const int jq = threadIdx.x & 3;  // columns jq*16 .. jq*16+15
  const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)k * ld + jq * 16);
  v4i *dst = reinterpret_cast<v4i *>(&Bs[k][jq * 16]);
#pragma unroll
  for (int c = 0; c < 4; ++c) dst[c] = src[c];
}

__global__ __launch_bounds__(256) void fw_phase1_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (r >= d.nb) return;
  __shared__ int D[kT][kT];
  gint *g = tile_ptr(d, r, r);
  load_tile(D, g, d.ld);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int k = 0; k < kT; ++k) {
    __syncthreads();
    // Row k and column k do not change during step k (D[k][k] >= 0), so one
    // barrier per step orders the in-place update.
    const int4 rk = *reinterpret_cast<const int4 *>(&D[k][tx * 4]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = ty * 4 + q;
      const int dik = D[i][k];
      int4 v = *reinterpret_cast<int4 *>(&D[i][tx * 4]);
      v.x = min(v.x, dik + rk.x);
      v.y = min(v.y, dik + rk.y);
      v.z = min(v.z, dik + rk.z);
      v.w = min(v.w, dik + rk.w);
      *reinterpret_cast<int4 *>(&D[i][tx * 4]) = v;
    }
  }
  __syncthreads();
  const int k = threadIdx.x >> 2, jq = threadIdx.x & 3;
  gv4i *dst = reinterpret_cast<gv4i *>(g + (int64_t)k * d.ld + jq * 16);
  const v4i *src = reinterpret_cast<const v4i *>(&D[k][jq * 16]);
#pragma unroll
  for (int c = 0; c < 4; ++c) dst[c] = src[c];
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

// B tile rows k0 .. k0+31, columns 0..63 -> Bs[k][j].
__device__ __forceinline__ void load_b_chunk(int (*Bs)[kT], const gint *g, int64_t ld, int k0) {
  const int k = threadIdx.x >> 1;
  const int half = threadIdx.x & 1;
  const gv4i *src = reinterpret_cast<const gv4i *>(g + (int64_t)(k0 + k) * ld + half * 32);
  v4i *dst = reinterpret_cast<v4i *>(&Bs[k][half * 32]);
#pragma unroll
  for (int c = 0; c < 8; ++c) dst[c] = src[c];
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

// acc = min(acc, A (x) B) for 64 x 64 tiles A (row-major at a) and B (at b).
__device__ __forceinline__ void tile_product(int (&acc)[8][8], const gint *a, const gint *b,
                                             int64_t ld, int (*As)[kT], int (*Bs)[kT], int tr,
                                             int tc) {
#pragma unroll
  for (int k0 = 0; k0 < kT; k0 += kKc) {
    if (k0 > 0) __syncthreads();
    load_a_chunk(As, a, ld, k0);
    load_b_chunk(Bs, b, ld, k0);
    __syncthreads();
    mp_accumulate<8, 8, kKc>(acc, &As[0][0], kT, &Bs[0][0], kT, tr * 4, 32, tc * 4, 32);
  }
}

__global__ __launch_bounds__(64) void fw_phase2_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (r >= d.nb) return;
  const int others = d.nb - 1;
  int t = blockIdx.x;
  if (t >= 2 * others) return;
  const bool is_row = t < others;
  int o = is_row ? t : t - others;
  o = o < r ? o : o + 1;
  __shared__ int As[kKc][kT];
  __shared__ int Bs[kKc][kT];
  const gint *pivot = tile_ptr(d, r, r);
  gint *target = is_row ? tile_ptr(d, r, o) : tile_ptr(d, o, r);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  int acc[8][8];
  load_acc8(acc, target, d.ld, tr, tc);
  // row tile: target = P (x) target;  column tile: target = target (x) P.
  tile_product(acc, is_row ? pivot : target, is_row ? target : pivot, d.ld, As, Bs, tr, tc);
  store_acc8(acc, target, d.ld, tr, tc);
}

__global__ __launch_bounds__(64) void fw_phase3_kernel(const FwDesc *__restrict__ descs, int r) {
  const FwDesc d = descs[blockIdx.y];
  if (r >= d.nb) return;
  const int others = d.nb - 1;
  const int t = blockIdx.x;
  if (t >= others * others) return;
  int i = t / others, j = t - (t / others) * others;
  i = i < r ? i : i + 1;
  j = j < r ? j : j + 1;
  __shared__ int As[kKc][kT];
  __shared__ int Bs[kKc][kT];
  gint *c = tile_ptr(d, i, j);
  const int tr = threadIdx.x >> 3, tc = threadIdx.x & 7;
  int acc[8][8];
  load_acc8(acc, c, d.ld, tr, tc);
  tile_product(acc, tile_ptr(d, i, r), tile_ptr(d, r, j), d.ld, As, Bs, tr, tc);
  store_acc8(acc, c, d.ld, tr, tc);
}

int launch_fw_round(const FwDesc *d_descs, int ndesc, int max_nb, int r, hipStream_t stream) {
  if (ndesc <= 0 || r >= max_nb) return FAP_OK;
  hipLaunchKernelGGL(fw_phase1_kernel, dim3(1, ndesc), dim3(256), 0, stream, d_descs, r);
  FAP_CHECK_LAUNCH();
  const int others = max_nb - 1;
  if (others > 0) {
    hipLaunchKernelGGL(fw_phase2_kernel, dim3(2 * others, ndesc), dim3(64), 0, stream, d_descs,
                       r);
    FAP_CHECK_LAUNCH();
    hipLaunchKernelGGL(fw_phase3_kernel, dim3(others * others, ndesc), dim3(64), 0, stream,
                       d_descs, r);
    FAP_CHECK_LAUNCH();
  }
  return FAP_OK;
}

int launch_fw(const FwDesc *d_descs, int ndesc, int max_nb, hipStream_t stream) {
  for (int r = 0; r < max_nb; ++r) FAP_TRY(launch_fw_round(d_descs, ndesc, max_nb, r, stream));
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

}  // namespace fap_hip

using namespace fap_hip;

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
