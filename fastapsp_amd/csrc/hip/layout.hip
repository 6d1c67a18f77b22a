// Layout kernels of the part pipeline: sub-matrix build, boundary-row decode and
// the gather of boundary sources.
//
// Build replaces MysubMatBuild_path (fap/utils/util_Myalgorithm.h:44-84) and its
// three-kernel GPU twin (buildFromCSR / buildFromSubGraph / setDiagonal, :87-191,
// which builds the MAXVALUE init on the host and copies it over).  Here one
// workgroup assembles one row in LDS and writes it once, coalesced.
#include <algorithm>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"

namespace fap_hip {

constexpr int kBuildChunk = 8192;  // LDS row buffer: 32 KB

template <typename D>
__device__ __forceinline__ int find_by_begin(const D *__restrict__ d, int nd, int key,
                                             int D::*field) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].*field <= key) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One workgroup per padded row of every part's FW matrix.
//   row i <  bdy : FW[i][j] = dist_rows[i][st2ed[offset + j]]   (dist_rows set)
//   other rows   : FAP_INF, then the part's own out-arcs of st2ed[offset + i] in
//                  CSR order (a later arc overwrites an earlier one)
//   diagonal     : 0;  padding (j >= s or i >= s): FAP_INF.
// With d.fw16 set, the row is written as 16-bit entries min(value, kCap16).  Rows
// longer than the LDS buffer are assembled in chunks of kBuildChunk columns.
template <bool kVec>
__global__ __launch_bounds__(256) void build_kernel(const BuildDesc *__restrict__ descs, int nd,
                                                    int total_rows,
                                                    const int *__restrict__ row_off,
                                                    const int *__restrict__ col,
                                                    const int *__restrict__ w,
                                                    const int *__restrict__ graph_id,
                                                    const int *__restrict__ st2ed,
                                                    const int *__restrict__ ed2st) {
  const int grow = blockIdx.x;
  if (grow >= total_rows) return;
  const BuildDesc d = descs[find_by_begin(descs, nd, grow, &BuildDesc::row_begin)];
  const int i = grow - d.row_begin;
  __shared__ int rowbuf[kBuildChunk];
  const int tid = threadIdx.x;
  const bool from_rows = i < d.s && i < d.bdy && d.dist_rows;
  for (int c0 = 0; c0 < d.S; c0 += kBuildChunk) {
    const int cw = min(kBuildChunk, d.S - c0);  // a multiple of 64
    if (c0 > 0) __syncthreads();  // the previous chunk has been written out
    if (from_rows) {
      const int *src = d.dist_rows + (int64_t)i * d.ld_dist;
      for (int j = tid; j < cw; j += 256)
        rowbuf[j] = (c0 + j < d.s) ? src[st2ed[d.offset + c0 + j]] : kInf;
    } else {
      for (int j = tid; j < cw; j += 256) rowbuf[j] = kInf;
    }
    __syncthreads();
    if (tid == 0 && i < d.s) {
      if (!from_rows) {
        const int ver = st2ed[d.offset + i];
        const int e1 = row_off[ver + 1];
        for (int e = row_off[ver]; e < e1; ++e) {
          const int nb = col[e];
          if (graph_id[nb] != d.part_id) continue;
          const int idx = ed2st[nb] - d.offset - c0;
          if (idx >= 0 && idx < cw && c0 + idx < d.s) rowbuf[idx] = w ? w[e] : 1;
        }
      }
      if (i >= c0 && i < c0 + cw) rowbuf[i - c0] = 0;
    }
    __syncthreads();
    if (d.fw16) {
      // Capped 16-bit row, two entries per 32-bit word (chunks are multiples of 64).
      unsigned *dst16 = reinterpret_cast<unsigned *>(d.fw16 + (int64_t)i * d.S + c0);
      for (int x = tid; x < cw / 2; x += 256) {
        const unsigned lo = (unsigned)min(rowbuf[2 * x], kCap16);
        const unsigned hi = (unsigned)min(rowbuf[2 * x + 1], kCap16);
        dst16[x] = lo | (hi << 16);
      }
      continue;
    }
    int *dst = d.fw + (int64_t)i * d.ldf + c0;
    if (kVec) {
      for (int j = tid * 4; j < cw; j += 1024)
        *reinterpret_cast<int4 *>(dst + j) = *reinterpret_cast<const int4 *>(&rowbuf[j]);
    } else {
      for (int j = tid; j < cw; j += 256) dst[j] = rowbuf[j];
    }
  }
}

// out[i][st2ed[offset + j]] = fw[i][j] for the decode rows of every part.
__global__ void decode_kernel(const DecodeDesc *__restrict__ descs, int nd, int64_t total,
                              const int *__restrict__ st2ed) {
  const int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (item >= total) return;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].item_begin <= item) lo = mid;
    else hi = mid - 1;
  }
  const DecodeDesc d = descs[lo];
  const int64_t local = item - d.item_begin;
  const int i = (int)(local / d.s);
  const int j = (int)(local - (int64_t)i * d.s);
  d.out[(int64_t)i * d.ld_out + st2ed[d.offset + j]] = d.fw[(int64_t)i * d.ldf + j];
}

// Boundary sources of a list of parts: src[k] = st2ed[offset_p + q],
// rows[k] = (out_row_p + q) * ld  for q < bdy_p, in part order.
__global__ void gather_sources_kernel(const int *__restrict__ offsets,
                                      const int *__restrict__ src_begin,
                                      const int64_t *__restrict__ out_rows, int nparts, int total,
                                      const int *__restrict__ st2ed, int64_t ld,
                                      int *__restrict__ src, int64_t *__restrict__ rows) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  int lo = 0, hi = nparts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (src_begin[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  const int q = k - src_begin[lo];
  src[k] = st2ed[offsets[lo] + q];
  rows[k] = (out_rows[lo] + q) * ld;
}

int launch_build(const BuildDesc *d_descs, int nd, int total_rows, bool vec, const int *row_off,
                 const int *col, const int *w, const int *graph_id, const int *st2ed,
                 const int *ed2st, hipStream_t stream) {
  if (nd <= 0 || total_rows <= 0) return FAP_OK;
  if (vec)
    hipLaunchKernelGGL(build_kernel<true>, dim3(total_rows), dim3(256), 0, stream, d_descs, nd,
                       total_rows, row_off, col, w, graph_id, st2ed, ed2st);
  else
    hipLaunchKernelGGL(build_kernel<false>, dim3(total_rows), dim3(256), 0, stream, d_descs, nd,
                       total_rows, row_off, col, w, graph_id, st2ed, ed2st);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_decode(const DecodeDesc *d_descs, int nd, int64_t total, const int *st2ed,
                  hipStream_t stream) {
  if (nd <= 0 || total <= 0) return FAP_OK;
  hipLaunchKernelGGL(decode_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                     d_descs, nd, total, st2ed);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

int launch_gather_sources(const int *offsets, const int *src_begin, const int64_t *out_rows,
                          int nparts, int total, const int *st2ed, int64_t ld, int *src,
                          int64_t *rows, hipStream_t stream) {
  if (nparts <= 0 || total <= 0) return FAP_OK;
  hipLaunchKernelGGL(gather_sources_kernel, dim3((total + 255) / 256), dim3(256), 0, stream,
                     offsets, src_begin, out_rows, nparts, total, st2ed, ld, src, rows);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

}  // namespace fap_hip

using namespace fap_hip;

extern "C" int fap_hip_submat_build(int32_t *d_sub, int64_t ld_sub, const int32_t *d_dist_rows,
                                    int64_t ld_dist, const int32_t *d_row_offset,
                                    const int32_t *d_col, const int32_t *d_w,
                                    const int32_t *d_graph_id, const int32_t *d_st2ed,
                                    const int32_t *d_ed2st, int32_t n, int32_t start, int32_t s,
                                    int32_t bdy, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(s >= 0 && bdy >= 0 && bdy <= s && start >= 0 && start + s <= n && ld_sub >= s,
              "fap_hip_submat_build: bad sizes");
  if (s == 0) return FAP_OK;
  FAP_REQUIRE(s <= kMaxPart, "fap_hip_submat_build: s=%d above %d", s, kMaxPart);
  FAP_REQUIRE(d_sub && d_row_offset && d_graph_id && d_st2ed && d_ed2st &&
                  (bdy == 0 || (d_dist_rows && ld_dist >= n)),
              "fap_hip_submat_build: null pointer");
  hipStream_t st = resolve(stream);
  // Part id of the block = graph_id of its first vertex.
  int part_id = 0, first = 0;
  FAP_CHECK_HIP(hipMemcpyAsync(&first, d_st2ed + start, 4, hipMemcpyDeviceToHost, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  FAP_CHECK_HIP(hipMemcpyAsync(&part_id, d_graph_id + first, 4, hipMemcpyDeviceToHost, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  BuildDesc h{};
  h.fw = d_sub;
  h.ldf = ld_sub;
  h.S = s;
  h.s = s;
  h.bdy = bdy;
  h.offset = start;
  h.part_id = part_id;
  h.row_begin = 0;
  h.dist_rows = d_dist_rows;
  h.ld_dist = ld_dist;
  static Workspace ws;
  FAP_TRY(ws.ensure(sizeof(BuildDesc)));
  FAP_CHECK_HIP(hipMemcpyAsync(ws.ptr, &h, sizeof(h), hipMemcpyHostToDevice, st));
  FAP_TRY(launch_build(static_cast<BuildDesc *>(ws.ptr), 1, s, false, d_row_offset, d_col, d_w,
                       d_graph_id, d_st2ed, d_ed2st, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}

extern "C" int fap_hip_submat_decode(const int32_t *d_sub, int64_t ld_sub, int32_t *d_out,
                                     int64_t ld_out, const int32_t *d_st2ed, int32_t start,
                                     int32_t rows, int32_t s, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(rows >= 0 && s >= 0 && start >= 0 && ld_sub >= s, "fap_hip_submat_decode: bad sizes");
  if (rows == 0 || s == 0) return FAP_OK;
  FAP_REQUIRE(d_sub && d_out && d_st2ed, "fap_hip_submat_decode: null pointer");
  hipStream_t st = resolve(stream);
  DecodeDesc h{};
  h.fw = d_sub;
  h.ldf = ld_sub;
  h.out = d_out;
  h.ld_out = ld_out;
  h.rows = rows;
  h.s = s;
  h.offset = start;
  h.item_begin = 0;
  static Workspace ws;
  FAP_TRY(ws.ensure(sizeof(DecodeDesc)));
  FAP_CHECK_HIP(hipMemcpyAsync(ws.ptr, &h, sizeof(h), hipMemcpyHostToDevice, st));
  FAP_TRY(launch_decode(static_cast<DecodeDesc *>(ws.ptr), 1, (int64_t)rows * s, d_st2ed, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  return FAP_OK;
}

// ---- float <-> int32 distance conversion ---------------------------------------
// The reference's seam works on float distances with MAXVALUE = 1e8 for unreachable
// (fap/utils/parameter.h:26); the int32 kernels use FAP_INF.  For integer weights
// every finite reference distance is an integer below 2^24, so both maps are exact.
namespace {

__global__ void f32_to_i32_kernel(const float *__restrict__ in, int *__restrict__ out,
                                  int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float v = in[i];
    out[i] = v >= 1e8f ? kInf : (int)v;
  }
}

__global__ void i32_to_f32_kernel(const int *__restrict__ in, float *__restrict__ out,
                                  int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int v = in[i];
    out[i] = v >= kInf ? 1e8f : (float)v;
  }
}

unsigned conv_grid(int64_t count) {
  return (unsigned)std::min<int64_t>((count + 255) / 256, 8LL * device_cus() * 8);
}

}  // namespace

extern "C" int fap_hip_f32_to_i32(const float *d_src, int32_t *d_dst, int64_t count,
                                  void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(count >= 0, "fap_hip_f32_to_i32: negative count");
  if (count == 0) return FAP_OK;
  FAP_REQUIRE(d_src && d_dst, "fap_hip_f32_to_i32: null pointer");
  hipLaunchKernelGGL(f32_to_i32_kernel, dim3(conv_grid(count)), dim3(256), 0, resolve(stream),
                     d_src, d_dst, count);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

extern "C" int fap_hip_i32_to_f32(const int32_t *d_src, float *d_dst, int64_t count,
                                  void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(count >= 0, "fap_hip_i32_to_f32: negative count");
  if (count == 0) return FAP_OK;
  FAP_REQUIRE(d_src && d_dst, "fap_hip_i32_to_f32: null pointer");
  hipLaunchKernelGGL(i32_to_f32_kernel, dim3(conv_grid(count)), dim3(256), 0, resolve(stream),
                     d_src, d_dst, count);
  FAP_CHECK_LAUNCH();
  return FAP_OK;
}

// ---- layout validation -----------------------------------------------------------
// Sorted position pos (one thread each): v = st2ed[pos] in range, ed2st[v] == pos, v's
// label is the id of the part whose rows hold pos, and an inner vertex (local index >=
// the part's boundary count) has no arc to another part.  Any violation sets *bad.
namespace {

__global__ void check_layout_kernel(const int *__restrict__ row_off, const int *__restrict__ col,
                                    const int *__restrict__ graph_id, const int *__restrict__ st2ed,
                                    const int *__restrict__ ed2st, int n,
                                    const fap_part_desc *__restrict__ parts, int np,
                                    int *__restrict__ bad) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  int lo = 0, hi = np - 1;  // parts sorted by offset, tiling [0, n)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (parts[mid].offset <= pos) lo = mid;
    else hi = mid - 1;
  }
  const fap_part_desc q = parts[lo];
  const int v = st2ed[pos];
  bool ok = v >= 0 && v < n && pos >= q.offset && pos < q.offset + q.num;
  if (ok) ok = ed2st[v] == pos && graph_id[v] == q.part_id;
  if (ok && pos - q.offset >= q.bdy)
    for (int e = row_off[v]; e < row_off[v + 1] && ok; ++e) {
      const int u = col[e];
      ok = u >= 0 && u < n && graph_id[u] == q.part_id;
    }
  if (!ok) atomicOr(bad, 1);
}

}  // namespace

extern "C" int fap_hip_check_layout(const fap_graph_dev *g, const fap_part_desc *parts,
                                    int32_t nparts, int32_t *ok) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && ok && nparts >= 1 && parts && g->n > 0 && g->d_row_offset && g->d_graph_id &&
                  g->d_st2ed && g->d_ed2st && (g->m == 0 || g->d_col),
              "fap_hip_check_layout: null argument");
  *ok = 0;
  std::vector<fap_part_desc> ps(parts, parts + nparts);
  std::sort(ps.begin(), ps.end(),
            [](const fap_part_desc &a, const fap_part_desc &b) { return a.offset < b.offset; });
  int next = 0;
  for (const fap_part_desc &q : ps) {
    if (q.offset != next || q.num < 0 || q.bdy < 0 || q.bdy > q.num) return FAP_OK;  // no tiling
    next += q.num;
  }
  if (next != g->n) return FAP_OK;
  hipStream_t st = lib_stream();
  static Workspace ws;
  FAP_TRY(ws.ensure(256 + sizeof(fap_part_desc) * ps.size()));
  int *d_bad = static_cast<int *>(ws.ptr);
  fap_part_desc *d_parts = reinterpret_cast<fap_part_desc *>(static_cast<char *>(ws.ptr) + 256);
  FAP_TRY(fap_hip_fill_i32(d_bad, 0, 1, st));
  FAP_CHECK_HIP(hipMemcpyAsync(d_parts, ps.data(), sizeof(fap_part_desc) * ps.size(),
                               hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(check_layout_kernel, dim3((g->n + 255) / 256), dim3(256), 0, st,
                     g->d_row_offset, g->d_col, g->d_graph_id, g->d_st2ed, g->d_ed2st, g->n,
                     d_parts, (int)ps.size(), d_bad);
  FAP_CHECK_LAUNCH();
  int bad = 1;
  FAP_CHECK_HIP(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  *ok = bad == 0;
  return FAP_OK;
}
