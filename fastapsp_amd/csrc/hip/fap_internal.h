// Internal helpers shared by the HIP translation units of libfap_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "fap_hip.h"

namespace fap_hip {

// Work-unit shape of the streaming merge: rows per unit, unit order (1: the column
// blocks of one row block back to back) and resident waves per CU (0: no cap).
struct StreamShape {
  int rows, col_major, waves_per_cu;
};
StreamShape stream_shape();

// Thread-local message of the last failing C-ABI call.
void set_error(const char *fmt, ...);
void clear_error();

// Library stream of the current device (created by fap_hip_init).
hipStream_t lib_stream();
// Internal streams (i = 0, 1) used to overlap the stages of different parts.
hipStream_t aux_stream(int i);
// Compute units of the device (set by fap_hip_init).
int device_cus();

// Workgroup L of a grid of `total` is dispatched to XCD L % 8; returns the index that
// gives every XCD one contiguous range [x * total / 8, ...) (consecutive tiles of a part
// then share that XCD's L2).
__device__ __forceinline__ unsigned xcd_contiguous(unsigned L, unsigned total) {
  const unsigned per = total >> 3, rem = total & 7, x = L & 7, slot = L >> 3;
  return x < rem ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}
// Device the library runs on (set by fap_hip_init), -1 before.
int lib_device();
// Looks up the device address of the allocation-free Floyd-Warshall scratch on
// `device` (current); called once per device by fap_hip_init.
int floyd_scratch_setup(int device);
inline hipStream_t resolve(void *stream) {
  return stream ? static_cast<hipStream_t>(stream) : lib_stream();
}
bool initialized();

// Budgeted device allocation used by every internal workspace.
int dev_alloc(void **p, int64_t bytes);
int dev_free(void *p);

// Grow-only device workspace (one per purpose), released by the library at exit.
struct Workspace {
  void *ptr = nullptr;
  int64_t bytes = 0;
  int ensure(int64_t need);
};

// Ordering of the pipelines' shared workspaces across caller streams.  Every solve
// call makes its stream wait for the fence (the end of the previous call that used
// the workspaces, on whatever stream it ran) before touching them, and records the
// fence at its end.  The fence is an event owned by the library, so it stays valid
// when the caller destroys the stream it was recorded on.
int ws_fence_wait(hipStream_t st);
int ws_fence_record(hipStream_t st);

constexpr int kInf = FAP_INF;

// Stage timing of the solve pipelines, read by fap_hip_last_stage_times.  Stages:
// 0 sssp, 1 build, 2 Floyd-Warshall, 3 merge, 4 decode.
int timing_start(hipStream_t st);                   // new call: reset, record start
int timing_record(hipStream_t on, hipEvent_t *e);   // pooled event recorded on `on`
void timing_add_span(int stage, hipEvent_t a, hipEvent_t b);
int timing_finish(hipStream_t st);                  // record the end of the call
void timing_note_keyed(int merge_parts, int fw_parts);  // path solves: keyed forms taken
void timing_note_bgraph();  // boundary rows came from the boundary graph

// Full solve of one graph through the batch stages (batch.hip: boundary graph closed
// by Floyd-Warshall, any boundary count), timed as one span of stage `stage`.
int solve_one_batched(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                      void *d_out, int64_t ld_out, bool is_float, int stage, hipStream_t st);

// Boundary rows of parts covering a unit-weight graph with paired arcs from the
// boundary graph (boundary.hip); FAP_ERR_INVALID, nothing launched, when the parts do
// not qualify.
int boundary_rows_unit(const fap_graph_dev *g, const fap_part_desc *act, int P, int *const *fw,
                       const int *S, int32_t *d_out, int64_t ld_out, hipStream_t st);

}  // namespace fap_hip

#define FAP_CHECK_HIP(call)                                                            \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      fap_hip::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),        \
                         __FILE__, __LINE__);                                          \
      return FAP_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define FAP_CHECK_LAUNCH()                                                             \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) {                                                            \
      fap_hip::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),   \
                         __FILE__, __LINE__);                                          \
      return FAP_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define FAP_REQUIRE(cond, ...)                                                         \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      fap_hip::set_error(__VA_ARGS__);                                                 \
      return FAP_ERR_INVALID;                                                          \
    }                                                                                  \
  } while (0)

#define FAP_TRY(call)                                                                  \
  do {                                                                                 \
    int s_ = (call);                                                                   \
    if (s_ != FAP_OK) return s_;                                                       \
  } while (0)
