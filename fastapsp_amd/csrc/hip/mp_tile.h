// Register-tiled min-plus micro-kernel on LDS panels (VALU: v_add_u32 + v_min3_i32).
//
// Min-plus is not a multiply-accumulate, so it runs on the vector ALUs, not on
// MFMA.  Each thread owns a TM x TN block of accumulators made of 4-wide groups:
// A rows arow + g*astride + 0..3 (g < TM/4) and B columns bcol + g*bstride + 0..3
// (g < TN/4).  Both panels are stored k-major in LDS (As[k][i], Bs[k][j]) so every
// operand group is one 16-byte LDS read; groups spaced by bstride let the lanes of
// a wave read one contiguous run per instruction (no bank conflicts).  Two k-steps
// are folded into one v_min3:  acc = min(acc, a0 + b0, a1 + b1), i.e. 1.5 VALU
// instructions per (add, min) step.
//
// Distances are int32 in [0, FAP_INF]; FAP_INF + FAP_INF = 2^31 - 2 does not
// overflow, and min() with an accumulator <= FAP_INF keeps every result <= FAP_INF.
#pragma once

#include <hip/hip_runtime.h>

namespace fap_hip {

// acc = min(acc, x, y) as exactly one v_min3_i32.  Written as a one-instruction ALU
// asm statement because the compiler otherwise re-associates the running minimum
// across unrolled k-steps into v_min + v_min3 pairs (1.75 instead of 1.5 VALU
// instructions per min-plus step).  Register operands only; no memory access.
__device__ __forceinline__ int min3_i32(int acc, int x, int y) {
  int r;
  asm("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(acc), "v"(x), "v"(y));
  return r;
}

template <int N>
__device__ __forceinline__ void lds_load_groups(int (&dst)[N], const int *src, int stride) {
#pragma unroll
  for (int g = 0; g < N / 4; ++g) {
    const int4 t = *reinterpret_cast<const int4 *>(src + g * stride);
    dst[4 * g + 0] = t.x;
    dst[4 * g + 1] = t.y;
    dst[4 * g + 2] = t.z;
    dst[4 * g + 3] = t.w;
  }
}

template <int TM, int TN>
__device__ __forceinline__ void mp_step2(int (&acc)[TM][TN], const int *As, int lda, const int *Bs,
                                         int ldb, int arow, int astride, int bcol, int bstride,
                                         int k) {
  int a0[TM], a1[TM], b0[TN], b1[TN];
  lds_load_groups<TM>(a0, As + k * lda + arow, astride);
  lds_load_groups<TN>(b0, Bs + k * ldb + bcol, bstride);
  lds_load_groups<TM>(a1, As + (k + 1) * lda + arow, astride);
  lds_load_groups<TN>(b1, Bs + (k + 1) * ldb + bcol, bstride);
#pragma unroll
  for (int r = 0; r < TM; ++r)
#pragma unroll
    for (int c = 0; c < TN; ++c) acc[r][c] = min3_i32(acc[r][c], a0[r] + b0[c], a1[r] + b1[c]);
}

// acc[r][c] = min(acc[r][c], min_{k < KSTEPS} A[k][row r] + B[k][col c]), KSTEPS even.
template <int TM, int TN, int KSTEPS>
__device__ __forceinline__ void mp_accumulate(int (&acc)[TM][TN], const int *As, int lda,
                                              const int *Bs, int ldb, int arow, int astride,
                                              int bcol, int bstride) {
#pragma unroll 4
  for (int k = 0; k < KSTEPS; k += 2)
    mp_step2<TM, TN>(acc, As, lda, Bs, ldb, arow, astride, bcol, bstride, k);
}

// Run-time step count (kcount even; the panels hold FAP_INF beyond the real K).
template <int TM, int TN>
__device__ __forceinline__ void mp_accumulate_n(int (&acc)[TM][TN], const int *As, int lda,
                                                const int *Bs, int ldb, int arow, int astride,
                                                int bcol, int bstride, int kcount) {
#pragma unroll 2
  for (int k = 0; k < kcount; k += 2)
    mp_step2<TM, TN>(acc, As, lda, Bs, ldb, arow, astride, bcol, bstride, k);
}

}  // namespace fap_hip
