// Device descriptors shared by the kernels of the part pipeline.
#pragma once

#include <cstdint>

namespace fap_hip {

// Largest part the pipelines take: an int32 Floyd-Warshall matrix of 32768^2 entries is
// 4 GB, and every index of such a matrix fits 31 bits.
constexpr int kMaxPart = 32768;

// One padded S x S Floyd-Warshall matrix (S = ld = nb * 64, FAP_INF padding).
struct FwDesc {
  int *mat;
  int64_t ld;
  int nb;
};

// One FP64 matrix of the batched Floyd-Warshall (fp64.hip): S x S, ld = S = nb * 64.
struct Fw64Desc {
  double *mat;
  int nb;
};

// Cap of the 16-bit Floyd-Warshall: an entry holds min(distance, kCap16).  Twice the
// cap is 0x7C00, the bit pattern of +inf in IEEE half precision, so every stored
// value and every sum of two stays a non-negative half whose bit pattern orders as
// the integer does (floyd16.hip relies on this for v_pk_minimum3_f16).
constexpr int kCap16 = 0x3E00;

// The same matrix held as 16-bit distances (ld in elements, rows 128-byte aligned).
struct Fw16Desc {
  uint16_t *mat;
  int64_t ld;
  int nb;
};

// Expansion of one part's 16-bit result into its int32 FW matrix (both S x S,
// ld = S), 8 entries per work item.  Entries equal to the cap become FAP_INF; with
// component labels, a capped entry between two vertices of one connected component
// marks the part in flags[part_index] (its true distance is >= the cap: the part is
// redone in int32).  Without labels the caller guarantees that every finite entry
// is below the cap.
// One part of a bit-parallel in-part BFS launch (inpart.hip): its S x S int32 matrix,
// the launch's first workgroup for it (group_begin) and the source group that
// workgroup takes (g0: sources g0 * 64 ..).
struct InpartDesc {
  int *fw;
  int S, s, offset, part_id;
  int group_begin, g0;
  // The part's graph (CSR, part labels, sorted order), so one launch can serve parts
  // of many graphs.
  const int *row_off, *col, *graph_id, *st2ed, *ed2st;
};
constexpr int kInpartBfsMaxS = 2048;  // levels fit 11 bit-planes; 8 LDS bytes per vertex

struct UnpackDesc {
  const uint16_t *src;
  int *dst;
  int S, s, offset, part_index;
  int64_t item_begin;  // prefix over parts of S * S / 8
};

// One min-plus product C = A (x) B with overwrite semantics:
//   C[i][j] = min_k A[i][k] + B[k][j]  (FAP_INF when K == 0), i < M, j < N.
// When part_id > 0 the epilogue also performs the decode of the part's own
// columns: for a column j with graph_id[j] == part_id the stored value is
// min(C[i][j], fw[(fw_row0 + i) * ldf + ed2st[j] - offset]).  With boundary rows
// in the FW matrix (SSSP form) the FW value is the distance itself; with an
// in-part FW matrix the minimum adds the paths that leave the part.
struct MergeDesc {
  const int *A;
  int64_t lda;
  const int *B;
  int64_t ldb;
  int *C;
  int64_t ldc;
  int M, K;
  int row_blocks;   // ceil(M / 64)
  int tile_begin;   // first logical tile of this product
  int part_id;      // 0 = no fused decode
  int offset;       // first sorted position of the part
  const int *fw;    // Floyd-Warshall matrix of the part
  int64_t ldf;
  int fw_row0;      // row of the FW matrix that holds C's row 0
};

// Sub-matrix build of one part into its padded FW matrix.
struct BuildDesc {
  int *fw;
  int64_t ldf;
  uint16_t *fw16;   // non-null: write the capped 16-bit matrix here (ld = S) instead
  int S;            // padded edge (multiple of 64, <= kBuildMaxS)
  int s, bdy, offset, part_id;
  int row_begin;    // first global build row of this part
  const int *dist_rows;  // the part's SSSP rows (row i = boundary vertex i); NULL: every
                         // row from the part's own arcs (in-part distances)
  int64_t ld_dist;
};

// One part of the boundary graph (boundary.hip): its in-part Floyd-Warshall matrix
// (int32, ld = S; the leading bdy rows and columns are its boundary vertices), its
// first index bs among the boundary graph's vertices and its first output row.
struct BgPart {
  const int *fw;
  int S, bdy, offset, part_id;
  int bs;
  int cb;  // first entry of the part's clique block in the G_B clique weights
  int64_t out_row;
};

// One vertex of the boundary graph as the closure kernel reads it: its part's first
// G_B index and boundary count, its clique in-weights at woff (woff + c holds the
// in-part distance from boundary vertex c of the part to this one) and its first
// kBgCutW cut arcs (G_B target, weight); ncut > kBgCutW means the rest are read from
// the CSR.
constexpr int kBgCutW = 4;
struct BgVert {
  int bs, bdy, woff, ncut;
  int cut_t[kBgCutW];
  int cut_w[kBgCutW];
};

// Decode of the boundary rows' own columns of one part.
struct DecodeDesc {
  const int *fw;
  int64_t ldf;
  int *out;         // first row of the part's block
  int64_t ld_out;
  int rows, s, offset;
  int64_t item_begin;  // prefix over parts of rows * s
};

}  // namespace fap_hip
