/*
 * fap_hip.h -- C-ABI of the MI355X (gfx950) Fast-APSP hot path.
 *
 * Plain pointers, int64 sizes and strides, an int status return (never exit()),
 * explicit streams (a hipStream_t passed as void*, NULL = the library stream).
 * Every d_* pointer is device memory owned by the caller unless stated.
 *
 * Distances are int32 with FAP_INF (= INT32_MAX / 2) for "unreachable"; the
 * reference's float MAXVALUE = 1e8 (fap/utils/parameter.h:26) maps to FAP_INF and
 * every finite integer distance below 2^24 maps 1:1, so INF + INF never overflows.
 *
 * Each entry point names the reference interface it replaces (file:line inside
 * TomerHaham/FastAPSP).  The seam there is C++ free functions selected by
 * #ifdef WITH_CUDA / WITH_HIP (fap/kernel.h:15-17); here it is this C-ABI.
 */
#ifndef FAP_HIP_H_
#define FAP_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FAP_INF 1073741823 /* INT32_MAX / 2 */

enum fap_status {
  FAP_OK = 0,
  FAP_ERR_INVALID = 1, /* bad argument (sizes, null pointers, alignment) */
  FAP_ERR_NOMEM = 2,   /* device allocation failed                        */
  FAP_ERR_HIP = 3,     /* a HIP runtime call failed                       */
  FAP_ERR_BUDGET = 4,  /* the request exceeds the HBM budget              */
  FAP_ERR_IO = 5,      /* file missing or malformed                       */
  FAP_ERR_STATE = 6,   /* call out of order (e.g. solve before preCondition) */
  FAP_ERR_NODEV = 7    /* no usable GPU                                   */
};

/* Execution forms of the per-part pipeline (identical results).
 * PLAIN:   one launch sequence per part (SSSP, build, FW rounds, min-plus,
 *          decode), parts in turn -- the structure of fapGraph::run
 *          (fap/fap.cpp:195-503) called from solve() (fap/fap.cpp:507-518).
 * BATCHED: every stage launched once for all parts (one SSSP grid over all
 *          boundary sources, one build grid, three FW launches per round for all
 *          parts, one min-plus launch per padded-K group), stages in turn.
 * OVERLAP: SSSP and build as in BATCHED, then the parts in chunks on two internal
 *          streams so one chunk's FW rounds run beside another chunk's min-plus.
 *          On MI355X both stages are limited by the memory system and contend, so
 *          this form measures slower than BATCHED; it is kept for comparison. */
enum fap_mode { FAP_MODE_PLAIN = 0, FAP_MODE_BATCHED = 1, FAP_MODE_OVERLAP = 2 };
/* Flag OR-ed into a mode: run every Floyd-Warshall in int32 even where the 16-bit
 * form applies (see fap_hip_solve_parts).  Results are identical either way. */
#define FAP_MODE_FLAG_INT32_FW 0x100
/* Flag OR-ed into a mode: compute the boundary rows with one SSSP per boundary vertex
 * even where the BATCHED form would take them from the boundary graph (see
 * fap_hip_solve_parts).  Results are identical either way. */
#define FAP_MODE_FLAG_SSSP 0x200
/* Flag OR-ed into a mode: compute the in-part distances of the boundary-graph form by
 * Floyd-Warshall even where the bit-parallel BFS applies (see fap_hip_solve_parts).
 * Results are identical either way. */
#define FAP_MODE_FLAG_FW 0x400

/* ------------------------------------------------------------------ runtime */

/* Select the device and create the library stream.  Idempotent.
 * Replaces the implicit CUDA context use of fap/fap.cpp:222-388. */
int fap_hip_init(int device);
/* Message of the last failing call on this host thread ("" if none). */
const char *fap_hip_last_error(void);
/* Device properties: CU count, LDS bytes per workgroup, total HBM bytes. */
int fap_hip_device_info(int *cu_count, int64_t *lds_bytes, int64_t *hbm_bytes, char *name, int name_len);

/* HBM budget for fap_hip_malloc (0 = no limit).  Allocations beyond it fail with
 * FAP_ERR_BUDGET before touching the allocator.  Replaces the cudaMemGetInfo
 * accounting of fap/fap.cpp:240-243,296-299,356-358. */
int fap_hip_set_budget(int64_t bytes);
int fap_hip_mem_stats(int64_t *in_use, int64_t *high_water, int64_t *budget);
int fap_hip_malloc(void **d_ptr, int64_t bytes);
int fap_hip_free(void *d_ptr);
int fap_hip_memcpy_h2d(void *d_dst, const void *h_src, int64_t bytes, void *stream);
int fap_hip_memcpy_d2h(void *h_dst, const void *d_src, int64_t bytes, void *stream);
int fap_hip_memcpy_d2d(void *d_dst, const void *d_src, int64_t bytes, void *stream);
/* Strided D2H of rows x cols int32 (pitch in elements). */
int fap_hip_memcpy2d_d2h_i32(int32_t *h_dst, int64_t h_ld, const int32_t *d_src, int64_t d_ld,
                             int64_t rows, int64_t cols, void *stream);
int fap_hip_fill_i32(int32_t *d_ptr, int32_t value, int64_t count, void *stream);
/* Plain device copy kernel (16-byte accesses when aligned); also the known-size
 * reference for HBM traffic counters. */
int fap_hip_copy_i32(int32_t *d_dst, const int32_t *d_src, int64_t count, void *stream);
int fap_hip_stream_sync(void *stream);
int fap_hip_device_sync(void);

/* Events for timing on a given stream (ms between two recorded events). */
int fap_hip_event_create(void **ev);
int fap_hip_event_destroy(void *ev);
int fap_hip_event_record(void *ev, void *stream);
int fap_hip_event_elapsed_ms(void *ev_start, void *ev_end, float *ms);
/* *done = 1 when all work recorded before the event has completed (no wait). */
int fap_hip_event_query(void *ev, int32_t *done);
/* Additional non-blocking streams (default priority) for work that may overlap. */
int fap_hip_stream_create(void **stream);
int fap_hip_stream_destroy(void *stream);

/* ------------------------------------------------------- kernel-level seam */

/* Boundary SSSP: one shortest-path row per source over the WHOLE graph.
 * d_out[i * ld_out + v] = dist(d_src[i], v), FAP_INF if unreachable.
 * d_w == NULL: unit weights (BFS).  Otherwise int32 weights >= 0.
 * CSR: d_row_offset[n + 1], d_col[m], d_w[m].  ld_out >= n.
 * Replaces handle_boundry_path / handle_boundry_path_data_on_gpu
 * (fap/kernel/batch_sssp.h:254-303) -> batched_sssp_path (:232-252) ->
 * dijkstra_path (:79-114); the reference seam takes st2ed + offset, here
 * d_src = d_st2ed + offset. */
int fap_hip_sssp_batch(const int32_t *d_row_offset, const int32_t *d_col, const int32_t *d_w,
                       int32_t n, int64_t m, const int32_t *d_src, int32_t nsrc,
                       int32_t *d_out, int64_t ld_out, void *stream);

/* Floyd-Warshall closure of one s x s int32 matrix in place (row stride ld >= s).
 * Entries must be >= 0 or FAP_INF with a 0 diagonal.  Result equals the
 * k-i-j triple loop of floyd_CPU_path (fap/kernel/floyd.h:38-49) for distances.
 * Replaces floyd_path_gpu / floyd_path (fap/kernel/floyd.h:150-170). */
int fap_hip_floyd(int32_t *d_mat, int32_t s, int64_t ld, void *stream);

/* Allocation-free variant of fap_hip_floyd.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and result as
 * fap_hip_floyd; the padded matrix and the descriptor live in a scratch array that
 * the library holds on every device it runs on until the process ends
 * (4 * (64 + FAP_FLOYD_NOALLOC_MAX_S^2) bytes = 64 MiB), so the call allocates
 * nothing, copies nothing from the host and does not wait for the stream: it only
 * enqueues kernels on `stream`.  s > FAP_FLOYD_NOALLOC_MAX_S is refused with
 * FAP_ERR_INVALID.  Every call on a device shares that device's scratch: calls that
 * overlap on one device (on different streams without ordering between them, or
 * from several host threads) race; order them on one stream. */
#define FAP_FLOYD_NOALLOC_MAX_S 4096
int fap_hip_floyd_noalloc(int32_t *d_mat, int32_t s, int64_t ld, void *stream);

/* Min-plus product with overwrite semantics:
 *   C[i][j] = min(FAP_INF, min_k A[i][k] + B[k][j]),  0 <= i < M, 0 <= j < N.
 * K == 0 gives all FAP_INF.  Strides in elements.
 * Replaces min_plus_path_advanced / min_plus_path_advanced_gpu
 * (fap/kernel/minplus.h:81-105) -> min_plus_path_CPU_advanced (:51-64). */
int fap_hip_minplus(const int32_t *d_A, int64_t lda, const int32_t *d_B, int64_t ldb,
                    int32_t *d_C, int64_t ldc, int32_t M, int32_t N, int32_t K, void *stream);

/* Seam entries with predecessor ("path") matrices, as the reference's seam carries
 * them.  Distances are int32 (>= 0, FAP_INF for unreachable); the results equal the
 * reference CPU branch's distances and paths bit for bit, tie-breaking included.
 *
 * Floyd-Warshall closure with paths, in place: where d[i][k] + d[k][j] < d[i][j]
 * for pivots k in ascending order, d[i][j] takes the sum and path[i][j] = path[k][j]
 * (the k-i-j loop with strict '>').  s <= 8192.  Replaces floyd_path
 * (fap/kernel/floyd.h:162-170) -> floyd_CPU_path (:38-49). */
int fap_hip_floyd_path(int32_t *d_mat, int32_t *d_path, int32_t s, int64_t ld, void *stream);

/* Min-plus with paths, overwrite semantics: C[i][j] = min_k A[i][k] + B[k][j] over
 * finite sums (FAP_INF if none) and Cpath[i][j] = Bpath[k][j] for the smallest k
 * reaching it (-1 if none).  Replaces min_plus_path_advanced[_gpu]
 * (fap/kernel/minplus.h:81-105) -> min_plus_path_CPU_advanced (:51-64). */
int fap_hip_minplus_path(const int32_t *d_A, int64_t lda, const int32_t *d_B, int64_t ldb,
                         const int32_t *d_Bpath, int64_t ldbp, int32_t *d_C, int64_t ldc,
                         int32_t *d_Cpath, int64_t ldcp, int32_t M, int32_t N, int32_t K,
                         void *stream);

/* Boundary SSSP with predecessors: the rows of fap_hip_sssp_batch, and
 * d_path[i * ld_out + v] = the vertex before v on the path Dijkstra with a
 * (distance, vertex) min-heap and strict '>' relaxation picks (the first-settled
 * in-neighbour u with d(u) + w(u, v) = d(v)), the source on its own entry, -1 when
 * unreachable.  d_rrow_offset / d_rcol / d_rw: the reverse CSR (in-arcs of every
 * vertex; d_rw NULL exactly when d_w is NULL).  The predecessors are the reference's
 * for positive weights; with zero-weight arcs they are valid shortest-path
 * predecessors but Dijkstra's choice among equal distances depends on its heap
 * order (as for fap_hip_solve_parts_path).  Replaces handle_boundry_path
 * (fap/kernel/batch_sssp.h:280-303) -> dijkstra_path (:79-114). */
int fap_hip_sssp_batch_path(const int32_t *d_row_offset, const int32_t *d_col,
                            const int32_t *d_w, const int32_t *d_rrow_offset,
                            const int32_t *d_rcol, const int32_t *d_rw, int32_t n, int64_t m,
                            const int32_t *d_src, int32_t nsrc, int32_t *d_out, int32_t *d_path,
                            int64_t ld_out, void *stream);

/* Distance conversion at the seam's edges: float (MAXVALUE = 1e8 for unreachable,
 * fap/utils/parameter.h:26) -> int32 maps v >= 1e8 to FAP_INF and truncates others
 * (exact for the integer-valued distances of integer weights); int32 -> float maps
 * FAP_INF to 1e8. */
int fap_hip_f32_to_i32(const float *d_src, int32_t *d_dst, int64_t count, void *stream);
int fap_hip_i32_to_f32(const int32_t *d_src, float *d_dst, int64_t count, void *stream);

/* Dense s x s sub-matrix of one part (row stride ld_sub >= s):
 *   rows i <  bdy : d_dist_rows[i * ld_dist + st2ed[start + j]]  (the SSSP rows)
 *   rows i >= bdy : FAP_INF, then each intra-part arc of st2ed[start+i] in CSR
 *                   order writes its weight (a later arc wins)
 *   diagonal      : 0.
 * d_w == NULL: unit weights.  Replaces MysubMatBuild_path
 * (fap/utils/util_Myalgorithm.h:44-84) / LaunchMysubMatBuildKernel (:143-191). */
int fap_hip_submat_build(int32_t *d_sub, int64_t ld_sub, const int32_t *d_dist_rows, int64_t ld_dist,
                         const int32_t *d_row_offset, const int32_t *d_col, const int32_t *d_w,
                         const int32_t *d_graph_id, const int32_t *d_st2ed, const int32_t *d_ed2st,
                         int32_t n, int32_t start, int32_t s, int32_t bdy, void *stream);

/* Scatter rows x s of the sub-matrix into final columns:
 *   d_out[i * ld_out + st2ed[start + j]] = d_sub[i * ld_sub + j].
 * Replaces MysubMatDecode_path (fap/utils/util_Myalgorithm.h:236-247) /
 * LaunchMysubMatDecodePathKernel (:215-231). */
int fap_hip_submat_decode(const int32_t *d_sub, int64_t ld_sub, int32_t *d_out, int64_t ld_out,
                          const int32_t *d_st2ed, int32_t start, int32_t rows, int32_t s, void *stream);

/* ----------------------------------------------------- fused part pipeline */

/* Device view of one graph after preCondition (all arrays device-resident). */
typedef struct fap_graph_dev {
  int32_t n;
  int64_t m;
  const int32_t *d_row_offset; /* [n + 1] */
  const int32_t *d_col;        /* [m]     */
  const int32_t *d_w;          /* [m] or NULL for unit weights */
  const int32_t *d_graph_id;   /* [n], part ids 1..K */
  const int32_t *d_st2ed;      /* [n], sorted position -> vertex */
  const int32_t *d_ed2st;      /* [n], vertex -> sorted position */
  /* [n] connected-component label of each vertex, or NULL.  Only for graphs in which
   * every arc u->v has a reverse arc v->u (so reachability is symmetric); it enables
   * the 16-bit Floyd-Warshall of fap_hip_solve_parts. */
  const int32_t *d_comp;
  /* [m] float32 arc weights for fap_hip_solve_parts_f32, or NULL (unit weights). */
  const float *d_wf;
} fap_graph_dev;

/* One part: C_BlockVer_offset / C_BlockVer_num / C_BlockBdy_num of
 * fap/utils/util_preprocess.h:98-105 and the first output row of its block. */
typedef struct fap_part_desc {
  int32_t part_id; /* 1-based, as graph_id */
  int32_t offset;  /* first sorted position of the part */
  int32_t num;     /* s_p: vertices in the part */
  int32_t bdy;     /* b_p: boundary vertices (first b_p sorted positions) */
  int64_t out_row; /* first row of the part's s_p x n block in d_out */
} fap_part_desc;

/* Device time of the last fap_hip_solve_parts call (ms): total wall time on the
 * stream, and per stage the summed duration of that stage's launches (in the
 * OVERLAP form stages of different chunks overlap, so the stages can sum to more
 * than the total). */
typedef struct fap_stage_times {
  float sssp_ms, build_ms, floyd_ms, minplus_ms, decode_ms, total_ms;
  int32_t parts_fw16;   /* parts whose Floyd-Warshall ran in the 16-bit form */
  int32_t parts_redone; /* of those, parts redone in int32 (a distance reached the cap) */
  int32_t boundary_graph; /* 1: boundary rows from the boundary graph, 0: boundary SSSP */
  int32_t parts_keyed;  /* path solves: parts whose min-plus ran in the keyed form */
  int32_t parts_fw_keyed; /* path solves: parts whose Floyd-Warshall ran in the keyed form */
  int32_t parts_bfs;    /* parts whose in-part distances came from the bit-parallel BFS */
  int32_t merge_packed; /* boundary-graph form on unit weights: 1 when the merge ran on
                           packed 16-bit rows, 0 when on int32 rows; -1 otherwise */
} fap_stage_times;

/* Optional check of a C-ABI caller's layout before fap_hip_solve_parts (whose
 * boundary-graph form relies on it; fapGraph builds it correctly): the parts tile
 * [0, n) in sorted positions, st2ed / ed2st are inverse permutations, every vertex's
 * label is the id of the part holding its position, and no inner vertex (local index
 * >= the part's boundary count) has an arc to another part (util_preprocess.h:32-106).
 * *ok = 1 when all hold.  Synchronous; O(n + m).  With a layout that fails it, solves
 * stay within their buffers but their results are undefined. */
int fap_hip_check_layout(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                         int32_t *ok);

/* The whole per-part pipeline of fapGraph::run (fap/fap.cpp:195-503) for a list
 * of parts: boundary SSSP -> sub-matrix build -> Floyd-Warshall -> min-plus of
 * inner rows against the boundary rows, with the decode of in-part columns fused
 * into the min-plus epilogue.  Part p writes rows out_row .. out_row + s_p - 1 of
 * d_out (row stride ld_out >= n; ld_out % 4 == 0 and 16-byte aligned d_out give
 * vector stores).  Host-side descriptors; the call is asynchronous on `stream`.
 * mode: FAP_MODE_PLAIN, FAP_MODE_BATCHED or FAP_MODE_OVERLAP, optionally with
 * FAP_MODE_FLAG_INT32_FW (identical results).
 * 16-bit Floyd-Warshall: when g->d_comp is set (and the flag is absent), every part's
 * matrix is closed on 16-bit entries min(d, 15872) (packed VALU arithmetic, twice
 * the int32 rate).  Entries below the cap are exact; an entry at the cap is
 * unreachable when its two vertices lie in different components.  A part in which
 * a capped entry joins two vertices of one component is redone in int32 before the
 * call returns, so the results never differ from the int32 form.  In this form the
 * call waits for the stream before it returns (it reads the per-part flags).
 * Boundary rows: in the BATCHED form, when the parts cover every vertex and hold at
 * most 4096 boundary vertices in all (and FAP_MODE_FLAG_SSSP is absent), the boundary
 * rows come from the boundary graph instead of one SSSP per
 * boundary vertex: each part's in-part distances are computed on its in-part arcs, the
 * graph of all boundary vertices (in-part boundary distances plus cut arcs) is closed
 * by Bellman-Ford sweeps (for unit weights with work lists in LDS), and one min-plus
 * pass writes every boundary row.  Unit-weight graphs
 * then always take the 16-bit Floyd-Warshall (no flag check, no wait), others int32.
 * The stage times report this stage as sssp_ms.  In this form, on unit-weight graphs
 * with component labels (every arc paired with its reverse) and parts of at most 2048
 * vertices, the in-part distances come from a bit-parallel BFS (64 sources per 64-bit
 * word) instead of Floyd-Warshall, unless FAP_MODE_FLAG_FW or FAP_MODE_FLAG_INT32_FW
 * is set; the stage times report it as floyd_ms. */
int fap_hip_solve_parts(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                        int32_t *d_out, int64_t ld_out, int32_t mode, void *stream);

/* A sweep of solveSubGraph calls (the reference solves every part of its rank in turn,
 * app/MoreTaskMpiExample.cpp:91-96, each call one run() of fap/fap.cpp:195-503), split
 * into a once-per-graph half and a per-part half.  No single reference function: this
 * pair replaces the repeated handle_boundry_path + MysubMatBuild_path + floyd_path of
 * fap.cpp:405-428 with one batched pass.
 * fap_hip_prepare_parts: boundary SSSP, sub-matrix build and int32 Floyd-Warshall of
 * every listed part in one batched pass.  Part p's boundary rows go to rows out_row ..
 * out_row + bdy - 1 of d_rows (stride ld_rows >= n); its padded S x S Floyd-Warshall
 * matrix (S = s_p rounded up to 64) to d_fw + fw_off[p] (fw_off[p] = -1 for an empty
 * part).  With d_fw == NULL only fw_off and *fw_entries (int32 entries of d_fw) are
 * computed.  Asynchronous on `stream`. */
int fap_hip_prepare_parts(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                          int32_t *d_rows, int64_t ld_rows, int32_t *d_fw, int64_t *fw_off,
                          int64_t *fw_entries, void *stream);

/* Per part of `parts` (boundary rows out_row .. out_row + bdy - 1 of d_rows, as written by
 * fap_hip_prepare_parts): counts[p] = the columns outside the part that one of its
 * boundary rows reaches (an entry below FAP_INF).  Waits for the stream. */
int fap_hip_count_reached(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                          const int32_t *d_rows, int64_t ld_rows, int32_t *counts, void *stream);

/* One part's s_p x n block (rows out_row .. of d_out, stride ld_out >= n) from its
 * prepared Floyd-Warshall matrix d_fw (S x S) and boundary rows d_rows (bdy rows, stride
 * ld_rows): the boundary rows copied with their own columns from FW (MysubMatDecode_path,
 * fap/fap.cpp:492-496), the inner rows by min-plus against the boundary rows
 * (min_plus_path_advanced, fap.cpp:437-485) with own columns from FW.  reached < 0: the
 * dense product over every column (stage times: row copy as sssp_ms, product as
 * minplus_ms, boundary rows' own columns as decode_ms).  reached >= 0 (the part's
 * fap_hip_count_reached count): the product over the reached columns only, into a
 * compact matrix, then one pass writes every row (slot map as sssp_ms, product as
 * minplus_ms, the row pass as decode_ms).  The block is identical in both forms.
 * Asynchronous on `stream`. */
int fap_hip_merge_prepared(const fap_graph_dev *g, const fap_part_desc *part,
                           const int32_t *d_fw, const int32_t *d_rows, int64_t ld_rows,
                           int64_t reached, int32_t *d_out, int64_t ld_out, void *stream);
int fap_hip_last_stage_times(fap_stage_times *t);

/* Work-unit shape of the streaming min-plus merge of fap_hip_solve_parts (the
 * MatrixMulCUDA6_path launches of fap/kernel/cuda/minplus.cu, for K <= 32): rows per
 * unit (8..1024, a multiple of 8; one wave keeps a K x 256-column panel of boundary rows
 * in registers and writes `rows` rows of 256 columns), unit order (1: the column blocks
 * of one row block back to back, 0: the row blocks of one column block) and a cap on
 * the resident waves per CU (0: none, else 4..32).  Results do not depend on the shape.  Default
 * 128 / 0 / 0.  A solve reads the shape once, when it plans its merge; set it between
 * solves.  No counterpart in the reference (its tile is fixed at compile time). */
int fap_hip_set_merge_shape(int32_t rows, int32_t unit_order, int32_t waves_per_cu);
int fap_hip_get_merge_shape(int32_t *rows, int32_t *unit_order, int32_t *waves_per_cu);

/* Write probe: stores into `rows` x `cols` int32 entries (row stride ld) with the
 * streaming merge's store pattern (units of the merge shape's rows x 256 columns, one
 * 1 KB row segment per wave and row), only every sample_stride-th row block, and
 * returns the device time in *ms (the call waits for the stream).  Overwrites the
 * buffer.  How fast the merge writes a result buffer depends on where the buffer
 * lies in device memory; the probe measures it.  No counterpart in the reference. */
int fap_hip_write_probe(int32_t *d, int64_t ld, int32_t rows, int32_t cols,
                        int32_t sample_stride, void *stream, float *ms);

/* Allocates a rows x ld int32 result buffer (as fap_hip_malloc; free it with
 * fap_hip_free) placed where the merge writes fastest: up to `candidates` buffers of
 * that size are allocated side by side (as many as memory and the budget allow, at
 * least one), each is timed with the write probe on one row block in 8, the fastest
 * is kept and the others are freed.  On MI355X a linear fill of 52.5 GB measured 9.5 ms
 * in some ranges of device memory and 7.9-8.1 ms in others, and the headline merge
 * 10.6 against 9.0 ms; which range a new allocation lands in varies.  *probe_ms (may
 * be NULL) receives the kept buffer's probe time (0 when candidates == 1).  The buffer
 * content is undefined.  The high-water mark counts the candidates held at once.  No
 * counterpart in the reference. */
int fap_hip_malloc_placed(void **d_ptr, int64_t rows, int64_t ld, int32_t candidates,
                          float *probe_ms);

/* The same pipeline with predecessor matrices, as fapGraph::run with is_path_needed
 * (fap/fap.cpp:195-503) computes them on its CPU path: d_path[row][v] = the vertex
 * before v on the chosen shortest path from the row's source, -1 if unreachable, the
 * source itself on the diagonal, with the reference's tie-breaking at every stage
 * (Dijkstra's heap order for the boundary rows, k-i-j order with strict '>' for
 * Floyd-Warshall, first k for the min-plus).  Integer weights only (g->d_wf NULL).
 * d_rrow_offset / d_rcol / d_rw: the reverse CSR (in-arcs of every vertex; d_rw NULL
 * exactly when g->d_w is NULL).  d_path has the layout of d_out.  Weights must be
 * positive: with a zero-weight arc the vertex Dijkstra settles first among equal
 * distances depends on when it entered the heap, a tie-break no rule on the
 * distance rows reproduces (the distances are still exact, the predecessors are
 * then valid shortest-path predecessors, not necessarily the reference's).  One
 * batched form.  Per part, the Floyd-Warshall panels and the
 * min-plus run in a keyed form (distance and path source in one 32-bit key, one
 * unsigned minimum per candidate) when the part's values stay below the key limits,
 * in a compare-and-select form otherwise; the results are the same.  The call waits
 * for the stream twice to read the maxima that choose the forms.  Boundary-row
 * distances: from the boundary graph (as in fap_hip_solve_parts, with the in-part BFS of
 * the boundary sources) when the parts cover a unit-weight graph with component labels,
 * hold at most 4096 boundary vertices and at most 2048 vertices each, else by SSSP; the
 * predecessors follow from the distance rows either way.  Stage times as for
 * fap_hip_solve_parts, with parts_keyed / parts_fw_keyed counting the keyed parts and
 * boundary_graph set when the boundary graph gave the boundary rows. */
int fap_hip_solve_parts_path(const fap_graph_dev *g, const int32_t *d_rrow_offset,
                             const int32_t *d_rcol, const int32_t *d_rw,
                             const fap_part_desc *parts, int32_t nparts, int32_t *d_out,
                             int32_t *d_path, int64_t ld_out, void *stream);

/* The same pipeline for float32 weights (g->d_wf), the reference's float path:
 * distances are carried in FP64 (exact sums of float32 weights while a path sum
 * spans fewer than 53 significant bits) and written as float32, with MAXVALUE = 1e8
 * (fap/parameter.h:26) for unreachable entries and distances >= 1e8.  Output rows as
 * in fap_hip_solve_parts.  OVERLAP runs as BATCHED.  Workspace: FP64 FW matrices and
 * FP64 SSSP rows (sum of b_p x n doubles). */
int fap_hip_solve_parts_f32(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                            float *d_out, int64_t ld_out, int32_t mode, void *stream);

/* ------------------------------------------------------------ many graphs */
/* One graph of a batch: its device view, its parts (part p writes rows out_row ..
 * out_row + s_p - 1; out_row = offset gives the n x n layout of solve()), the result
 * buffer (int32, or float32 when is_float: weights in g.d_wf) with row stride ld_out. */
typedef struct fap_batch_item {
  fap_graph_dev g;
  const fap_part_desc *parts;
  int32_t nparts;
  int32_t is_float;
  void *d_out;
  int64_t ld_out;
  int32_t flags; /* FAP_ITEM_* */
  /* > 0: the parts tile the graph but only the rows of the part with this id are
   * produced (a sub-graph block, solveSubGraph(id)): boundary rows from the boundary
   * graph (every part's in-part distances and the closed boundary graph) instead of
   * one SSSP per boundary vertex; the other parts' out_row are ignored. */
  int32_t only_part;
} fap_batch_item;
/* Boundary rows by SSSP even when the parts tile the graph: for graphs with parallel
 * arcs inside a part whose later arc (CSR order) is heavier, where the reference's
 * sub-matrix build keeps the later arc and its boundary rows the lightest
 * (fapGraph sets it from preCondition). */
#define FAP_ITEM_BOUNDARY_SSSP 0x1
/* Floyd-Warshall of this (float-weight) item's parts and boundary graph on 32 x 32 tiles
 * (the per-graph form of FAP_BATCH_FW_TILE32; an autotuner knob per size class). */
#define FAP_ITEM_FW_TILE32 0x2

/* A prepared batch: the work tables of every part of every graph and optionally the
 * captured launch sequence (FAP_BATCH_GRAPH: fap_hip_batch_run replays a hipGraph).
 * Per distance type (int32, FP64) every stage is one launch over all graphs: build,
 * batched Floyd-Warshall of the parts, and for items whose parts cover their graph
 * the boundary graph (built, closed by Floyd-Warshall, then one min-plus pass writing
 * the boundary rows); items with a subset of the parts take one SSSP per boundary
 * vertex instead.  The merge of the inner rows is a persistent grid pulling (part, row
 * block, column block) units from a device counter, largest parts first.  Parts of at
 * most 32768 vertices.  The item arrays are copied; device buffers must outlive the
 * batch.  No counterpart in the reference (its solve() loops over parts of one graph,
 * fap/fap.cpp:507-518). */
typedef struct fap_batch fap_batch;
#define FAP_BATCH_GRAPH 0x1
/* flags bits 8..15: persistent merge workgroups per CU (0 = as many as stay resident). */
#define FAP_BATCH_PER_CU(x) (((x)&0xFF) << 8)
/* flags bits 16..17: boundary sources per unit of the boundary-row product
 * (0 = 16, 1 = 8, 2 = 16, 3 = 32); bits 18..19: inner rows per wave of the int32 merge
 * (0 = 16, 1 = 8, 2 = 16). */
#define FAP_BATCH_SRC_ROWS(code) (((code)&0x3) << 16)
#define FAP_BATCH_MERGE_ROWS(code) (((code)&0x3) << 18)
/* The int32 merge of unit-weight graphs with n <= 15872 vertices runs on packed 16-bit
 * distances (exact: every finite distance is below the 16-bit cap) unless this flag
 * is set. */
#define FAP_BATCH_NO_PACKED 0x100000
/* In-part distances of the parts of full items on unit-weight graphs with component
 * labels (every arc paired with its reverse) come from the bit-parallel BFS of
 * fap_hip_solve_parts (parts of at most 2048 vertices) instead of build + Floyd-Warshall,
 * unless this flag is set.  The results are the same. */
#define FAP_BATCH_NO_BFS 0x200000
/* The inner-row merge as a plain grid (one workgroup per unit, in unit order) instead
 * of the persistent work queue; with both distance types in a batch their launch
 * sequences run side by side, and a plain grid lets the other type's launches take
 * workgroup slots as merge units end. */
#define FAP_BATCH_UNIT_GRID 0x400000
/* Floyd-Warshall of the FP64 (float-weight) parts and boundary graphs on 32 x 32 tiles
 * (matrices padded to multiples of 32) instead of 64 x 64: less padding for small parts,
 * more rounds; an autotuner knob per size class.  Results are the same. */
#define FAP_BATCH_FW_TILE32 0x800000
/* In-part distances of weighted parts (of items whose parts tile the graph) by
 * Floyd-Warshall on the whole part even above 160 vertices, instead of the nested
 * closure (sub-parts of at most 64 vertices closed one tile each, then their crossing
 * graph, then two min-plus passes): the same matrices.  The nested closure is planned on
 * the host at creation (a copy of each graph's CSR, a BFS bisection of each part, about
 * 0.1 ms per part), so it pays for batches that are replayed; the one-run forms
 * (fap_hip_solve_batch, the single-graph routes through the batch stages, the request
 * server) set this flag. */
#define FAP_BATCH_NO_NESTED 0x1000000
int fap_hip_batch_create(fap_batch **b, const fap_batch_item *items, int32_t nitems,
                         int32_t flags);
int fap_hip_batch_run(fap_batch *b, void *stream);
/* Parts of the batch whose in-part distances come from the nested closure (see
 * FAP_BATCH_NO_NESTED), or a negative status.  No counterpart in the reference. */
int fap_hip_batch_nested_parts(const fap_batch *b);
int fap_hip_batch_destroy(fap_batch *b);
/* Solve every item: FAP_MODE_PLAIN = graph after graph, each the per-part launch
 * sequence (int32 FW); FAP_MODE_BATCHED / OVERLAP = one persistent launch (create,
 * run, wait, destroy). */
int fap_hip_solve_batch(const fap_batch_item *items, int32_t nitems, int32_t mode, void *stream);

/* --------------------------------------------------------------- graph API */
/* C wrapper of the C++ fap::fapGraph (include/fap/fap.h) for FFI callers
 * (ctypes, cgo).  Mirrors fap/fap.h:100-150.  Opaque handle. */
typedef struct fap_graph fap_graph;

/* (file, directed, weighted, K, partitioner) ctor, fap/fap.h:100-101.
 * file: "<name>" opened as graph/<(un)weight>-<(un)directed>/<name>.mtx relative
 * to the working directory (fap/utils/readMatFile.h:43-51), or a path ending in
 * ".mtx" opened as given. partitioner: "metis" (default when NULL or ""). */
int fap_graph_create_from_file(fap_graph **g, const char *file, int directed, int weighted,
                               int32_t K, const char *partitioner);
/* The .mtx reader alone (readVerEdges + readMatFile, fap/utils/readMatFile.h:36-173):
 * call with NULL arrays to get n and m, then with row_offset[n+1], col[m], weight[m]. */
int fap_mtx_read(const char *file, int directed, int weighted, int32_t *n, int64_t *m,
                 int32_t *row_offset, int32_t *col, float *weight);
/* From a CSR in the reference's layout (row_offset[n+1], col_val[m], weight[m]
 * floats, weight may be NULL = unit).  fap/fap.h:102-104. */
int fap_graph_create_from_csr(fap_graph **g, int32_t n, int64_t m, const int32_t *row_offset,
                              const int32_t *col_val, const float *weight, int directed, int32_t K);
/* Synthetic graphs: kind 0 = road grid (W x H grid, drop p_a of edges, add
 * down-right diagonals with p_b), 1 = road-like percolated grid trimmed to
 * exactly n_target vertices (keep p_a, chords at rate p_b), 2 = directed k-NN
 * geometric graph (n_target points, k = W nearest, int weights in [1, H]).
 * Seeded std::mt19937_64; identical on every machine. */
int fap_graph_generate(fap_graph **g, int32_t kind, int32_t W, int32_t H, int32_t n_target,
                       double p_a, double p_b, uint64_t seed, int32_t K);
int fap_graph_destroy(fap_graph *g);
/* Replace the partition with a caller's (1-based ids, e.g. METIS output) before
 * preCondition; with the reference's partition the row layout is the reference's.
 * Replaces the readIdFile_METIS result of fap/fap.cpp:87-89. */
int fap_graph_set_partition(fap_graph *g, const int32_t *graph_id);
/* findBoundry + sort_and_encode (fap/utils/util_preprocess.h:32-106) for a given
 * partition: st2ed[n], ed2st[n], per-part num / offset / boundary count [K+1]. */
int fap_layout_for_partition(int32_t n, const int32_t *row_offset, const int32_t *col, int32_t K,
                             const int32_t *graph_id, int32_t *st2ed, int32_t *ed2st,
                             int32_t *blk_num, int32_t *blk_off, int32_t *bdy_num);
int fap_graph_precondition(fap_graph *g);
int fap_graph_is_split(fap_graph *g, int *is_split);
int fap_graph_set_mode(fap_graph *g, int32_t mode);
int fap_graph_solve(fap_graph *g);
int fap_graph_solve_subgraph(fap_graph *g, int32_t id);
/* fapGraph::setSubGraphSweep (after_calls >= 0; a negative value changes nothing) and
 * *ready = 0 before the sweep's batched pass has run, 1 after it (dense product), 2 after
 * it (sparse product), 3 when the blocks come from a full solve(); ready may be NULL.  No reference
 * counterpart: the reference recomputes every part's SSSP and Floyd-Warshall on each
 * solveSubGraph call (fap/fap.cpp:405-428). */
int fap_graph_set_subgraph_sweep(fap_graph *g, int32_t after_calls, int32_t *ready);
/* solve(is_path_needed) / solveSubGraph(id, is_path_needed) (fap/fap.h:127-128). */
int fap_graph_solve_ex(fap_graph *g, int32_t is_path_needed);
int fap_graph_solve_subgraph_ex(fap_graph *g, int32_t id, int32_t is_path_needed);
/* Rows of the predecessors of solve(true) / of the last solveSubGraph(id, true). */
int fap_graph_get_path_rows(fap_graph *g, int64_t row0, int64_t rows, int32_t *h_out);
int fap_graph_get_subgraph_path_rows(fap_graph *g, int32_t id, int64_t row0, int64_t rows,
                                     int32_t *h_out);
int fap_graph_sizes(fap_graph *g, int32_t *n, int64_t *m, int32_t *K);
/* Host copies: CSR (row_offset[n+1], col[m], weight[m] as int32), layout arrays
 * (graph_id[n], st2ed[n], ed2st[n]) and per-part counts (num[K+1], offset[K+1],
 * bdy[K+1], index 0 unused as in the reference).  NULL outputs are skipped. */
int fap_graph_get_csr(fap_graph *g, int32_t *row_offset, int32_t *col, int32_t *weight);
/* The arc weights as stored (float32, CSR order); fap_graph_get_csr truncates to int32. */
int fap_graph_get_weights_f32(fap_graph *g, float *weight);
int fap_graph_get_layout(fap_graph *g, int32_t *graph_id, int32_t *st2ed, int32_t *ed2st,
                         int32_t *blk_num, int32_t *blk_off, int32_t *bdy_num);
/* Rows [row0, row0 + rows) of the full n x n result of solve(), int32 (D2H). */
int fap_graph_get_dist_rows(fap_graph *g, int64_t row0, int64_t rows, int32_t *h_out);
/* The s_p x n block of the last solveSubGraph(id). */
int fap_graph_get_subgraph_dist(fap_graph *g, int32_t id, int32_t *h_out);
/* Rows [row0, row0 + rows) of that block (D2H of a slice instead of the whole block). */
int fap_graph_get_subgraph_rows(fap_graph *g, int32_t id, int64_t row0, int64_t rows,
                                int32_t *h_out);
/* *is_float = 1 when some weight is not an integer below 2^24: solves then run
 * fap_hip_solve_parts_f32 and the int32 getters return FAP_ERR_STATE. */
int fap_graph_is_float(fap_graph *g, int32_t *is_float);
/* 1 when preCondition found parallel arcs inside a part whose later arc is heavier:
 * solve() then takes the boundary rows from SSSP (FAP_MODE_FLAG_SSSP), the reference's
 * semantics (its sub-matrix build keeps the later arc, util_Myalgorithm.h:44-84, its
 * Dijkstra the lightest). */
int fap_graph_boundary_by_sssp(fap_graph *g, int32_t *by_sssp);
/* Float32 rows (MAXVALUE = 1e8 for unreachable) of the solve() result / of the last
 * solveSubGraph block, for any graph. */
int fap_graph_get_dist_rows_f32(fap_graph *g, int64_t row0, int64_t rows, float *h_out);
int fap_graph_get_subgraph_rows_f32(fap_graph *g, int32_t id, int64_t row0, int64_t rows,
                                    float *h_out);
/* Device pointer and row stride of the resident solve() result. */
int fap_graph_device_dist(fap_graph *g, int32_t **d_dist, int64_t *ld);
/* Stage times (s): partitioning, sssp, floyd, min-plus, data transformation. */
int fap_graph_times(fap_graph *g, double *t_part, double *t_sssp, double *t_floyd,
                    double *t_minplus, double *t_data);
int fap_graph_edge_cut(fap_graph *g, double *edge_cut);
/* Batches of graphs (after preCondition): fap_hip_batch_create / fap_hip_solve_batch
 * over the graphs' device views and their solve() result buffers, so the graphs'
 * getters return the batch results. */
int fap_graph_batch_create(fap_batch **b, fap_graph **graphs, int32_t count, int32_t flags);
/* The same with per-graph item flags (FAP_ITEM_*, OR-ed into each graph's item; NULL =
 * none), e.g. the Floyd-Warshall tile the autotuner picked for each graph's size class. */
int fap_graph_batch_create_ex(fap_batch **b, fap_graph **graphs, int32_t count, int32_t flags,
                              const int32_t *item_flags);
int fap_graph_solve_batch(fap_graph **graphs, int32_t count, int32_t mode);

/* ---------------------------------------------------------------- serving */
/* A request server over resident graphs (no counterpart in the reference).  Graphs
 * are registered after preCondition and stay resident; requests ask for a full APSP
 * (sub_id 0: the n x n result in the layout of solve()) or one sub-graph block
 * (sub_id = part id: the s_p x n rows of solveSubGraph).  A scheduler thread admits
 * waiting requests in arrival order while their result buffers and batch scratch fit
 * in budget_bytes next to everything admitted and not yet released, puts every
 * admitted request into one new batch (fap_hip_batch_*) on a free stream of its own,
 * and marks requests done when their batch's event has passed.  Results stay in HBM
 * until fap_server_release.  Thread-safe; a request that can never fit fails. */
typedef struct fap_server fap_server;
typedef struct fap_server_stats {
  int64_t submitted, completed, failed, batches;
  int64_t reserved_bytes, peak_reserved_bytes;
  double entries;  /* result entries of completed requests */
  double span_s;   /* first submit -> last completion */
  double p50_ms, p99_ms, mean_ms;  /* submit -> completion latency */
} fap_server_stats;
/* Request states reported by fap_server_wait. */
enum { FAP_REQ_PENDING = 0, FAP_REQ_RUNNING = 1, FAP_REQ_DONE = 2, FAP_REQ_FAILED = 3 };
int fap_server_create(fap_server **s, int64_t budget_bytes, int32_t nstreams);
int fap_server_destroy(fap_server *s); /* drains pending and running requests */
int fap_server_add_graph(fap_server *s, fap_graph *g, int32_t *index);
int fap_server_submit(fap_server *s, int32_t graph, int32_t sub_id, int64_t *ticket);
int fap_server_wait(fap_server *s, int64_t ticket, double timeout_s, int32_t *state);
int fap_server_result(fap_server *s, int64_t ticket, void **d_out, int64_t *rows, int64_t *cols,
                      int64_t *ld, int32_t *is_float);
int fap_server_release(fap_server *s, int64_t ticket);
int fap_server_get_stats(fap_server *s, fap_server_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* FAP_HIP_H_ */
