// The per-part Fast-APSP pipeline behind fap_hip_solve_parts.
//
// Mirrors fapGraph::run (fap/fap.cpp:195-503) for a list of parts:
//   1 boundary SSSP            handle_boundry_path            (fap/fap.cpp:405-409)
//   2 sub-matrix build         MysubMatBuild_path             (fap/fap.cpp:417-423)
//   3 Floyd-Warshall           floyd_path                     (fap/fap.cpp:428)
//   4 min-plus of inner rows   MyMat1Build + min_plus_path_advanced (:437-485)
//   5 decode of own columns    MysubMatDecode_path            (fap/fap.cpp:492-496)
// Step 5 is fused into the min-plus epilogue for the inner rows; for the boundary
// rows it is a small scatter kernel.  Results stay in device memory.
//
// PLAIN   launches the five steps part after part (the reference's structure:
//         solve() calls run() once per part, fap/fap.cpp:512-516).
// BATCHED launches every step once for all parts: one SSSP grid over all boundary
//         sources, one build grid over all rows, three FW launches per round for
//         all parts, one min-plus launch per padded-K group, one decode grid.
// OVERLAP runs SSSP and build as in BATCHED, then splits the parts into chunks
//         on two internal streams, so the FW rounds of one chunk run beside the
//         min-plus of another (measured slower on MI355X: both contend for memory).
//
// With component labels (symmetric graphs) the build writes capped 16-bit
// matrices, Floyd-Warshall runs in the packed 16-bit form (floyd16.hip) and an
// unpack step writes the int32 matrices the merge reads.  After the launches the
// call reads one flag per part; a part whose 16-bit closure reached the cap between
// two vertices of one component is solved again with int32 Floyd-Warshall.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fap_internal.h"
#include "fw_desc.h"

namespace fap_hip {

int launch_sssp(const int *d_row_off, const int *d_col, const int *d_w, int n, const int *d_src,
                int nsrc, const int64_t *d_rows, int *d_out, int64_t ld, bool vec_fill,
                hipStream_t stream);
int launch_fw(const FwDesc *d_descs, int ndesc, int max_nb, hipStream_t stream);
int launch_bg_vert(const BgPart *d_parts, int np, int nb_total, const int *d_pidx_by_pid, int npid,
                   const int *row_off, const int *col, const int *w, const int *graph_id,
                   const int *st2ed, const int *ed2st, BgVert *d_verts, int *d_wt,
                   int64_t *d_orow, hipStream_t stream);
int launch_bg_closure(const BgVert *d_verts, const int *d_wt, int cb_total, bool unit,
                      int nb_total, int Sb, const BgPart *d_parts, int np,
                      const int *d_pidx_by_pid, int npid, const int *row_off, const int *col,
                      const int *w,
                      const int *graph_id, const int *st2ed, const int *ed2st, int *d_mat,
                      hipStream_t stream);
int launch_bg_rows(const BgPart *d_parts, const int64_t *d_orow, int nb_total, const int *d_mat,
                   int Sb, int max_bdy, const int *d_pidx_by_pid, int npid, const int *graph_id,
                   const int *ed2st, int n, int *d_out, int64_t ld_out, int *d_max_b,
                   bool sym, hipStream_t stream);
int launch_fw16(const Fw16Desc *d_descs, int ndesc, int max_nb, hipStream_t stream);
int launch_fw16_unpack(const UnpackDesc *d_descs, int nd, int64_t total_items,
                       const int *d_comp_sorted, int *d_flags, hipStream_t stream);
int launch_comp_sorted(const int *d_comp, const int *d_st2ed, int n, int *d_out,
                       hipStream_t stream);
int launch_ell_build(const int *d_row_off, const int *d_col, int n, int4 *d_ell,
                     hipStream_t stream);
int launch_bfs_lds(const int4 *d_ell, const int *d_row_off, const int *d_col, int n,
                   const int *d_src, int nsrc, const int64_t *d_rows, int *d_out, int64_t ld,
                   bool vec_fill, hipStream_t stream);
int launch_merge(const MergeDesc *d_descs, int nd, int total_tiles, int N, bool vec,
                 const int *d_graph_id, const int *d_ed2st, hipStream_t stream);
int merge_tiles(int M, int N);
int merge_stream_units(int M, int N, const StreamShape &sh);
int merge_stream_row_blocks(int M, const StreamShape &sh);
int merge_stream_kpad(int K);
int launch_merge_stream(const MergeDesc *d_descs, int nd, int total_units, int N, int kp,
                        const int *d_graph_id, const int *d_ed2st, const int *d_max_b, int lim16,
                        const StreamShape &sh, hipStream_t stream);
int launch_inpart_bfs(const InpartDesc *d_descs, int nd, int total_groups, int s_max,
                      hipStream_t stream);
int launch_build(const BuildDesc *d_descs, int nd, int total_rows, bool vec, const int *row_off,
                 const int *col, const int *w, const int *graph_id, const int *st2ed,
                 const int *ed2st, hipStream_t stream);
int launch_decode(const DecodeDesc *d_descs, int nd, int64_t total, const int *st2ed,
                  hipStream_t stream);
int launch_gather_sources(const int *offsets, const int *src_begin, const int64_t *out_rows,
                          int nparts, int total, const int *st2ed, int64_t ld, int *src,
                          int64_t *rows, hipStream_t stream);

constexpr int kTile = 64;
constexpr int kStreamMaxK = 32;  // larger boundary counts use the LDS-tiled product
// Boundary-graph form of the boundary rows (boundary.hip): at most this many boundary
// vertices in all (one workgroup per boundary vertex holds a distance row in LDS).
constexpr int kBgMaxTotal = 4096;

// Merge kernel of a product: 0 = LDS-tiled, else the streaming form with this
// padded K (A rows of the padded FW matrices always hold >= 64 entries).
int merge_kind(int K) { return K <= kStreamMaxK ? merge_stream_kpad(K) : 0; }

struct MergeGroup {
  int kind;   // 0 = tiled, else padded K of the streaming kernel
  int first;  // first descriptor of the group
  int count;  // descriptors in the group
  int units;  // work units (tiles) of the group
  int chunk;  // chunk of parts (BATCHED form)
};

namespace {

Workspace g_ws_fw;    // Floyd-Warshall matrices
Workspace g_ws_fw16;  // 16-bit Floyd-Warshall matrices
Workspace g_ws_meta;  // descriptors, sources, row indices
Workspace g_ws_ell;   // ELL adjacency of the unit-weight graph
Workspace g_ws_bg;    // closed boundary-graph matrix
Workspace g_ws_bgv;   // boundary-graph vertices and clique weights
// Set only inside fap_hip_prepare_parts: the Floyd-Warshall matrices go to this
// caller-owned buffer instead of g_ws_fw and the pass stops after Floyd-Warshall.
int *g_prep_fw = nullptr;

struct EventPool {
  std::vector<hipEvent_t> ev;
  int used = 0;
  int take(hipEvent_t *e) {
    if (used == (int)ev.size()) {
      hipEvent_t x;
      FAP_CHECK_HIP(hipEventCreate(&x));
      ev.push_back(x);
    }
    *e = ev[used++];
    return FAP_OK;
  }
};
EventPool g_events;

// Timed spans of the last call: (stage, start event, end event); stage 0..4 =
// sssp, build, fw, merge, decode.  g_total spans the whole call on the caller stream.
struct Span {
  int stage;
  hipEvent_t a, b;
};
std::vector<Span> g_spans;
hipEvent_t g_total[2] = {nullptr, nullptr};
constexpr int kChunks = 8;
int g_parts16 = 0, g_redone = 0, g_bgraph = 0, g_keyed = 0, g_fw_keyed = 0, g_bfs = 0;
// Packed-merge report: the largest finite boundary-row entry, copied to pinned host
// memory at the end of the pass, against the pass's limit (g_lim16 < 0: no report).
int *g_maxb_host = nullptr;
int g_lim16 = -1;

// What the last pass left for the flag check of the 16-bit form.
struct Pass16 {
  std::vector<fap_part_desc> parts;  // parts solved in the 16-bit form
  const int *d_flags = nullptr;      // one flag per entry of parts
};
Pass16 g_pass16;

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

void timing_note_keyed(int merge_parts, int fw_parts) {
  g_keyed = merge_parts;
  g_fw_keyed = fw_parts;
}
void timing_note_bgraph() { g_bgraph = 1; }

// Stage timing shared with the float-weight pipeline (fp64.hip).
int timing_start(hipStream_t st) {
  g_events.used = 0;
  g_spans.clear();
  g_total[0] = g_total[1] = nullptr;
  g_parts16 = g_redone = g_bgraph = g_keyed = g_fw_keyed = g_bfs = 0;
  g_lim16 = -1;
  FAP_TRY(g_events.take(&g_total[0]));
  FAP_CHECK_HIP(hipEventRecord(g_total[0], st));
  return FAP_OK;
}

int timing_record(hipStream_t on, hipEvent_t *e) {
  FAP_TRY(g_events.take(e));
  FAP_CHECK_HIP(hipEventRecord(*e, on));
  return FAP_OK;
}

void timing_add_span(int stage, hipEvent_t a, hipEvent_t b) { g_spans.push_back(Span{stage, a, b}); }

int timing_finish(hipStream_t st) {
  FAP_TRY(g_events.take(&g_total[1]));
  FAP_CHECK_HIP(hipEventRecord(g_total[1], st));
  return FAP_OK;
}

}  // namespace fap_hip

using namespace fap_hip;

// One pass of the pipeline over `parts`.  append: a second pass of the same call
// (its spans are added to the first pass's, the total spans both).
static int solve_pass(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                      int32_t *d_out, int64_t ld_out, int32_t mode, void *stream, bool append) {
  g_pass16 = Pass16{};
  const int n = g->n;
  FAP_REQUIRE(n > 0 && g->m >= 0 && ld_out >= n && g->d_row_offset && g->d_graph_id &&
                  g->d_st2ed && g->d_ed2st && (g->m == 0 || g->d_col),
              "fap_hip_solve_parts: bad graph view");
  hipStream_t st = resolve(stream);

  // ---- host plan ---------------------------------------------------------
  std::vector<fap_part_desc> act;
  for (int p = 0; p < nparts; ++p) {
    const fap_part_desc &q = parts[p];
    FAP_REQUIRE(q.part_id >= 1 && q.num >= 0 && q.bdy >= 0 && q.bdy <= q.num && q.offset >= 0 &&
                    (int64_t)q.offset + q.num <= n && q.out_row >= 0,
                "fap_hip_solve_parts: bad part %d (id %d off %d num %d bdy %d)", p, q.part_id,
                q.offset, q.num, q.bdy);
    FAP_REQUIRE(q.num <= kMaxPart,
                "fap_hip_solve_parts: part %d has %d vertices (limit %d); use a larger K",
                q.part_id, q.num, kMaxPart);
    if (q.num > 0) act.push_back(q);
  }
  const int P = (int)act.size();
  if (!append) {
    g_spans.clear();
    g_total[0] = g_total[1] = nullptr;
  }
  if (P == 0) return FAP_OK;
  const int form = mode & 0xFF;
  const bool overlap = form == FAP_MODE_OVERLAP;
  const bool batched = form != FAP_MODE_PLAIN;  // one descriptor table for all parts
  // Boundary rows from the boundary graph (boundary.hip) instead of one SSSP per
  // boundary vertex: BATCHED form over parts that cover every vertex, with few
  // boundary vertices.  The Floyd-Warshall matrices then hold in-part distances.
  bool hier = false;
  int bg_total = 0, bg_maxb = 0, max_pid = 0;
  if (form == FAP_MODE_BATCHED && !(mode & FAP_MODE_FLAG_SSSP)) {
    std::vector<std::pair<int, int>> span_of;  // (offset, num) of every non-empty part
    for (const fap_part_desc &q : act) {
      span_of.push_back({q.offset, q.num});
      bg_total += q.bdy;
      bg_maxb = std::max(bg_maxb, q.bdy);
      max_pid = std::max(max_pid, q.part_id);
    }
    std::sort(span_of.begin(), span_of.end());
    bool cover = true;
    int next = 0;
    for (const auto &x : span_of) {
      cover = cover && x.first == next;
      next = x.first + x.second;
    }
    hier = cover && next == n && bg_total <= kBgMaxTotal;
    // More boundary vertices than the single-graph boundary-graph form holds: the
    // batch stages (boundary graph closed by Floyd-Warshall) instead of boundary SSSP.
    if (cover && next == n && !hier && !append) {
      g_events.used = 0;
      g_spans.clear();
      FAP_TRY(timing_record(st, &g_total[0]));
      FAP_TRY(solve_one_batched(g, act.data(), (int)act.size(), d_out, ld_out, false, 3, st));
      return timing_record(st, &g_total[1]);
    }
  }
  // 16-bit Floyd-Warshall: on in-part matrices of unit-weight graphs every finite
  // entry is at most s - 1, below the cap for parts of at most kCap16 vertices, so the
  // cap means "unreachable inside the part" exactly; on SSSP-form matrices of graphs
  // with component labels the
  // cap is checked against the labels (and the part redone in int32 if needed).
  // Bit-parallel in-part BFS (inpart.hip) instead of build + Floyd-Warshall: the
  // boundary-graph form needs in-part distances only, which on unit weights with
  // paired arcs are BFS distances.
  int s_max_all = 0;
  for (const fap_part_desc &q : act) s_max_all = std::max(s_max_all, q.num);
  const bool bfs_parts = hier && g->d_w == nullptr && g->d_comp != nullptr &&
                         !(mode & (FAP_MODE_FLAG_INT32_FW | FAP_MODE_FLAG_FW)) &&
                         s_max_all <= kInpartBfsMaxS;
  // In the boundary-graph form the 16-bit closure has no cap check: it needs every
  // in-part distance (at most s - 1 for unit weights) below the cap.
  const bool use16 = !(mode & FAP_MODE_FLAG_INT32_FW) && !bfs_parts &&
                     (hier ? g->d_w == nullptr && s_max_all <= kCap16 : g->d_comp != nullptr);
  const bool flags16 = use16 && !hier;
  const int nchunks = overlap ? std::min(kChunks, P) : 1;
  std::vector<int> chunk_of(P);
  for (int p = 0; p < P; ++p) chunk_of[p] = (int)((int64_t)p * nchunks / P);
  const bool vec = (ld_out % 4 == 0) && ((uintptr_t)d_out % 16 == 0);

  std::vector<int> S(P), nb(P);
  std::vector<int64_t> fw_off(P);
  int64_t fw_total = 0, fw_max = 0;  // entries
  int max_nb = 0;
  for (int p = 0; p < P; ++p) {
    S[p] = (int)align_up(act[p].num, kTile);
    nb[p] = S[p] / kTile;
    max_nb = std::max(max_nb, nb[p]);
    const int64_t sz = (int64_t)S[p] * S[p];
    fw_off[p] = batched ? fw_total : 0;
    fw_total += sz;
    fw_max = std::max(fw_max, sz);
  }
  int *fw_base = g_prep_fw;
  if (!fw_base) {
    FAP_TRY(g_ws_fw.ensure((batched ? fw_total : fw_max) * 4));
    fw_base = static_cast<int *>(g_ws_fw.ptr);
  }
  uint16_t *fw16_base = nullptr;
  if (use16) {
    FAP_TRY(g_ws_fw16.ensure((batched ? fw_total : fw_max) * 2));
    fw16_base = static_cast<uint16_t *>(g_ws_fw16.ptr);
  }

  std::vector<BuildDesc> bd(P);
  std::vector<FwDesc> fd(P);
  std::vector<MergeDesc> md(P);
  std::vector<DecodeDesc> dd(P);
  std::vector<Fw16Desc> fd16(P);
  std::vector<UnpackDesc> ud(P);
  // In-part BFS in two launches: the source groups holding each part's boundary
  // vertices (all the boundary stage reads, the in-part matrices being symmetric),
  // then the rest, which runs beside the boundary stage.
  std::vector<InpartDesc> ipa, ipb;
  int ipa_groups = 0, ipb_groups = 0;
  std::vector<int> offsets(P), src_begin(P);
  std::vector<int64_t> out_rows(P);
  int rows_total = 0, src_total = 0;
  int64_t dec_total = 0, unpack_total = 0;
  std::vector<int64_t> chunk_dec(nchunks, 0), chunk_unpack(nchunks, 0);
  for (int p = 0; p < P; ++p) {
    const fap_part_desc &q = act[p];
    int *fw = fw_base + fw_off[p];
    uint16_t *fw16 = use16 ? fw16_base + fw_off[p] : nullptr;
    const int M = q.num - q.bdy;
    bd[p] = BuildDesc{fw, S[p], fw16, S[p], q.num, q.bdy, q.offset, q.part_id,
                      batched ? rows_total : 0, hier ? nullptr : d_out + q.out_row * ld_out,
                      ld_out};
    fd[p] = FwDesc{fw, S[p], nb[p]};
    fd16[p] = Fw16Desc{fw16, S[p], nb[p]};
    MergeDesc m{};
    m.A = fw + (int64_t)q.bdy * S[p];
    m.lda = S[p];
    m.B = d_out + q.out_row * ld_out;
    m.ldb = ld_out;
    m.C = d_out + (q.out_row + q.bdy) * ld_out;
    m.ldc = ld_out;
    m.M = M;
    m.K = q.bdy;
    m.row_blocks = 0;  // set when the merge groups are formed
    m.tile_begin = 0;
    m.part_id = q.part_id;
    m.offset = q.offset;
    m.fw = fw;
    m.ldf = S[p];
    m.fw_row0 = q.bdy;
    md[p] = m;
    if (p > 0 && chunk_of[p] != chunk_of[p - 1]) dec_total = unpack_total = 0;
    ud[p] = UnpackDesc{fw16, fw, S[p], q.num, q.offset, p, batched ? unpack_total : 0};
    {  // nb[p] = ceil(s / 64) groups of 64 sources
      const int ga = std::min(nb[p], (q.bdy + 63) / 64);
      if (ga > 0)
        ipa.push_back(InpartDesc{fw, S[p], q.num, q.offset, q.part_id, ipa_groups, 0,
                                 g->d_row_offset, g->d_col, g->d_graph_id, g->d_st2ed,
                                 g->d_ed2st});
      if (nb[p] > ga)
        ipb.push_back(InpartDesc{fw, S[p], q.num, q.offset, q.part_id, ipb_groups, ga,
                                 g->d_row_offset, g->d_col, g->d_graph_id, g->d_st2ed,
                                 g->d_ed2st});
      ipa_groups += ga;
      ipb_groups += nb[p] - ga;
    }
    unpack_total += (int64_t)S[p] * S[p] / 8;
    chunk_unpack[chunk_of[p]] += (int64_t)S[p] * S[p] / 8;
    dd[p] = DecodeDesc{fw, S[p], d_out + q.out_row * ld_out, ld_out, q.bdy, q.num, q.offset,
                       batched ? dec_total : 0};
    offsets[p] = q.offset;
    src_begin[p] = batched ? src_total : 0;
    out_rows[p] = q.out_row;
    rows_total += S[p];
    src_total += q.bdy;
    dec_total += (int64_t)q.bdy * q.num;
    chunk_dec[chunk_of[p]] += (int64_t)q.bdy * q.num;
  }

  // ---- merge groups: products with M > 0, ordered by kernel kind ------------
  std::vector<MergeDesc> mdg;
  std::vector<MergeGroup> groups;
  std::vector<int> part_merge(P, -1);  // plain form: the group of part p
  const StreamShape shape = stream_shape();  // one snapshot for the plan and the launches
  {
    std::vector<int> order;
    for (int p = 0; p < P; ++p)
      if (md[p].M > 0) order.push_back(p);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      if (chunk_of[a] != chunk_of[b]) return chunk_of[a] < chunk_of[b];
      return merge_kind(md[a].K) < merge_kind(md[b].K);
    });
    for (int p : order) {
      MergeDesc m = md[p];
      const int kind = merge_kind(m.K);
      const bool stream_kernel = kind > 0;
      m.row_blocks = stream_kernel ? merge_stream_row_blocks(m.M, shape) : (m.M + 63) / 64;
      const int64_t units = stream_kernel ? merge_stream_units(m.M, n, shape) : merge_tiles(m.M, n);
      const bool new_group = !batched || groups.empty() || groups.back().kind != kind ||
                             groups.back().chunk != chunk_of[p];
      if (new_group) groups.push_back(MergeGroup{kind, (int)mdg.size(), 0, 0, chunk_of[p]});
      MergeGroup &gr = groups.back();
      FAP_REQUIRE((int64_t)gr.units + units < (1LL << 30), "fap_hip_solve_parts: too many tiles");
      m.tile_begin = gr.units;
      gr.units += (int)units;
      gr.count += 1;
      part_merge[p] = (int)groups.size() - 1;
      mdg.push_back(m);
    }
  }
  const int nmd = (int)mdg.size();

  // ---- upload descriptors (one buffer) -------------------------------------
  int64_t off = 0;
  auto carve = [&](int64_t bytes) {
    const int64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const int64_t o_bd = carve(sizeof(BuildDesc) * P), o_fd = carve(sizeof(FwDesc) * P);
  const int64_t o_md = carve(sizeof(MergeDesc) * std::max(nmd, 1));
  const int64_t o_dd = carve(sizeof(DecodeDesc) * P);
  const int64_t o_fd16 = carve(sizeof(Fw16Desc) * P), o_ud = carve(sizeof(UnpackDesc) * P);
  const int64_t o_ip = carve(sizeof(InpartDesc) * 2 * P);
  const int64_t o_off = carve(4LL * P), o_sb = carve(4LL * P), o_or = carve(8LL * P);
  const int64_t o_src = carve(4LL * std::max(src_total, 1));
  const int64_t o_rows = carve(8LL * std::max(src_total, 1));
  // Boundary graph: part table, part id -> table index, the matrix.
  std::vector<BgPart> bgp;
  std::vector<int> pidx_by_pid;
  const int Sb = (int)align_up(bg_total, kTile);
  int *bg_mat = nullptr, *bg_wt = nullptr;
  BgVert *bg_verts = nullptr;
  int64_t *bg_orow = nullptr;
  int bg_cb = 0;  // clique weights of all G_B vertices
  if (hier) {
    FAP_TRY(g_ws_bg.ensure(std::max<int64_t>(1, (int64_t)Sb * Sb) * 4));
    bg_mat = static_cast<int *>(g_ws_bg.ptr);
    pidx_by_pid.assign(max_pid + 1, 0);
    int bs = 0, cb = 0;
    for (int p = 0; p < P; ++p) {
      const fap_part_desc &q = act[p];
      bgp.push_back(
          BgPart{fw_base + fw_off[p], S[p], q.bdy, q.offset, q.part_id, bs, cb, q.out_row});
      pidx_by_pid[q.part_id] = p;
      bs += q.bdy;
      cb += q.bdy * q.bdy;  // <= 4096^2
    }
    bg_cb = cb;
    const int64_t vbytes = align_up((int64_t)sizeof(BgVert) * std::max(bg_total, 1), 256);
    const int64_t obytes = align_up(8LL * std::max(bg_total, 1), 256);
    FAP_TRY(g_ws_bgv.ensure(vbytes + obytes + 4LL * (cb + 64)));  // + slack read past the end
    char *base = static_cast<char *>(g_ws_bgv.ptr);
    bg_verts = reinterpret_cast<BgVert *>(base);
    bg_orow = reinterpret_cast<int64_t *>(base + vbytes);
    bg_wt = reinterpret_cast<int *>(base + vbytes + obytes);
  }
  const int64_t o_bgp = carve(sizeof(BgPart) * std::max<size_t>(bgp.size(), 1));
  const int64_t o_pidx = carve(4LL * std::max<size_t>(pidx_by_pid.size(), 1));
  const int64_t o_flags = carve(4LL * P), o_comp = carve(4LL * n);  // device-written
  const int64_t o_maxb = carve(4);
  FAP_TRY(g_ws_meta.ensure(off));
  char *meta = static_cast<char *>(g_ws_meta.ptr);
  std::vector<char> host(off);
  std::memcpy(host.data() + o_bd, bd.data(), sizeof(BuildDesc) * P);
  std::memcpy(host.data() + o_fd, fd.data(), sizeof(FwDesc) * P);
  if (nmd) std::memcpy(host.data() + o_md, mdg.data(), sizeof(MergeDesc) * nmd);
  std::memcpy(host.data() + o_dd, dd.data(), sizeof(DecodeDesc) * P);
  std::memcpy(host.data() + o_fd16, fd16.data(), sizeof(Fw16Desc) * P);
  std::memcpy(host.data() + o_ud, ud.data(), sizeof(UnpackDesc) * P);
  if (!ipa.empty()) std::memcpy(host.data() + o_ip, ipa.data(), sizeof(InpartDesc) * ipa.size());
  if (!ipb.empty())
    std::memcpy(host.data() + o_ip + sizeof(InpartDesc) * P, ipb.data(),
                sizeof(InpartDesc) * ipb.size());
  std::memcpy(host.data() + o_off, offsets.data(), 4LL * P);
  std::memcpy(host.data() + o_sb, src_begin.data(), 4LL * P);
  std::memcpy(host.data() + o_or, out_rows.data(), 8LL * P);
  if (hier) {
    std::memcpy(host.data() + o_bgp, bgp.data(), sizeof(BgPart) * bgp.size());
    std::memcpy(host.data() + o_pidx, pidx_by_pid.data(), 4 * pidx_by_pid.size());
  }
  FAP_CHECK_HIP(hipMemcpyAsync(meta, host.data(), o_flags, hipMemcpyHostToDevice, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  const BuildDesc *d_bd = reinterpret_cast<const BuildDesc *>(meta + o_bd);
  const FwDesc *d_fd = reinterpret_cast<const FwDesc *>(meta + o_fd);
  const MergeDesc *d_md = reinterpret_cast<const MergeDesc *>(meta + o_md);
  const DecodeDesc *d_dd = reinterpret_cast<const DecodeDesc *>(meta + o_dd);
  const int *d_offsets = reinterpret_cast<const int *>(meta + o_off);
  const int *d_sb = reinterpret_cast<const int *>(meta + o_sb);
  const int64_t *d_or = reinterpret_cast<const int64_t *>(meta + o_or);
  int *d_src = reinterpret_cast<int *>(meta + o_src);
  int64_t *d_rows = reinterpret_cast<int64_t *>(meta + o_rows);
  const Fw16Desc *d_fd16 = reinterpret_cast<const Fw16Desc *>(meta + o_fd16);
  const UnpackDesc *d_ud = reinterpret_cast<const UnpackDesc *>(meta + o_ud);
  const InpartDesc *d_ipa = reinterpret_cast<const InpartDesc *>(meta + o_ip);
  const InpartDesc *d_ipb = d_ipa + P;
  hipEvent_t ipb_done = nullptr;  // the second in-part BFS launch (internal stream)
  const BgPart *d_bgp = reinterpret_cast<const BgPart *>(meta + o_bgp);
  const int *d_pidx = reinterpret_cast<const int *>(meta + o_pidx);
  int *d_flags = reinterpret_cast<int *>(meta + o_flags);
  int *d_maxb = reinterpret_cast<int *>(meta + o_maxb);
  // Packed 16-bit merge (unit weights, boundary-graph form): in-part distances are
  // at most s_max - 1, so the merge may run packed while every finite boundary-row
  // entry is at most kCap16 - s_max (checked on the device against d_maxb).
  const bool merge16 = hier && g->d_w == nullptr;
  int s_max = 0;
  for (const fap_part_desc &q : act) s_max = std::max(s_max, q.num);
  const int lim16 = kCap16 - s_max;
  int *d_comp_sorted = reinterpret_cast<int *>(meta + o_comp);

  // ---- launches -------------------------------------------------------------
  if (!append) {
    g_events.used = 0;
    g_spans.clear();
  }
  auto event = [&](hipStream_t on, hipEvent_t *e) -> int {
    FAP_TRY(g_events.take(e));
    FAP_CHECK_HIP(hipEventRecord(*e, on));
    return FAP_OK;
  };
  // Runs fn on stream `on` inside a timed span of `stage`.
  auto span = [&](int stage, hipStream_t on, auto &&fn) -> int {
    Span sp{stage, nullptr, nullptr};
    FAP_TRY(event(on, &sp.a));
    FAP_TRY(fn(on));
    FAP_TRY(event(on, &sp.b));
    g_spans.push_back(sp);
    return FAP_OK;
  };
  if (!append) FAP_TRY(event(st, &g_total[0]));
  if (flags16) {
    FAP_TRY(fap_hip_fill_i32(d_flags, 0, P, st));
    FAP_TRY(launch_comp_sorted(g->d_comp, g->d_st2ed, n, d_comp_sorted, st));
  }
  // Floyd-Warshall of parts [p0, p1) with max tile count mnb; unpack_items: the
  // prefix end of those parts' unpack items.
  auto run_fw = [&](int p0, int p1, int mnb, int64_t unpack_items, hipStream_t on) -> int {
    if (bfs_parts) {  // boundary-graph form: one chunk, all parts
      if (ipb_groups > 0) {
        const hipStream_t side = aux_stream(1);
        hipEvent_t fork;
        FAP_TRY(event(on, &fork));
        FAP_CHECK_HIP(hipStreamWaitEvent(side, fork, 0));
        FAP_TRY(launch_inpart_bfs(d_ipb, (int)ipb.size(), ipb_groups, s_max_all, side));
        FAP_TRY(event(side, &ipb_done));
      }
      return launch_inpart_bfs(d_ipa, (int)ipa.size(), ipa_groups, s_max_all, on);
    }
    if (!use16) return launch_fw(d_fd + p0, p1 - p0, mnb, on);
    FAP_TRY(launch_fw16(d_fd16 + p0, p1 - p0, mnb, on));
    return launch_fw16_unpack(d_ud + p0, p1 - p0, unpack_items,
                              flags16 ? d_comp_sorted : nullptr, d_flags, on);
  };

  // Boundary SSSP: the LDS-bitmap BFS for unit weights, else label correcting.
  int4 *ell = nullptr;
  if (!g->d_w && !hier) {
    FAP_TRY(g_ws_ell.ensure((int64_t)n * sizeof(int4)));
    ell = static_cast<int4 *>(g_ws_ell.ptr);
    FAP_TRY(launch_ell_build(g->d_row_offset, g->d_col, n, ell, st));
  }
  auto run_sssp = [&](const int *src, int nsrc, hipStream_t on) -> int {
    if (ell) {
      const int s = launch_bfs_lds(ell, g->d_row_offset, g->d_col, n, src, nsrc, d_rows, d_out, 1,
                                   vec, on);
      if (s != FAP_ERR_INVALID) return s;
    }
    return launch_sssp(g->d_row_offset, g->d_col, g->d_w, n, src, nsrc, d_rows, d_out, 1, vec, on);
  };
  auto run_group = [&](const MergeGroup &gr, hipStream_t on) -> int {
    if (gr.kind == 0)
      return launch_merge(d_md + gr.first, gr.count, gr.units, n, vec, g->d_graph_id, g->d_ed2st,
                          on);
    return launch_merge_stream(d_md + gr.first, gr.count, gr.units, n, gr.kind, g->d_graph_id,
                               g->d_ed2st,
                               merge16 && bg_total > 0 ? d_maxb : nullptr, lim16, shape, on);
  };

  // Boundary graph: vertices, closure (its distance matrix), then every part's
  // boundary rows into d_out.
  auto run_boundary_graph = [&](hipStream_t on) -> int {
    if (bg_total == 0) return FAP_OK;
    FAP_TRY(fap_hip_fill_i32(d_maxb, 0, 1, on));
    const int npid = (int)pidx_by_pid.size();
    FAP_TRY(launch_bg_vert(d_bgp, P, bg_total, d_pidx, npid, g->d_row_offset, g->d_col, g->d_w,
                           g->d_graph_id, g->d_st2ed, g->d_ed2st, bg_verts, bg_wt, bg_orow, on));
    // Unit weights: in-part distances below s_max < 2^16, so the clique weights fit
    // 16 bits.
    FAP_TRY(launch_bg_closure(bg_verts, bg_wt, bg_cb, g->d_w == nullptr && s_max < 0xFFFF,
                              bg_total, Sb, d_bgp, P, d_pidx, npid, g->d_row_offset, g->d_col,
                              g->d_w,
                              g->d_graph_id, g->d_st2ed, g->d_ed2st, bg_mat, on));
    // Paired arcs (component labels) of unit weight: the in-part matrices are symmetric.
    return launch_bg_rows(d_bgp, bg_orow, bg_total, bg_mat, Sb, bg_maxb, d_pidx, npid, g->d_graph_id,
                          g->d_ed2st, n, d_out, ld_out, d_maxb,
                          g->d_comp != nullptr && g->d_w == nullptr, on);
  };

  if (batched) {
    if (!hier)
      FAP_TRY(span(0, st, [&](hipStream_t on) -> int {
        FAP_TRY(launch_gather_sources(d_offsets, d_sb, d_or, P, src_total, g->d_st2ed, ld_out,
                                      d_src, d_rows, on));
        return run_sssp(d_src, src_total, on);
      }));
    if (!bfs_parts)  // the BFS writes the whole in-part matrices itself
      FAP_TRY(span(1, st, [&](hipStream_t on) {
        return launch_build(d_bd, P, rows_total, true, g->d_row_offset, g->d_col, g->d_w,
                            g->d_graph_id, g->d_st2ed, g->d_ed2st, on);
      }));
    // Chunks of parts: [first part, end part) with their FW rounds, merges and decode.
    std::vector<int> cfirst(nchunks + 1, P);
    for (int p = P - 1; p >= 0; --p) cfirst[chunk_of[p]] = p;
    hipEvent_t ready = nullptr;
    if (overlap) FAP_TRY(event(st, &ready));
    std::vector<hipEvent_t> done;
    for (int c = 0; c < nchunks; ++c) {
      const int p0 = cfirst[c], p1 = cfirst[c + 1];
      if (p0 >= p1) continue;
      hipStream_t on = overlap ? aux_stream(c & 1) : st;
      if (overlap) FAP_CHECK_HIP(hipStreamWaitEvent(on, ready, 0));
      int cmax_nb = 0;
      int64_t cdec = 0;
      for (int p = p0; p < p1; ++p) cmax_nb = std::max(cmax_nb, nb[p]);
      cdec = overlap ? chunk_dec[c] : dec_total;
      const int64_t cunpack = overlap ? chunk_unpack[c] : unpack_total;
      FAP_TRY(span(2, on, [&](hipStream_t s2) { return run_fw(p0, p1, cmax_nb, cunpack, s2); }));
      if (g_prep_fw) continue;  // prepared parts: no merge, no decode
      if (hier) FAP_TRY(span(0, on, run_boundary_graph));  // hier: one chunk
      if (ipb_done) FAP_CHECK_HIP(hipStreamWaitEvent(on, ipb_done, 0));
      FAP_TRY(span(3, on, [&](hipStream_t s2) -> int {
        if (overlap) {
          for (const MergeGroup &gr : groups)
            if (gr.chunk == c) FAP_TRY(run_group(gr, s2));
          return FAP_OK;
        }
        // One launch per K-group, largest first, spread over the caller stream and
        // the two internal streams so one launch's tail overlaps the next launch
        // (merge stage 11.6 -> 10.7 ms on MI355X against one stream).
        std::vector<int> order;
        for (int k = 0; k < (int)groups.size(); ++k) order.push_back(k);
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return groups[x].units > groups[y].units; });
        const hipStream_t lanes[3] = {s2, aux_stream(0), aux_stream(1)};
        const int nl = std::min<int>(3, (int)order.size());
        hipEvent_t fork;
        FAP_TRY(event(s2, &fork));
        for (int l = 1; l < nl; ++l) FAP_CHECK_HIP(hipStreamWaitEvent(lanes[l], fork, 0));
        for (int k = 0; k < (int)order.size(); ++k) FAP_TRY(run_group(groups[order[k]], lanes[k % nl]));
        for (int l = 1; l < nl; ++l) {
          hipEvent_t joined;
          FAP_TRY(event(lanes[l], &joined));
          FAP_CHECK_HIP(hipStreamWaitEvent(s2, joined, 0));
        }
        return FAP_OK;
      }));
      if (!hier)  // hier: the boundary rows already hold their own columns
        FAP_TRY(span(4, on, [&](hipStream_t s2) {
          return launch_decode(d_dd + p0, p1 - p0, cdec, g->d_st2ed, s2);
        }));
      if (overlap) {
        hipEvent_t e;
        FAP_TRY(event(on, &e));
        done.push_back(e);
      }
    }
    for (hipEvent_t e : done) FAP_CHECK_HIP(hipStreamWaitEvent(st, e, 0));
  } else {
    for (int p = 0; p < P; ++p) {
      const fap_part_desc &q = act[p];
      FAP_TRY(span(0, st, [&](hipStream_t on) -> int {
        FAP_TRY(launch_gather_sources(d_offsets + p, d_sb + p, d_or + p, 1, q.bdy, g->d_st2ed,
                                      ld_out, d_src, d_rows, on));
        return run_sssp(d_src, q.bdy, on);
      }));
      FAP_TRY(span(1, st, [&](hipStream_t on) {
        return launch_build(d_bd + p, 1, S[p], true, g->d_row_offset, g->d_col, g->d_w,
                            g->d_graph_id, g->d_st2ed, g->d_ed2st, on);
      }));
      FAP_TRY(span(2, st, [&](hipStream_t on) {
        return run_fw(p, p + 1, nb[p], (int64_t)S[p] * S[p] / 8, on);
      }));
      FAP_TRY(span(3, st, [&](hipStream_t on) {
        return part_merge[p] >= 0 ? run_group(groups[part_merge[p]], on) : FAP_OK;
      }));
      FAP_TRY(span(4, st, [&](hipStream_t on) {
        return launch_decode(d_dd + p, 1, (int64_t)q.bdy * q.num, g->d_st2ed, on);
      }));
    }
  }
  if (merge16 && bg_total > 0 && !append) {
    if (!g_maxb_host) FAP_CHECK_HIP(hipHostMalloc((void **)&g_maxb_host, sizeof(int)));
    FAP_CHECK_HIP(hipMemcpyAsync(g_maxb_host, d_maxb, sizeof(int), hipMemcpyDeviceToHost, st));
    g_lim16 = lim16;
  }
  FAP_TRY(event(st, &g_total[1]));
  if (use16) g_parts16 += P;
  if (bfs_parts) g_bfs += P;
  if (hier) g_bgraph = 1;
  if (flags16) {
    g_pass16.parts = act;
    g_pass16.d_flags = d_flags;
  }
  return FAP_OK;
}

static int solve_parts(const fap_graph_dev *g, const fap_part_desc *parts, int32_t nparts,
                       int32_t *d_out, int64_t ld_out, int32_t mode, void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && nparts >= 0 && (nparts == 0 || parts) && d_out,
              "fap_hip_solve_parts: null argument");
  const int form = mode & 0xFF;
  FAP_REQUIRE((form == FAP_MODE_PLAIN || form == FAP_MODE_BATCHED || form == FAP_MODE_OVERLAP) &&
                  (mode & ~(0xFF | FAP_MODE_FLAG_INT32_FW | FAP_MODE_FLAG_SSSP |
                            FAP_MODE_FLAG_FW)) == 0,
              "fap_hip_solve_parts: unknown mode %d", mode);
  g_parts16 = g_redone = g_bgraph = g_keyed = g_fw_keyed = g_bfs = 0;
  g_lim16 = -1;
  FAP_TRY(solve_pass(g, parts, nparts, d_out, ld_out, mode, stream, false));
  if (g_pass16.parts.empty()) return FAP_OK;
  // Flag check of the 16-bit form: parts whose closure reached the cap inside one
  // component are solved again with int32 Floyd-Warshall.
  const Pass16 pass = g_pass16;
  hipStream_t st = resolve(stream);
  std::vector<int> flags(pass.parts.size());
  FAP_CHECK_HIP(hipMemcpyAsync(flags.data(), pass.d_flags, 4 * flags.size(),
                               hipMemcpyDeviceToHost, st));
  FAP_CHECK_HIP(hipStreamSynchronize(st));
  std::vector<fap_part_desc> redo;
  for (size_t i = 0; i < flags.size(); ++i)
    if (flags[i]) redo.push_back(pass.parts[i]);
  if (redo.empty()) return FAP_OK;
  g_redone = (int)redo.size();
  const int parts16 = g_parts16;
  FAP_TRY(solve_pass(g, redo.data(), (int32_t)redo.size(), d_out, ld_out,
                     mode | FAP_MODE_FLAG_INT32_FW, stream, true));
  g_parts16 = parts16;
  return FAP_OK;
}

extern "C" int fap_hip_solve_parts(const fap_graph_dev *g, const fap_part_desc *parts,
                                   int32_t nparts, int32_t *d_out, int64_t ld_out, int32_t mode,
                                   void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  const int s = solve_parts(g, parts, nparts, d_out, ld_out, mode, stream);
  const int f = ws_fence_record(st);  // also after a failed call: its launches used the workspaces
  return s != FAP_OK ? s : f;
}

extern "C" int fap_hip_prepare_parts(const fap_graph_dev *g, const fap_part_desc *parts,
                                     int32_t nparts, int32_t *d_rows, int64_t ld_rows,
                                     int32_t *d_fw, int64_t *fw_off, int64_t *fw_entries,
                                     void *stream) {
  if (!initialized()) {
    set_error("fap_hip_init has not been called");
    return FAP_ERR_STATE;
  }
  FAP_REQUIRE(g && nparts >= 0 && (nparts == 0 || parts) && fw_off && fw_entries,
              "fap_hip_prepare_parts: null argument");
  // Same offsets as the batched pass: non-empty parts in list order, S x S each.
  int64_t total = 0;
  for (int p = 0; p < nparts; ++p) {
    if (parts[p].num <= 0) {
      fw_off[p] = -1;
      continue;
    }
    const int64_t S = align_up(parts[p].num, kTile);
    fw_off[p] = total;
    total += S * S;
  }
  *fw_entries = total;
  if (!d_fw) return FAP_OK;  // size query
  FAP_REQUIRE(d_rows && ld_rows >= g->n, "fap_hip_prepare_parts: bad boundary-row buffer");
  hipStream_t st = resolve(stream);
  FAP_TRY(ws_fence_wait(st));
  g_parts16 = g_redone = g_bgraph = g_keyed = g_fw_keyed = g_bfs = 0;
  g_lim16 = -1;
  g_prep_fw = d_fw;
  const int s = solve_pass(g, parts, nparts, d_rows, ld_rows,
                           FAP_MODE_BATCHED | FAP_MODE_FLAG_SSSP | FAP_MODE_FLAG_INT32_FW, stream,
                           false);
  g_prep_fw = nullptr;
  const int f = ws_fence_record(st);
  return s != FAP_OK ? s : f;
}

extern "C" int fap_hip_last_stage_times(fap_stage_times *t) {
  FAP_REQUIRE(t, "fap_hip_last_stage_times: null");
  fap_stage_times acc{};
  if (g_total[1]) {
    FAP_CHECK_HIP(hipEventSynchronize(g_total[1]));
    float *slot[5] = {&acc.sssp_ms, &acc.build_ms, &acc.floyd_ms, &acc.minplus_ms, &acc.decode_ms};
    for (const Span &sp : g_spans) {
      float v;
      FAP_CHECK_HIP(hipEventElapsedTime(&v, sp.a, sp.b));
      *slot[sp.stage] += v;
    }
    FAP_CHECK_HIP(hipEventElapsedTime(&acc.total_ms, g_total[0], g_total[1]));
  }
  acc.parts_fw16 = g_parts16;
  acc.parts_redone = g_redone;
  acc.boundary_graph = g_bgraph;
  acc.parts_keyed = g_keyed;
  acc.parts_fw_keyed = g_fw_keyed;
  acc.parts_bfs = g_bfs;
  acc.merge_packed = g_lim16 < 0 || !g_total[1] ? -1 : (*g_maxb_host <= g_lim16 ? 1 : 0);
  *t = acc;
  return FAP_OK;
}
