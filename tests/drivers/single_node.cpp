// A caller shaped like the reference's single-node driver (app/singleNodeExample.cpp:
// 57-93): the (file, directed, weighted, K, partitioner, version, endTimeIo)
// constructor, preCondition, isSplit, solveSubGraph(1..3, true) in split mode or
// solve() otherwise, the sub-graph getters and check_result -- compiled against
// include/fap/fap.h and linked with libfap_hip.so, nothing else.  In non-split mode
// the reference's driver reads the sub-graph getters of id 0 (empty blocks) after
// solve(); this one solves with paths and checks the whole matrix instead.
//
// Extra flags for the parity test: -dump FILE writes the last solved block
// (int32 header: rows, n, sub graph id; then the source ids, the float distances
// and the int32 predecessors) so the test can compare it with the oracle; -p2p
// also exercises getDistanceP2P / getPathP2P and the getMemCons* accounting.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fap/fap.h"

int main(int argc, char **argv) {
  std::string file, partitioner = "metis", dump;
  int K = 8;
  bool directed = false, weighted = false, version = true, p2p = false;
  for (int i = 1; i < argc; ++i) {
    const bool has = i + 1 < argc;
    if (!std::strcmp(argv[i], "-f") && has) file = argv[++i];
    else if (!std::strcmp(argv[i], "-k") && has) K = std::stoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-direct") && has) directed = !std::strcmp(argv[++i], "true");
    else if (!std::strcmp(argv[i], "-weight") && has) weighted = !std::strcmp(argv[++i], "true");
    else if (!std::strcmp(argv[i], "-partitioner") && has) partitioner = argv[++i];
    else if (!std::strcmp(argv[i], "-version") && has) version = !std::strcmp(argv[++i], "true");
    else if (!std::strcmp(argv[i], "-dump") && has) dump = argv[++i];
    else if (!std::strcmp(argv[i], "-p2p")) p2p = true;
  }
  try {
    auto t0 = std::chrono::high_resolution_clock::now();
    std::chrono::time_point<std::chrono::high_resolution_clock> endTimeIo;
    fap::fapGraph G(file, directed, weighted, K, partitioner, version, endTimeIo);
    G.preCondition();
    if (G.isSplit() && K >= 3) {
      std::printf("Processing first three subgraphs in split mode\n");
      for (int id = 1; id <= 3; ++id) {
        G.solveSubGraph(id, true);
        std::printf("Subgraph %d processing complete\n", id);
      }
    } else {
      std::printf("Processing all subgraphs in non-split mode\n");
      G.solve(true);
      std::printf("All subgraphs processing complete\n");
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("Total runtime: %g seconds.\n", secs);

    int id = G.getCurrentSubGraphId();
    std::vector<float> dist;
    std::vector<int32_t> path, source;
    if (id > 0) {
      dist = G.getSubGraphDistance(id);
      path = G.getSubGraphPath(id);
      source = G.getSubGraphIndex(id);
    } else {  // non-split: the whole matrix, rows in sorted order
      dist = G.getAllDistance();
      path = G.getAllPath();
      source = G.getMapping();
    }
    const bool ok = G.check_result(dist.data(), path.data(), source.data(), (int)source.size(),
                                   (int)G.adj_size.size(), G.adj_size.data(),
                                   G.row_offset.data(), G.col_val.data(), G.weight.data(),
                                   G.getGraphId().data());
    if (ok) std::printf("the %d subGraph is right\n", id);
    else std::printf("the %d subGraph is wrong !!!\n", id);

    if (p2p) {
      const int u = source[source.size() / 2], v = G.num_vertexs - 1;
      const float d = G.getDistanceP2P(u, v);
      const std::vector<int32_t> route = G.getPathP2P(u, v);
      std::printf("p2p %d %d %.1f %zu", u, v, d, route.size());
      for (int32_t x : route) std::printf(" %d", x);
      std::printf("\n");
      std::printf("memcons %.6f %.6f %.6f\n", G.getMemConsSSSP(), G.getMemConsFloyd(),
                  G.getMemConsMinPlus());
      bool threw = false;
      try {
        (void)G.getDistanceP2P();
      } catch (const fap::FapError &) {
        threw = true;
      }
      std::printf("p2p_noargs_throws %d\n", threw ? 1 : 0);
    }
    if (!dump.empty()) {
      FILE *f = std::fopen(dump.c_str(), "wb");
      if (!f) return 3;
      const int32_t hdr[3] = {(int32_t)source.size(), G.num_vertexs, id};
      std::fwrite(hdr, 4, 3, f);
      std::fwrite(source.data(), 4, source.size(), f);
      std::fwrite(dist.data(), 4, dist.size(), f);
      std::fwrite(path.data(), 4, path.size(), f);
      std::fclose(f);
    }
    return ok ? 0 : 1;
  } catch (const fap::FapError &e) {
    std::fprintf(stderr, "FapError %d: %s\n", e.status(), e.what());
    return 2;
  }
}
