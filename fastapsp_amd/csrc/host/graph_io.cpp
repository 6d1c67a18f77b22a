// .mtx input and CSR construction (semantics of fap/utils/readMatFile.h:36-173).
#include <cmath>
#include <fstream>
#include <sstream>

#include "fap/fap.h"
#include "graph.h"

namespace fap {

Csr csr_from_arcs(int32_t n, const std::vector<Arc> &arcs, bool directed) {
  // Head insertion into per-vertex lists, then a walk of each list: the CSR of
  // vertex v lists its arcs in reverse order of insertion (readMatFile.h:33,147-162).
  const int64_t m = directed ? (int64_t)arcs.size() : 2 * (int64_t)arcs.size();
  std::vector<int32_t> head(n, -1), nxt(m), to(m);
  std::vector<float> wt(m);
  int64_t idx = 0;
  auto add = [&](int32_t a, int32_t b, float w) {
    to[idx] = b;
    wt[idx] = w;
    nxt[idx] = head[a];
    head[a] = (int32_t)idx++;
  };
  for (const Arc &e : arcs) {
    add(e.a, e.b, e.w);
    if (!directed) add(e.b, e.a, e.w);
  }
  Csr g;
  g.n = n;
  g.adj_size.assign(n, 0);
  g.row_offset.assign(n + 1, 0);
  g.col_val.resize(m);
  g.weight.resize(m);
  int64_t k = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t cnt = 0;
    for (int32_t j = head[i]; j != -1; j = nxt[j]) {
      g.col_val[k] = to[j];
      g.weight[k] = wt[j];
      ++k;
      ++cnt;
    }
    g.adj_size[i] = cnt;
    g.row_offset[i + 1] = g.row_offset[i] + cnt;
  }
  return g;
}

std::string mtx_path(const std::string &file, bool directed, bool weighted) {
  if (file.size() > 4 && file.compare(file.size() - 4, 4, ".mtx") == 0) return file;
  const char *dir = !directed && !weighted  ? "graph/unweight-undirected/"
                    : !directed && weighted ? "graph/weight-undirected/"
                    : directed && !weighted ? "graph/unweight-directed/"
                                            : "graph/weight-directed/";
  return std::string(dir) + file + ".mtx";
}

Csr read_mtx(const std::string &file, bool directed, bool weighted) {
  const std::string path = mtx_path(file, directed, weighted);
  std::ifstream in(path);
  if (!in) throw FapError(FAP_ERR_IO, "cannot open " + path);
  while (in.peek() == '%') in.ignore(1 << 20, '\n');
  long long n = 0, n2 = 0, m = 0;
  if (!(in >> n >> n2 >> m) || n <= 0 || n > INT32_MAX / 2 || m < 0)
    throw FapError(FAP_ERR_IO, "malformed .mtx header in " + path);
  std::vector<Arc> arcs;
  arcs.reserve((size_t)m);
  long long a, b;
  double c;
  while (true) {
    if (weighted) {
      if (!(in >> a >> b >> c)) break;
      c = std::fabs(c);
    } else {
      if (!(in >> a >> b)) break;
      c = 1.0;
    }
    if (a < 1 || a > n || b < 1 || b > n)
      throw FapError(FAP_ERR_IO, "vertex id out of range in " + path);
    arcs.push_back(Arc{(int32_t)(a - 1), (int32_t)(b - 1), (float)c});
  }
  if (!in.eof()) throw FapError(FAP_ERR_IO, "malformed entry in " + path);
  return csr_from_arcs((int32_t)n, arcs, directed);
}

}  // namespace fap
