// `convert <snap.txt> <prefix>`: SNAP edge list -> the matcher's binary CSR
// (<prefix>.meta.txt / .vertex.bin / .edge.bin / .label.bin), in one pass
// and without temporary files.  Same result as the reference's pipeline
// graph_converter/convert.sh (snapToBin.cpp, makeLists.cpp, compactify.cpp,
// labelize.cpp):
//   * lines that are empty or start with '#' are skipped, each other line
//     gives "src dst" (snapToBin.cpp:12-20);
//   * vertex ids are renumbered in order of first appearance, src before dst
//     (snapToBin.cpp:22-31);
//   * every edge is stored in both directions; rows are sorted and
//     deduplicated (makeLists.cpp, compactify.cpp:10-20);
//   * meta = "n m 4 max_degree" (compactify.cpp:45-48);
//   * labels = rand() % 10 after srand(0) (labelize.cpp:45-50).
// One deliberate difference: self-loops are dropped (the reference keeps
// them as a vertex in its own row, which the matcher assumes absent);
// --keep-self-loops keeps them.  --labels L --seed S draws uniform labels
// from the library's generator instead.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pumatch.h"

static void usage() {
  std::fprintf(stderr,
               "usage: convert <snap.txt> <out-prefix> [--keep-self-loops] [--labels L --seed S]\n");
}

int main(int argc, char** argv) {
  if (argc < 3) {
    usage();
    return 2;
  }
  bool keep_loops = false;
  uint32_t nlabels = 0;
  uint64_t seed = 0;
  for (int i = 3; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--keep-self-loops") keep_loops = true;
    else if (a == "--labels" && i + 1 < argc) nlabels = (uint32_t)std::atoi(argv[++i]);
    else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
    else {
      usage();
      return 2;
    }
  }
  FILE* f = std::fopen(argv[1], "r");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  std::unordered_map<uint64_t, uint32_t> id;
  std::vector<uint32_t> src, dst;
  auto get = [&](uint64_t raw) {
    auto it = id.find(raw);
    if (it != id.end()) return it->second;
    const uint32_t v = (uint32_t)id.size();
    id.emplace(raw, v);
    return v;
  };
  char line[1 << 12];
  uint64_t lines = 0;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n' || line[0] == '\r' || line[0] == 0) continue;
    char* p = line;
    char* e1 = nullptr;
    char* e2 = nullptr;
    const uint64_t a = std::strtoull(p, &e1, 10);
    if (e1 == p) continue;
    const uint64_t b = std::strtoull(e1, &e2, 10);
    if (e2 == e1) continue;
    const uint32_t u = get(a);
    const uint32_t v = get(b);
    src.push_back(u);
    dst.push_back(v);
    lines++;
  }
  std::fclose(f);
  if (id.size() > 0xFFFFFFF0ull) {
    std::fprintf(stderr, "too many vertices for 32-bit ids\n");
    return 1;
  }
  const uint64_t n = id.size();
  std::vector<uint64_t> rowptr(n + 1, 0);
  for (size_t i = 0; i < src.size(); i++) {
    if (src[i] == dst[i] && !keep_loops) continue;
    rowptr[src[i] + 1]++;
    if (src[i] != dst[i]) rowptr[dst[i] + 1]++;
  }
  for (uint64_t v = 0; v < n; v++) rowptr[v + 1] += rowptr[v];
  std::vector<uint32_t> col(rowptr[n]);
  std::vector<uint64_t> fill(rowptr.begin(), rowptr.end() - 1);
  for (size_t i = 0; i < src.size(); i++) {
    if (src[i] == dst[i] && !keep_loops) continue;
    col[fill[src[i]]++] = dst[i];
    if (src[i] != dst[i]) col[fill[dst[i]]++] = src[i];
  }
  // sort and deduplicate each row, then compact
  std::vector<uint64_t> out_ptr(n + 1, 0);
  uint64_t w = 0;
  for (uint64_t v = 0; v < n; v++) {
    const uint64_t b = rowptr[v], e = rowptr[v + 1];
    std::sort(col.begin() + (std::ptrdiff_t)b, col.begin() + (std::ptrdiff_t)e);
    uint32_t prev = 0;
    bool first = true;
    for (uint64_t i = b; i < e; i++) {
      if (!first && col[i] == prev) continue;
      col[w++] = col[i];
      prev = col[i];
      first = false;
    }
    out_ptr[v + 1] = w;
  }
  col.resize(w);
  std::vector<uint32_t> labels(n);
  if (nlabels == 0) {
    std::srand(0);  // the reference's labels: rand() % 10 in vertex order
    for (uint64_t v = 0; v < n; v++) labels[v] = (uint32_t)(std::rand() % 10);
  }
  pm_graph* g = nullptr;
  int rc = pm_graph_from_csr(n, w, out_ptr.data(), col.empty() ? nullptr : col.data(), labels.data(),
                             PM_GRAPH_HOST_ONLY, &g);
  if (rc == PM_OK && nlabels) rc = pm_graph_set_uniform_labels(g, nlabels, seed);
  if (rc == PM_OK) rc = pm_graph_write(g, argv[2]);
  if (rc != PM_OK) {
    std::fprintf(stderr, "convert failed: %s\n", pm_last_error());
    if (g) pm_graph_destroy(g);
    return 1;
  }
  pm_graph_info gi;
  pm_graph_info_get(g, &gi);
  std::printf("%llu lines, %llu vertices, %llu directed edges, max degree %u\n",
              (unsigned long long)lines, (unsigned long long)gi.n, (unsigned long long)gi.m,
              gi.max_degree);
  pm_graph_destroy(g);
  return 0;
}
