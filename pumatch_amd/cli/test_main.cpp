// `test <graph-prefix> <pattern.g>`: the reference's entry point (test.cu:15-212)
// on libpumatch.so.  Same two positional arguments, same final stdout line
// "<ms> <count>" (test.cu:200) and the same result.txt record
// "<graph>\t<pattern>\t<ms>\t<count>" (test.cu:202-209).  LABELED is a
// runtime switch whose default matches src/config.h:14 (true).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/pumatch.h"

static void usage() {
  std::fprintf(stderr,
               "usage: test <graph-prefix> <pattern.g> [--unlabeled] [--vertex-induced]\n"
               "            [--waves N] [--chunk N] [--repeat N] [--no-result-file]\n");
}

int main(int argc, char** argv) {
  if (argc < 3) {
    usage();
    return 2;
  }
  int labeled = 1, induced = 0, write_result = 1;
  pm_opts opts;
  pm_opts_default(&opts);
  for (int i = 3; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--unlabeled") labeled = 0;
    else if (a == "--labeled") labeled = 1;
    else if (a == "--vertex-induced") induced = 1;  // EDGE_INDUCED=false
    else if (a == "--no-result-file") write_result = 0;
    else if (a == "--waves" && i + 1 < argc) opts.waves = std::atoi(argv[++i]);
    else if (a == "--chunk" && i + 1 < argc) opts.chunk = std::atoi(argv[++i]);
    else if (a == "--repeat" && i + 1 < argc) opts.repeats = std::atoi(argv[++i]);
    else {
      usage();
      return 2;
    }
  }
  pm_graph* g = nullptr;
  int rc = pm_graph_open(argv[1], labeled ? PM_GRAPH_LABELS : 0, &g);
  if (rc) {
    std::fprintf(stderr, "graph %s: %s\n", argv[1], pm_last_error());
    return 1;
  }
  pm_graph_info gi;
  pm_graph_info_get(g, &gi);
  std::printf("graph: %llu vertices, %llu directed edges, max degree %u\n",
              (unsigned long long)gi.n, (unsigned long long)gi.m, gi.max_degree);
  pm_pattern* p = nullptr;
  rc = pm_pattern_load(argv[2], (labeled ? PM_LABELED : 0) | (induced ? PM_VERTEX_INDUCED : 0),
                       &p);
  if (rc) {
    std::fprintf(stderr, "pattern %s: %s\n", argv[2], pm_last_error());
    return 1;
  }
  pm_result r;
  rc = pm_match(g, p, &opts, &r);
  if (rc) {
    std::fprintf(stderr, "match: %s\n", pm_last_error());
    return 1;
  }
  std::printf("jobs: %llu  set ops: %llu  waves: %d\n", (unsigned long long)r.jobs,
              (unsigned long long)r.set_ops, r.waves);
  std::printf("%f %llu\n", r.kernel_ms, (unsigned long long)r.count);
  if (write_result) {
    FILE* f = std::fopen("result.txt", "a");
    if (f) {
      std::fprintf(f, "%s\t%s\t%f\t%llu\n", argv[1], argv[2], r.kernel_ms,
                   (unsigned long long)r.count);
      std::fclose(f);
    } else {
      std::fprintf(stderr, "cannot append to result.txt\n");
    }
  }
  pm_pattern_destroy(p);
  pm_graph_destroy(g);
  return 0;
}
