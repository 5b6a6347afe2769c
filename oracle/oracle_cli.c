/*
 * pm_oracle <graph-prefix> <pattern.g> [--unlabeled] [--vertex-induced] [--threads N]
 *
 * TEST INFRASTRUCTURE ONLY: the CPU oracle behind the reference's CLI shape
 * (test.cu:15-212).  It reads the CSR files the way
 * GraphPreprocessor::read_bin_file does (src/graph.h:228-254) and the .g
 * file the way PatternPreprocessor::readfile does (src/pattern.h:67-100),
 * then prints "<ms> <count>" (test.cu:200).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OR_MAXK 7
typedef struct {
  int k;
  int adj[OR_MAXK][OR_MAXK];
  int vlabel[OR_MAXK];
} or_pattern;

int or_count(uint64_t n, uint64_t m, const uint64_t* rowptr, const uint32_t* colidx,
             const uint32_t* labels, const or_pattern* p, int labeled, uint64_t edge_lo,
             uint64_t edge_hi, int threads, uint64_t* per_edge, uint64_t* out_count,
             uint64_t* out_jobs);

static void* slurp(const char* path, size_t bytes) {
  FILE* f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "Failed to open file: %s\n", path); exit(1); }
  void* p = malloc(bytes ? bytes : 1);
  if (bytes && fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "short read: %s\n", path); exit(1); }
  fclose(f);
  return p;
}

/* src/pattern.h:67-100 */
static int read_pattern(const char* path, int labeled, or_pattern* p) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[512];
  memset(p, 0, sizeof(*p));
  int state = 0; /* 0 header, 1 vertices, 2 edges */
  while (fgets(line, sizeof line, f)) {
    if (state == 0 && line[0] == '#') continue;
    if (state <= 1 && line[0] == 'v') {
      char t; int id, lab;
      if (sscanf(line, " %c %d %d", &t, &id, &lab) != 3 || p->k >= OR_MAXK) { fclose(f); return -1; }
      p->vlabel[p->k++] = labeled ? lab : 1;
      state = 1;
      continue;
    }
    state = 2;
    char t; int a, b, lab;
    if (sscanf(line, " %c %d %d %d", &t, &a, &b, &lab) != 4) continue;
    if (a < 0 || b < 0 || a >= p->k || b >= p->k) { fclose(f); return -1; }
    p->adj[a][b] = lab;
    p->adj[b][a] = lab;
  }
  fclose(f);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: pm_oracle <graph-prefix> <pattern.g> [--unlabeled] [--vertex-induced] "
                    "[--threads N]\n");
    return 2;
  }
  int labeled = 1, induced = 0, threads = 0;
  for (int i = 3; i < argc; i++) {
    if (!strcmp(argv[i], "--unlabeled")) labeled = 0;
    else if (!strcmp(argv[i], "--vertex-induced")) induced = 1;
    else if (!strcmp(argv[i], "--threads") && i + 1 < argc) threads = atoi(argv[++i]);
  }
  char path[4096];
  snprintf(path, sizeof path, "%s.meta.txt", argv[1]);
  FILE* mf = fopen(path, "r");
  if (!mf) { fprintf(stderr, "Failed to open file: %s\n", path); return 1; }
  long long n, m, vid, md;
  if (fscanf(mf, "%lld %lld %lld %lld", &n, &m, &vid, &md) != 4 || vid != 4) { fclose(mf); return 1; }
  fclose(mf);
  snprintf(path, sizeof path, "%s.vertex.bin", argv[1]);
  uint64_t* rowptr = (uint64_t*)slurp(path, sizeof(uint64_t) * (n + 1));
  snprintf(path, sizeof path, "%s.edge.bin", argv[1]);
  uint32_t* colidx = (uint32_t*)slurp(path, sizeof(uint32_t) * m);
  uint32_t* labels = NULL;
  if (labeled) {
    snprintf(path, sizeof path, "%s.label.bin", argv[1]);
    labels = (uint32_t*)slurp(path, sizeof(uint32_t) * n);
  }
  or_pattern p;
  if (read_pattern(argv[2], labeled, &p)) { fprintf(stderr, "bad pattern %s\n", argv[2]); return 1; }
  uint64_t count = 0, jobs = 0;
  double t0 = omp_get_wtime();
  int rc = or_count((uint64_t)n, (uint64_t)m, rowptr, colidx, labels, &p, labeled | (induced << 1), 0, 0, threads,
                    NULL, &count, &jobs);
  double t1 = omp_get_wtime();
  if (rc) { fprintf(stderr, "oracle failed (%d)\n", rc); return 1; }
  printf("jobs: %llu threads: %d\n", (unsigned long long)jobs, threads > 0 ? threads : omp_get_max_threads());
  printf("%f %llu\n", (t1 - t0) * 1e3, (unsigned long long)count);
  return 0;
}
