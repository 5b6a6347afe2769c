// Host CSR: the reference's on-disk format and seeded synthetic generators.
//
// On-disk layout (src/graph.h:228-254, written by
// graph_converter/compactify.cpp:10-51 and labelize.cpp:39-51):
//   <prefix>.meta.txt   "n m 4 maxdeg"
//   <prefix>.vertex.bin u64[n+1] row offsets
//   <prefix>.edge.bin   u32[m]   sorted, deduplicated neighbour ids
//   <prefix>.label.bin  u32[n]   vertex labels
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "pm_internal.hpp"

namespace pm {

namespace {

// splitmix64 finaliser: the counter-based stream every generator draws from,
// so a draw depends only on (seed, index) and not on the thread count.
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline uint64_t stream(uint64_t seed, uint64_t idx, uint64_t sub) {
  return mix64(mix64(seed * 0xD1B54A32D192ED03ull + sub) ^ (idx * 0x9E3779B97F4A7C15ull));
}
inline double unit53(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

void finalize_graph(HostGraph* g) {
  uint32_t md = 0;
#pragma omp parallel for reduction(max : md) schedule(static)
  for (int64_t v = 0; v < (int64_t)g->n; v++) {
    uint64_t d = g->rowptr[v + 1] - g->rowptr[v];
    if (d > md) md = (uint32_t)d;
  }
  g->max_degree = md;
}

// out[v] = max degree over N(v): the job filter's "some w in N(c) with
// deg(w) >= d_P(2)" (src/job_queue.h:48-54) becomes one comparison.
void compute_max_neighbour_degree(const HostGraph& g, std::vector<uint32_t>* out) {
  out->assign(g.n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)g.n; v++) {
    uint32_t mx = 0;
    for (uint64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; e++) {
      const uint32_t w = g.colidx[e];
      const uint32_t d = (uint32_t)(g.rowptr[w + 1] - g.rowptr[w]);
      mx = d > mx ? d : mx;
    }
    (*out)[v] = mx;
  }
}

void build_label_grouped(const HostGraph& g, std::vector<uint32_t>* lcol,
                         std::vector<uint32_t>* loff, uint32_t* nl_out) {
  uint32_t maxl = 0;
  for (uint32_t l : g.labels) maxl = l > maxl ? l : maxl;
  const uint32_t nl = maxl + 2;
  *nl_out = nl;
  lcol->resize(g.m);
  loff->assign((size_t)g.n * nl, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < (int64_t)g.n; v++) {
    uint32_t* o = loff->data() + (size_t)v * nl;
    const uint64_t b = g.rowptr[v], e = g.rowptr[v + 1];
    for (uint64_t i = b; i < e; i++) o[g.labels[g.colidx[i]] + 1]++;
    for (uint32_t l = 1; l < nl; l++) o[l] += o[l - 1];
    std::vector<uint32_t> cur(o, o + nl);
    for (uint64_t i = b; i < e; i++) {
      const uint32_t x = g.colidx[i];
      (*lcol)[b + cur[g.labels[x]]++] = x;
    }
  }
}

// pairs = (src << 32 | dst) for directed edges, any order, duplicates and
// self-loops allowed; builds the sorted, deduplicated, loop-free CSR.
int build_csr_from_pairs(HostGraph* g, uint64_t n, std::vector<uint64_t>& pairs) {
  const uint64_t np = pairs.size();
  std::vector<std::atomic<uint64_t>> cnt(n + 1);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v <= (int64_t)n; v++) cnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)np; i++) {
    uint64_t s = pairs[i] >> 32, d = pairs[i] & 0xFFFFFFFFull;
    if (s != d) cnt[s].fetch_add(1, std::memory_order_relaxed);
  }
  std::vector<uint64_t> off(n + 1);
  uint64_t acc = 0;
  for (uint64_t v = 0; v < n; v++) {
    off[v] = acc;
    acc += cnt[v].load(std::memory_order_relaxed);
  }
  off[n] = acc;
  std::vector<uint32_t> raw(acc);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v <= (int64_t)n; v++) cnt[v].store(off[v], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)np; i++) {
    uint64_t s = pairs[i] >> 32, d = pairs[i] & 0xFFFFFFFFull;
    if (s != d) raw[cnt[s].fetch_add(1, std::memory_order_relaxed)] = (uint32_t)d;
  }
  std::vector<uint64_t>().swap(pairs);
  std::vector<uint64_t> deg(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < (int64_t)n; v++) {
    uint32_t* b = raw.data() + off[v];
    uint32_t* e = raw.data() + off[v + 1];
    std::sort(b, e);
    deg[v] = (uint64_t)(std::unique(b, e) - b);
  }
  g->n = n;
  g->rowptr.assign(n + 1, 0);
  for (uint64_t v = 0; v < n; v++) g->rowptr[v + 1] = g->rowptr[v] + deg[v];
  g->m = g->rowptr[n];
  g->colidx.resize(g->m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < (int64_t)n; v++)
    std::memcpy(g->colidx.data() + g->rowptr[v], raw.data() + off[v], sizeof(uint32_t) * deg[v]);
  finalize_graph(g);
  return PM_OK;
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004) with Graph500 conventions:
// 2^scale vertices, edge_factor * 2^scale generated edges, each edge placed by
// `scale` recursive quadrant choices with probabilities (a, b, c, 1-a-b-c).
// No vertex permutation.  Every edge is stored in both directions.
int generate_rmat(HostGraph* g, int scale, double edge_factor, double a, double b, double c,
                  uint64_t seed) {
  if (scale < 1 || scale > 31 || edge_factor <= 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return PM_ERR_ARG;
  const uint64_t n = 1ull << scale;
  const uint64_t ne = (uint64_t)std::llround(edge_factor * (double)n);
  const double ab = a + b, abc = a + b + c;
  std::vector<uint64_t> pairs(2 * ne);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)ne; i++) {
    uint64_t s = 0, d = 0;
    for (int l = 0; l < scale; l++) {
      double u = unit53(stream(seed, (uint64_t)i, (uint64_t)l));
      uint64_t sb = 0, db = 0;
      if (u < a) {
      } else if (u < ab) {
        db = 1;
      } else if (u < abc) {
        sb = 1;
      } else {
        sb = 1;
        db = 1;
      }
      s = (s << 1) | sb;
      d = (d << 1) | db;
    }
    pairs[2 * i] = (s << 32) | d;
    pairs[2 * i + 1] = (d << 32) | s;
  }
  return build_csr_from_pairs(g, n, pairs);
}

// Chung-Lu: vertex i has weight w_i = dmax * (1 + i/i0)^(-1/(gamma-1)) with
// i0 chosen so the mean weight is `avg`; n*avg/2 edges, each endpoint drawn
// with probability proportional to its weight (Walker alias table).
int generate_chunglu(HostGraph* g, uint64_t n, double gamma, double dmax, double avg,
                     uint64_t seed) {
  if (n < 2 || n > 0x7FFFFFFFull || gamma <= 1.0 || dmax <= 0 || avg <= 0 || avg > dmax)
    return PM_ERR_ARG;
  const double ex = -1.0 / (gamma - 1.0);
  auto mean_for = [&](double i0) {
    // mean of dmax*(1+x/i0)^ex over x in [0,n): closed form of the integral
    double t = ex + 1.0;
    return dmax * i0 / (t * (double)n) * (std::pow(1.0 + (double)n / i0, t) - 1.0);
  };
  double lo = 1e-6, hi = (double)n * 1e6;
  for (int it = 0; it < 200; it++) {
    double mid = std::sqrt(lo * hi);
    if (mean_for(mid) < avg) lo = mid; else hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  std::vector<double> w(n);
  double tot = 0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    w[i] = dmax * std::pow(1.0 + (double)i / i0, ex);
    tot += w[i];
  }
  // alias table (Vose)
  std::vector<float> prob(n);
  std::vector<uint32_t> alias(n);
  {
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    std::vector<double> sc(n);
    for (uint64_t i = 0; i < n; i++) {
      sc[i] = w[i] * (double)n / tot;
      (sc[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      uint32_t s = small.back(); small.pop_back();
      uint32_t l = large.back();
      prob[s] = (float)sc[s];
      alias[s] = l;
      sc[l] = (sc[l] + sc[s]) - 1.0;
      if (sc[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
    for (uint32_t i : large) { prob[i] = 1.0f; alias[i] = i; }
    for (uint32_t i : small) { prob[i] = 1.0f; alias[i] = i; }
  }
  std::vector<double>().swap(w);
  const uint64_t ne = (uint64_t)std::llround((double)n * avg / 2.0);
  std::vector<uint64_t> pairs(2 * ne);
  auto draw = [&](uint64_t r) -> uint64_t {
    uint64_t col = (r >> 32) % n;  // n < 2^31: bias below 2^-32 relative
    double u = (double)(r & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
    return u < (double)prob[col] ? col : alias[col];
  };
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)ne; i++) {
    uint64_t s = draw(stream(seed, (uint64_t)i, 0));
    uint64_t d = draw(stream(seed, (uint64_t)i, 1));
    pairs[2 * i] = (s << 32) | d;
    pairs[2 * i + 1] = (d << 32) | s;
  }
  return build_csr_from_pairs(g, n, pairs);
}

void assign_uniform_labels(HostGraph* g, uint32_t nlabels, uint64_t seed) {
  g->labels.resize(g->n);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)g->n; v++)
    g->labels[v] = (uint32_t)(stream(seed, (uint64_t)v, 0x1abe1) % nlabels);
}

}  // namespace pm

// ---------------------------------------------------------------------------
// file I/O (C ABI entry points live in api.cpp; these are the bodies)
// ---------------------------------------------------------------------------
namespace pm {

static bool read_exact(const std::string& path, void* dst, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  size_t got = 0;
  char* p = (char*)dst;
  while (got < bytes) {
    size_t r = std::fread(p + got, 1, bytes - got, f);
    if (r == 0) break;
    got += r;
  }
  std::fclose(f);
  return got == bytes;
}

static bool write_exact(const std::string& path, const void* src, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = std::fwrite(src, 1, bytes, f) == bytes;
  ok = (std::fclose(f) == 0) && ok;
  return ok;
}

int read_graph_files(const std::string& prefix, bool labels, HostGraph* g) {
  std::ifstream meta(prefix + ".meta.txt");
  if (!meta.good()) {
    set_error("cannot open " + prefix + ".meta.txt");
    return PM_ERR_IO;
  }
  long long n = -1, m = -1, vid = -1, maxdeg = -1;
  meta >> n >> m >> vid >> maxdeg;
  if (!meta || n < 0 || m < 0 || vid != 4) {
    set_error("malformed " + prefix + ".meta.txt (expected 'n m 4 maxdeg')");
    return PM_ERR_FORMAT;
  }
  g->n = (uint64_t)n;
  g->m = (uint64_t)m;
  g->rowptr.resize(g->n + 1);
  g->colidx.resize(g->m);
  if (!read_exact(prefix + ".vertex.bin", g->rowptr.data(), sizeof(uint64_t) * (g->n + 1))) {
    set_error("cannot read " + prefix + ".vertex.bin");
    return PM_ERR_IO;
  }
  if (!read_exact(prefix + ".edge.bin", g->colidx.data(), sizeof(uint32_t) * g->m)) {
    set_error("cannot read " + prefix + ".edge.bin");
    return PM_ERR_IO;
  }
  if (g->rowptr[0] != 0 || g->rowptr[g->n] != g->m) {
    set_error("row offsets do not span the edge file");
    return PM_ERR_FORMAT;
  }
  for (uint64_t v = 0; v < g->n; v++)
    if (g->rowptr[v + 1] < g->rowptr[v]) {
      set_error("row offsets decrease");
      return PM_ERR_FORMAT;
    }
  bool bad = false;
#pragma omp parallel for reduction(|| : bad) schedule(static)
  for (int64_t e = 0; e < (int64_t)g->m; e++) bad = bad || g->colidx[e] >= g->n;
  if (bad) {
    set_error("neighbour id out of range");
    return PM_ERR_FORMAT;
  }
  if (labels) {
    g->labels.resize(g->n);
    if (!read_exact(prefix + ".label.bin", g->labels.data(), sizeof(uint32_t) * g->n)) {
      set_error("cannot read " + prefix + ".label.bin");
      return PM_ERR_IO;
    }
  }
  finalize_graph(g);
  return PM_OK;
}

int write_graph_files(const std::string& prefix, const HostGraph& g) {
  {
    std::ofstream meta(prefix + ".meta.txt");
    if (!meta.good()) {
      set_error("cannot write " + prefix + ".meta.txt");
      return PM_ERR_IO;
    }
    // one field per line, as graph_converter/compactify.cpp:45-48 writes it
    meta << g.n << "\n" << g.m << "\n4\n" << g.max_degree << "\n";
  }
  if (!write_exact(prefix + ".vertex.bin", g.rowptr.data(), sizeof(uint64_t) * (g.n + 1)) ||
      !write_exact(prefix + ".edge.bin", g.colidx.data(), sizeof(uint32_t) * g.m)) {
    set_error("cannot write CSR files for " + prefix);
    return PM_ERR_IO;
  }
  if (!g.labels.empty() &&
      !write_exact(prefix + ".label.bin", g.labels.data(), sizeof(uint32_t) * g.n)) {
    set_error("cannot write " + prefix + ".label.bin");
    return PM_ERR_IO;
  }
  return PM_OK;
}

}  // namespace pm
