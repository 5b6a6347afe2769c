// C ABI of libpumatch.so (declared in include/pumatch.h).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>

#include "pm_internal.hpp"

namespace pm {
int read_graph_files(const std::string& prefix, bool labels, HostGraph* g);
int write_graph_files(const std::string& prefix, const HostGraph& g);
int device_update_labels(pm_graph* g);

namespace {
thread_local std::string t_err;
}
void set_error(const std::string& msg) { t_err = msg; }
const char* last_error() { return t_err.c_str(); }
}  // namespace pm

using namespace pm;

namespace {

int finish_graph(pm_graph* g, uint32_t flags, pm_graph** out) {
  if (!(flags & PM_GRAPH_HOST_ONLY)) {
    int rc = device_upload(g);
    if (rc) {
      delete g;
      return rc;
    }
  }
  *out = g;
  return PM_OK;
}

int check_labels(const HostGraph& h) {
  for (uint32_t l : h.labels)
    if (l > 31) {
      set_error("vertex label above 31 (labels are used as 1<<label, src/graph.h:251)");
      return PM_ERR_FORMAT;
    }
  return PM_OK;
}

}  // namespace

extern "C" {

const char* pm_last_error(void) { return pm::last_error(); }
const char* pm_version(void) { return "pumatch-amd 0.1 (gfx950)"; }

int pm_graph_open(const char* prefix, uint32_t flags, pm_graph** out) {
  if (!prefix || !out) return PM_ERR_ARG;
  pm_graph* g = new (std::nothrow) pm_graph();
  if (!g) return PM_ERR_NOMEM;
  int rc = read_graph_files(prefix, (flags & PM_GRAPH_LABELS) != 0, &g->h);
  if (rc == PM_OK) rc = check_labels(g->h);
  if (rc) {
    delete g;
    return rc;
  }
  return finish_graph(g, flags, out);
}

int pm_graph_from_csr(uint64_t n, uint64_t m, const uint64_t* rowptr, const uint32_t* colidx,
                      const uint32_t* labels, uint32_t flags, pm_graph** out) {
  if (!rowptr || (!colidx && m) || !out || n >= 0xFFFFFFFFull) return PM_ERR_ARG;
  if (rowptr[0] != 0 || rowptr[n] != m) {
    set_error("rowptr must start at 0 and end at m");
    return PM_ERR_FORMAT;
  }
  pm_graph* g = new (std::nothrow) pm_graph();
  if (!g) return PM_ERR_NOMEM;
  g->h.n = n;
  g->h.m = m;
  g->h.rowptr.assign(rowptr, rowptr + n + 1);
  g->h.colidx.assign(colidx, colidx + m);
  if (labels) g->h.labels.assign(labels, labels + n);
  for (uint64_t v = 0; v < n; v++) {
    if (g->h.rowptr[v + 1] < g->h.rowptr[v]) {
      delete g;
      set_error("row offsets decrease");
      return PM_ERR_FORMAT;
    }
    for (uint64_t e = g->h.rowptr[v]; e < g->h.rowptr[v + 1]; e++) {
      if (g->h.colidx[e] >= n || (e > g->h.rowptr[v] && g->h.colidx[e] <= g->h.colidx[e - 1])) {
        delete g;
        set_error("adjacency lists must be strictly increasing vertex ids < n");
        return PM_ERR_FORMAT;
      }
    }
  }
  int rc = check_labels(g->h);
  if (rc) {
    delete g;
    return rc;
  }
  finalize_graph(&g->h);
  return finish_graph(g, flags, out);
}

int pm_graph_generate_rmat(int32_t scale, double edge_factor, double a, double b, double c,
                           uint64_t seed, uint32_t flags, pm_graph** out) {
  if (!out) return PM_ERR_ARG;
  pm_graph* g = new (std::nothrow) pm_graph();
  if (!g) return PM_ERR_NOMEM;
  int rc = generate_rmat(&g->h, scale, edge_factor, a, b, c, seed);
  if (rc) {
    delete g;
    set_error("bad R-MAT parameters");
    return rc;
  }
  return finish_graph(g, flags, out);
}

int pm_graph_generate_chunglu(uint64_t n, double gamma, double dmax, double avg, uint64_t seed,
                              uint32_t flags, pm_graph** out) {
  if (!out) return PM_ERR_ARG;
  pm_graph* g = new (std::nothrow) pm_graph();
  if (!g) return PM_ERR_NOMEM;
  int rc = generate_chunglu(&g->h, n, gamma, dmax, avg, seed);
  if (rc) {
    delete g;
    set_error("bad Chung-Lu parameters");
    return rc;
  }
  return finish_graph(g, flags, out);
}

int pm_graph_set_uniform_labels(pm_graph* g, uint32_t nlabels, uint64_t seed) {
  if (!g || nlabels == 0 || nlabels > 32) return PM_ERR_ARG;
  assign_uniform_labels(&g->h, nlabels, seed);
  g->stats_ok = false;
  return device_update_labels(g);
}

int pm_graph_write(const pm_graph* g, const char* prefix) {
  if (!g || !prefix) return PM_ERR_ARG;
  return write_graph_files(prefix, g->h);
}

int pm_graph_info_get(const pm_graph* g, pm_graph_info* info) {
  if (!g || !info) return PM_ERR_ARG;
  std::memset(info, 0, sizeof(*info));
  info->n = g->h.n;
  info->m = g->h.m;
  info->max_degree = g->h.max_degree;
  info->has_labels = g->h.labels.empty() ? 0 : 1;
  info->on_device = g->on_device ? 1 : 0;
  info->device_bytes = g->on_device ? g->d.bytes : 0;
  return PM_OK;
}

int pm_graph_host_csr(const pm_graph* g, const uint64_t** rowptr, const uint32_t** colidx,
                      const uint32_t** labels) {
  if (!g) return PM_ERR_ARG;
  if (rowptr) *rowptr = g->h.rowptr.data();
  if (colidx) *colidx = g->h.colidx.data();
  if (labels) *labels = g->h.labels.empty() ? nullptr : g->h.labels.data();
  return PM_OK;
}

int pm_graph_to_device(pm_graph* g) {
  if (!g) return PM_ERR_ARG;
  return device_upload(g);
}

void pm_graph_destroy(pm_graph* g) {
  if (!g) return;
  device_release(g);
  delete g;
}

int pm_pattern_parse(const char* text, int32_t labeled, pm_pattern** out) {
  if (!text || !out) return PM_ERR_ARG;
  if (labeled & ~(PM_LABELED | PM_VERTEX_INDUCED)) {
    set_error("pattern mode: unknown bits (PM_LABELED | PM_VERTEX_INDUCED)");
    return PM_ERR_ARG;
  }
  pm_pattern* p = new (std::nothrow) pm_pattern();
  if (!p) return PM_ERR_NOMEM;
  std::string err;
  int rc = parse_pattern_text(text, labeled, &p->plan, &err);
  if (rc == PM_OK) rc = build_plan(&p->plan, &err);
  if (rc) {
    set_error(err);
    delete p;
    return rc;
  }
  *out = p;
  return PM_OK;
}

int pm_pattern_load(const char* path, int32_t labeled, pm_pattern** out) {
  if (!path || !out) return PM_ERR_ARG;
  std::ifstream f(path);
  if (!f.good()) {
    set_error(std::string("cannot open pattern file ") + path);
    return PM_ERR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  int rc = pm_pattern_parse(ss.str().c_str(), labeled, out);
  if (rc == PM_OK) (*out)->path = path;
  return rc;
}

int pm_pattern_info_get(const pm_pattern* p, pm_pattern_info* info) {
  if (!p || !info) return PM_ERR_ARG;
  std::memset(info, 0, sizeof(*info));
  const Plan& pl = p->plan;
  info->k = pl.k;
  info->labeled = pl.labeled;
  info->automorphisms = pl.aut;
  info->rc_order = pl.rc_order;
  info->nslots = pl.nslots;
  for (int u = 0; u < pl.k; u++) {
    info->degree[u] = pl.deg[u];
    info->label[u] = pl.vlabel[u];
    info->sym_mask[u] = pl.B[u];
    info->nbr_mask[u] = pl.A[u];
  }
  info->vertex_induced = pl.induced;
  return PM_OK;
}

void pm_pattern_destroy(pm_pattern* p) { delete p; }

void pm_opts_default(pm_opts* o) {
  if (o) std::memset(o, 0, sizeof(*o));
}

int pm_match(pm_graph* g, const pm_pattern* p, const pm_opts* opts, pm_result* res) {
  if (!g || !p || !res) return PM_ERR_ARG;
  pm_opts def;
  pm_opts_default(&def);
  if (!opts) opts = &def;
  if (p->plan.labeled && g->h.labels.empty()) {
    set_error("labeled matching needs vertex labels (PM_GRAPH_LABELS)");
    return PM_ERR_ARG;
  }
  return device_match(g, p->plan, opts, res);
}

int pm_match_batch(pm_graph* g, const pm_pattern* const* p, int32_t n, const pm_opts* opts,
                   pm_result* results) {
  if (!g || !p || n < 0 || (n && !results)) return PM_ERR_ARG;
  if (n == 0) return PM_OK;
  pm_opts def;
  pm_opts_default(&def);
  if (!opts) opts = &def;
  std::vector<const Plan*> plans((size_t)n);
  for (int32_t i = 0; i < n; i++) {
    if (!p[i]) return PM_ERR_ARG;
    plans[i] = &p[i]->plan;
  }
  return device_match_batch(g, plans.data(), n, opts, results);
}

int pm_model_create(uint64_t seed, pm_model** out) {
  if (!out) return PM_ERR_ARG;
  pm_model* m = new (std::nothrow) pm_model();
  if (!m) return PM_ERR_NOMEM;
  model_init(m, seed);
  *out = m;
  return PM_OK;
}

void pm_model_destroy(pm_model* m) { delete m; }

int pm_model_get(const pm_model* m, float* params) {
  if (!m || !params) return PM_ERR_ARG;
  std::memcpy(params, m->w.data(), sizeof(float) * PM_MODEL_PARAMS);
  return PM_OK;
}

int pm_model_set(pm_model* m, const float* params) {
  if (!m || !params) return PM_ERR_ARG;
  std::memcpy(m->w.data(), params, sizeof(float) * PM_MODEL_PARAMS);
  std::fill(m->adam_m.begin(), m->adam_m.end(), 0.0f);
  std::fill(m->adam_v.begin(), m->adam_v.end(), 0.0f);
  m->step = 0;
  m->version = next_model_version();
  return PM_OK;
}

int pm_graph_features(pm_graph* g, float* out) {
  if (!g || (!out && g->h.n)) return PM_ERR_ARG;
  return graph_features_host(g, out);
}

int pm_model_predict(const pm_model* m, const float* X, uint64_t n, float* Y) {
  if (!m || (n && (!X || !Y))) return PM_ERR_ARG;
  return model_predict_host(m, X, n, Y);
}

int pm_model_grad(const pm_model* m, const float* X, const float* Y, const float* W, uint64_t n,
                  float* grad, double* loss) {
  if (!m || !grad || (n && (!X || !Y))) return PM_ERR_ARG;
  return model_grad(m, X, Y, W, n, grad, loss);
}

int pm_model_train(pm_model* m, const float* X, const float* Y, const float* W, uint64_t n,
                   const pm_train_opts* o, double* final_loss) {
  if (!m || (n && (!X || !Y))) return PM_ERR_ARG;
  return model_train(m, X, Y, W, n, o, final_loss);
}

int pm_autotune(pm_graph* g, pm_pattern* const* p, int32_t n, const pm_opts* base,
                const pm_tune_opts* t, pm_opts* best, pm_tune_report* report) {
  if (!g || !p || n <= 0 || !best || !report) return PM_ERR_ARG;
  pm_opts def;
  pm_opts_default(&def);
  if (!base) base = &def;
  for (int32_t i = 0; i < n; i++)
    if (!p[i]) return PM_ERR_ARG;
  return autotune(g, p, n, base, t, best, report);
}

}  // extern "C"
