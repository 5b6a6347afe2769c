// Autotuner: measured choice of core orders and launch granularities.
//
// The reference fixes its granularities at compile time (JOB_CHUNK_SIZE,
// GRID_DIM/BLOCK_DIM, src/config.h:9-25) and matches in file order.  Here
// every such choice leaves the count unchanged, so it can be picked per
// (graph, pattern set) by timing: each trial is a launch over a fixed sample
// of the jobs (every s-th block of 4096 CSR edges), its fastest of a few
// repeats is compared, and its per-pattern counts must equal the first
// trial's (a built-in consistency check: the sampled job set is the same in
// every trial).
#include <algorithm>
#include <cstring>
#include <vector>

#include "pm_internal.hpp"

namespace pm {
namespace {

struct Trial {
  double ms = 0;
  std::vector<uint64_t> counts;
};

int run_trial(pm_graph* g, const std::vector<const Plan*>& plans, const pm_opts& o, Trial* t) {
  std::vector<pm_result> res(plans.size());
  const int rc = device_match_batch(g, plans.data(), (int)plans.size(), &o, res.data());
  if (rc) return rc;
  t->ms = res.empty() ? 0.0 : res[0].min_ms;
  t->counts.resize(res.size());
  for (size_t i = 0; i < res.size(); i++) t->counts[i] = res[i].count;
  return PM_OK;
}

}  // namespace

int autotune(pm_graph* g, pm_pattern* const* p, int n, const pm_opts* base,
             const pm_tune_opts* topt, pm_opts* best, pm_tune_report* rep) {
  pm_tune_opts t;
  std::memset(&t, 0, sizeof(t));
  if (topt) t = *topt;
  const HostGraph& h = g->h;
  const uint64_t blocks = (h.m + 4095) / 4096;
  // about 256 sampled blocks (1M edges) unless the caller says otherwise
  const int sample = t.sample > 0 ? t.sample : (int)std::max<uint64_t>(1, blocks / 256);
  const int max_orders = t.max_orders > 0 ? t.max_orders : 4;
  const int repeats = t.repeats > 0 ? t.repeats : 2;
  pm_opts o = *base;
  o.per_edge = nullptr;
  o.sample = sample;
  o.repeats = repeats;
  o.order = 0;
  std::memset(rep, 0, sizeof(*rep));
  rep->sample = sample;
  if (!g->stats_ok) {
    g->stats = order_stats(h);
    g->stats_ok = true;
  }

  // (1) core orders, one labeled pattern at a time
  for (int i = 0; i < n; i++) {
    Plan& pl = p[i]->plan;
    pl.tuned_graph = nullptr;
    if (!pl.labeled || pl.kc < 4) continue;
    const auto cands = rank_core_orders(pl, g->stats, max_orders);
    const std::vector<int> file(pl.core_v + 2, pl.core_v + pl.kc);
    double best_ms = 0;
    uint64_t want = 0;
    int best_c = -1;
    for (size_t c = 0; c < cands.size(); c++) {
      Plan trial = pl;
      if (!apply_core_order(&trial, cands[c])) continue;
      trial.tuned_graph = (const void*)g;
      for (int j = 0; j < trial.kc; j++) trial.tuned_core[j] = trial.core_v[j];
      Trial r;
      const int rc = run_trial(g, {&trial}, o, &r);
      if (rc) return rc;
      rep->trials += repeats;
      if (best_c < 0) want = r.counts[0];
      if (r.counts[0] != want) {
        set_error("autotune: core orders disagree on a count");
        return PM_ERR_DEVICE;
      }
      if (best_c < 0 || r.ms < best_ms) {
        best_ms = r.ms;
        best_c = (int)c;
      }
    }
    if (best_c >= 0) {
      pl.tuned_graph = (const void*)g;
      pl.tuned_core[0] = pl.core_v[0];
      pl.tuned_core[1] = pl.core_v[1];
      for (size_t j = 0; j < cands[best_c].size(); j++) pl.tuned_core[j + 2] = cands[best_c][j];
      if (cands[best_c] != file) rep->reordered++;
    }
  }

  // (2) launch granularities of the whole batch, coordinate descent from
  // the caller's values
  std::vector<const Plan*> plans(n);
  for (int i = 0; i < n; i++) plans[i] = &p[i]->plan;
  Trial first;
  int rc = run_trial(g, plans, o, &first);
  if (rc) return rc;
  rep->trials += repeats;
  {
    pm_opts b0 = *base;
    b0.per_edge = nullptr;
    b0.sample = sample;
    b0.repeats = repeats;
    Trial r0;
    // the caller's options with file/cost order as they would run untuned
    std::vector<Plan> untuned(n);
    std::vector<const Plan*> up(n);
    for (int i = 0; i < n; i++) {
      untuned[i] = p[i]->plan;
      untuned[i].tuned_graph = nullptr;
      up[i] = &untuned[i];
    }
    rc = run_trial(g, up, b0, &r0);
    if (rc) return rc;
    rep->trials += repeats;
    if (r0.counts != first.counts) {
      set_error("autotune: tuned and untuned runs disagree on a count");
      return PM_ERR_DEVICE;
    }
    rep->base_ms = r0.ms;
  }
  double cur = first.ms;
  const int chunks[] = {0, 4, 16, 64};
  const int splits[] = {0, 2048, 32768, -1};
  const int hubs[] = {0, 2, 4, 32};
  for (int pass = 0; pass < 2; pass++) {
    for (int knob = 0; knob < 3; knob++) {
      const int* vals = knob == 0 ? chunks : knob == 1 ? splits : hubs;
      const int nv = 4;
      for (int v = 0; v < nv; v++) {
        pm_opts c = o;
        int32_t* field = knob == 0 ? &c.chunk : knob == 1 ? &c.split : &c.hub_factor;
        const int32_t now = knob == 0 ? o.chunk : knob == 1 ? o.split : o.hub_factor;
        if (vals[v] == now) continue;
        *field = vals[v];
        Trial r;
        rc = run_trial(g, plans, c, &r);
        if (rc) return rc;
        rep->trials += repeats;
        if (r.counts != first.counts) {
          set_error("autotune: launch options disagree on a count");
          return PM_ERR_DEVICE;
        }
        if (r.ms < cur * 0.97) {
          cur = r.ms;
          o = c;
        }
      }
    }
  }
  rep->best_ms = cur;
  *best = *base;
  best->chunk = o.chunk;
  best->split = o.split;
  best->hub_factor = o.hub_factor;
  best->order = 0;
  return PM_OK;
}

}  // namespace pm
