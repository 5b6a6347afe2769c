// Internal types of the MI355X subgraph matcher (host side).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pumatch.h"

namespace pm {

constexpr int kMaxK = PM_MAX_PATTERN;  // PAT_SIZE, src/config.h:9
constexpr int kMaxSlots = 32;

// Error reporting: the last error message of the calling thread.
void set_error(const std::string& msg);

// Host CSR in the reference's on-disk layout (src/graph.h:228-254):
// rowptr u64[n+1], colidx u32[m], raw labels u32[n].
struct HostGraph {
  uint64_t n = 0, m = 0;
  std::vector<uint64_t> rowptr;
  std::vector<uint32_t> colidx;
  std::vector<uint32_t> labels;  // empty when the graph carries no labels
  uint32_t max_degree = 0;
};

// Device-resident copy (HBM).  label8 holds the raw label (0..31) as a byte,
// maxnbr[v] = max degree over N(v) (evaluates the job filter's
// "exists w in N(c) with deg(w) >= d_P(2)" in O(1)).
// lcol/loff: the same rows regrouped by neighbour label (ids ascending
// within a label); the label-l part of row v is
// lcol[rowptr[v] + loff[v*nl + l] .. rowptr[v] + loff[v*nl + l + 1]).
struct DeviceGraph {
  uint64_t* rowptr = nullptr;
  uint32_t* colidx = nullptr;
  uint8_t* label8 = nullptr;
  uint32_t* maxnbr = nullptr;
  uint32_t* lcol = nullptr;
  uint32_t* loff = nullptr;
  uint32_t nl = 0;  // labels + 1 (0: no label-grouped copy)
  uint32_t* src64 = nullptr;  // source vertex of edge 64*i
  size_t bytes = 0;
};

}  // namespace pm

namespace pm {
// Graph figures the planner's cost model reads.
struct OrderStats {
  double n = 0, m = 0;
  double dnb = 0;                   // mean degree of a neighbour: sum d^2 / sum d
  std::vector<double> label_share;  // fraction of vertices per label
};
}  // namespace pm

struct pm_model {
  std::vector<float> w, adam_m, adam_v;  // PM_MODEL_PARAMS each
  int64_t step = 0;
  uint64_t version = 0;  // unique per parameter state (process-wide counter)
};

struct pm_graph {
  pm::HostGraph h;
  pm::DeviceGraph d;
  bool on_device = false;
  int device = 0;
  bool stats_ok = false;  // cleared when labels change
  pm::OrderStats stats;
  // model predictions per vertex (n x 2), for pred_model at pred_version
  const pm_model* pred_model = nullptr;
  uint64_t pred_version = 0;
  std::vector<float> pred;
  // learned job order: roots by predicted work, and the first job position
  // of each (device), for order_version
  uint32_t* d_root_order = nullptr;
  uint64_t* d_root_start = nullptr;
  uint64_t order_version = 0;
};

namespace pm {

// One set operation executed when vertex `depth` has just been matched.
//   kind VIEW:      slot = N(phi(depth))               (zero-copy CSR view)
//   kind INTERSECT: slot = slot[parent] n N(phi(depth)) (materialised)
struct SlotOp {
  int8_t kind;      // 0 = VIEW, 1 = INTERSECT
  int8_t slot;      // destination slot id
  int8_t parent;    // parent slot id (INTERSECT)
  int8_t pad;
  uint32_t labmask;  // labeled: OR of 1<<label over consumers (filter on materialise)
  // symmetry-breaking truncation: slot keeps x < max over consumers c of
  // min_{j in cons_b[c]} phi(j); a consumer with an empty mask disables it.
  uint8_t ncons;
  uint8_t cons_b[kMaxK];
};

struct Plan {
  // ---- the pattern as read (file order = matching order) ----
  int k = 0;
  int labeled = 1;
  int induced = 0;               // vertex-induced (EDGE_INDUCED=false)
  long long aut = 1;             // PatternMultiplicity, src/pattern.h:236
  int rc_order = 0;              // unlabeled job filter r > c
  int deg[kMaxK] = {};           // pattern degree (job filter uses deg[0..2])
  int vlabel[kMaxK] = {};        // pattern vertex label
  uint32_t A[kMaxK] = {};        // earlier neighbours
  uint32_t B[kMaxK] = {};        // symmetry pairs phi(u) < phi(j)
  uint32_t X[kMaxK] = {};        // earlier non-neighbours
  int adj[kMaxK][kMaxK] = {};    // edge labels as read

  // ---- pendant-tree folding (labeled mode only) ----
  // A vertex u >= 2 whose only unfolded neighbour is p < u is folded: its
  // subtree contributes the factor F_u(phi(p)), with
  //   F_u(v) = sum_{x in N(v), lab(x) = lab(u)} prod_{c child of u} F_c(x),
  // exact for homomorphism counting.  fold_v lists folded vertices children
  // first.
  int nfold = 0;
  int fold_v[kMaxK] = {};
  int fold_parent[kMaxK] = {};
  uint32_t fold_children[kMaxK] = {};  // by original id

  // ---- the core that the DFS enumerates (core index -> original id) ----
  int kc = 0;
  int core_v[kMaxK] = {};
  int clabel[kMaxK] = {};
  uint32_t cA[kMaxK] = {}, cB[kMaxK] = {}, cX[kMaxK] = {};

  // ---- execution plan over the core ----
  int nslots = 0;
  int slot_of[kMaxK][kMaxK];     // [depth][u] slot holding u's running set
  std::vector<SlotOp> ops[kMaxK];  // ops run at depth d (d = 0..kc-2)
  int n_materialised = 0;        // INTERSECT slots (arena sizing)

  // core order measured fastest by pm_autotune on graph `tuned_graph`
  // (positions 2..kc-1); used there instead of the cost model
  const void* tuned_graph = nullptr;
  int tuned_core[kMaxK] = {};
};

}  // namespace pm

struct pm_pattern {
  pm::Plan plan;
  std::string path;
};

namespace pm {

int parse_pattern_text(const std::string& text, int labeled, Plan* out, std::string* err);
int build_plan(Plan* p, std::string* err);
int build_core(Plan* p, std::string* err);
// labeled plans: cost-based order of core vertices 2..kc-1 (same count)
void reorder_core(Plan* p, const OrderStats& st);
std::vector<std::vector<int>> rank_core_orders(const Plan& p, const OrderStats& st, int max);
bool apply_core_order(Plan* p, const std::vector<int>& rest);
OrderStats order_stats(const HostGraph& h);

// generators (gen.cpp)
int generate_rmat(HostGraph* g, int scale, double edge_factor, double a, double b, double c,
                  uint64_t seed);
int generate_chunglu(HostGraph* g, uint64_t n, double gamma, double dmax, double avg,
                     uint64_t seed);
void assign_uniform_labels(HostGraph* g, uint32_t nlabels, uint64_t seed);
int build_csr_from_pairs(HostGraph* g, uint64_t n, std::vector<uint64_t>& pairs);
void finalize_graph(HostGraph* g);
void compute_max_neighbour_degree(const HostGraph& g, std::vector<uint32_t>* out);
// rows regrouped by neighbour label (stable, so ids stay ascending within a
// label); nl = max label + 2
void build_label_grouped(const HostGraph& g, std::vector<uint32_t>* lcol,
                         std::vector<uint32_t>* loff, uint32_t* nl);

// device (match.hip)
int device_upload(pm_graph* g);
int device_upload_rows(pm_graph* g, bool with_rows);
void device_release(pm_graph* g);
int device_match(pm_graph* g, const Plan& plan, const pm_opts* opts, pm_result* res);
// One launch over several patterns; res[i] gets pattern i's count and jobs
// and the launch-wide timings.
int device_match_batch(pm_graph* g, const Plan* const* plans, int npat, const pm_opts* opts,
                       pm_result* res);
int autotune(pm_graph* g, pm_pattern* const* p, int n, const pm_opts* base,
             const pm_tune_opts* t, pm_opts* best, pm_tune_report* rep);
// learned parts (learn.hip)
uint64_t next_model_version();
void model_init(pm_model* m, uint64_t seed);
int model_predict_host(const pm_model* m, const float* X, uint64_t n, float* Y);
int model_grad(const pm_model* m, const float* X, const float* Y, const float* Wt, uint64_t n,
               float* grad, double* loss);
int model_train(pm_model* m, const float* X, const float* Y, const float* Wt, uint64_t n,
                const pm_train_opts* o, double* final_loss);
int graph_features_host(pm_graph* g, float* out);
int graph_predictions(pm_graph* g, const pm_model* m, const std::vector<float>** out);

}  // namespace pm
