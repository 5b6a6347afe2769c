// .g pattern parsing and the matching plan.
//
// The matching order is the file order (identity, src/pattern.h:104-122).
// Automorphisms and the symmetry-breaking pairs follow the reference's
// stabiliser chain (src/pattern.h:198-262).  The set-operation plan is this
// library's own: for each depth d (vertices 0..d matched) and each future
// vertex u > d, u's running candidate set is the intersection of N(phi(j))
// over its earlier neighbours j <= d; identical masks share one slot, a slot
// whose mask did not change is reused, and the first neighbour's list is a
// zero-copy view of the CSR row.  The reference builds the same DAG of
// intersections with copy slots (src/pattern.h:143-196, convert2oned
// :329-355); the counts do not depend on how the intersections are shared.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "pm_internal.hpp"

namespace pm {

// src/pattern.h:67-100: skip leading '#' lines, then "v id label" lines in
// file order (the id field is ignored, labels are taken in line order), then
// "e a b label" lines until end of file; an edge sets both adj[a][b] and
// adj[b][a].  Blank or unparsable edge lines are rejected here (the
// reference would read uninitialised values).
int parse_pattern_text(const std::string& text, int labeled, Plan* out, std::string* err) {
  std::istringstream in(text);
  std::string line;
  std::vector<std::string> lines;
  while (std::getline(in, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t'))
      line.pop_back();
    lines.push_back(line);
  }
  size_t i = 0;
  while (i < lines.size() && !lines[i].empty() && lines[i][0] == '#') i++;
  Plan p;
  std::memset(p.adj, 0, sizeof(p.adj));
  int k = 0;
  while (i < lines.size() && !lines[i].empty() && lines[i][0] == 'v') {
    std::istringstream s(lines[i]);
    char tag;
    long long id, lab;
    if (!(s >> tag >> id >> lab)) {
      *err = "malformed vertex line: '" + lines[i] + "'";
      return PM_ERR_FORMAT;
    }
    if (k >= kMaxK) {
      *err = "pattern has more than 7 vertices (PAT_SIZE)";
      return PM_ERR_PATTERN;
    }
    if ((labeled & PM_LABELED) && (lab < 0 || lab > 31)) {
      *err = "vertex label outside 0..31";
      return PM_ERR_PATTERN;
    }
    p.vlabel[k] = (labeled & PM_LABELED) ? (int)lab : 1;
    k++;
    i++;
  }
  if (k == 0) {
    *err = "no vertex lines";
    return PM_ERR_FORMAT;
  }
  for (; i < lines.size(); i++) {
    if (lines[i].empty()) continue;
    if (lines[i][0] == '#') continue;
    std::istringstream s(lines[i]);
    char tag;
    long long a, b, lab;
    if (!(s >> tag >> a >> b >> lab) || tag != 'e') {
      *err = "malformed edge line: '" + lines[i] + "'";
      return PM_ERR_FORMAT;
    }
    if (a < 0 || b < 0 || a >= k || b >= k) {
      *err = "edge endpoint outside the vertex range";
      return PM_ERR_FORMAT;
    }
    p.adj[a][b] = (int)lab;
    p.adj[b][a] = (int)lab;
  }
  p.k = k;
  p.labeled = (labeled & PM_LABELED) ? 1 : 0;
  p.induced = (labeled & PM_VERTEX_INDUCED) ? 1 : 0;
  *out = p;
  return PM_OK;
}

namespace {

// A permutation pp is kept when mapping every vertex i to pp[i] carries the
// nonzero adjacency of i onto the label-1 adjacency of pp[i]
// (src/pattern.h:207-234).
bool keeps_adjacency(const Plan& p, const int* pp) {
  uint32_t image[kMaxK] = {};
  for (int i = 0; i < p.k; i++)
    for (int j = 0; j < p.k; j++)
      if (p.adj[i][j] != 0) image[pp[i]] |= 1u << pp[j];
  for (int i = 0; i < p.k; i++) {
    uint32_t ones = 0;
    for (int j = 0; j < p.k; j++)
      if (p.adj[i][j] == 1) ones |= 1u << j;
    if (ones != image[i]) return false;
  }
  return true;
}

}  // namespace

int build_plan(Plan* pp, std::string* err) {
  Plan& p = *pp;
  const int k = p.k;
  if (k < 3) {
    *err = "pattern needs at least 3 vertices";
    return PM_ERR_PATTERN;
  }
  for (int i = 0; i < k; i++) {
    int d = 0;
    for (int j = 0; j < k; j++)
      if (p.adj[i][j] > 0) d++;
    p.deg[i] = d;  // src/pattern.h:115-121
  }
  for (int u = 0; u < k; u++) {
    p.A[u] = p.X[u] = 0;
    for (int j = 0; j < u; j++) {
      if (p.adj[u][j] != 0) p.A[u] |= 1u << j;
      else p.X[u] |= 1u << j;
    }
  }
  if (!(p.A[1] & 1u)) {
    *err = "pattern vertices 0 and 1 must be adjacent (a job is an edge)";
    return PM_ERR_PATTERN;
  }
  for (int u = 2; u < k; u++)
    if (p.A[u] == 0) {
      *err = "every pattern vertex needs an earlier neighbour in file order";
      return PM_ERR_PATTERN;
    }

  // automorphisms: all k! permutations in lexicographic order
  std::vector<std::vector<int>> aut;
  std::vector<int> perm(k);
  for (int i = 0; i < k; i++) perm[i] = i;
  do {
    if (keeps_adjacency(p, perm.data())) aut.push_back(perm);
  } while (std::next_permutation(perm.begin(), perm.end()));
  p.aut = (long long)aut.size();

  // stabiliser chain (src/pattern.h:238-252): at step v, every kept
  // permutation that moves v contributes the pair (v, x[v]); those that fix
  // v form the next group.  The pair (j,u) means phi(u) < phi(j).
  for (int u = 0; u < k; u++) p.B[u] = 0;
  std::vector<std::vector<int>> group = aut;
  for (int v = 0; v < k; v++) {
    std::vector<std::vector<int>> fixed;
    for (auto& x : group) {
      if (x[v] == v) fixed.push_back(x);
      else p.B[x[v]] |= 1u << v;
    }
    group.swap(fixed);
  }
  if (p.labeled) {
    for (int u = 0; u < k; u++) p.B[u] = 0;  // ub = INT_MAX, src/gpu_match.cu:454-458
  }
  p.rc_order = (!p.labeled && (p.B[1] & 1u)) ? 1 : 0;

  // ---- folding of pendant trees (homomorphism semantics only; a
  // vertex-induced pendant vertex must also avoid the other vertices'
  // neighbourhoods, so nothing is folded there) ----
  bool alive[kMaxK];
  for (int u = 0; u < k; u++) alive[u] = true;
  p.nfold = 0;
  for (int u = 0; u < k; u++) p.fold_children[u] = 0;
  if (p.labeled && !p.induced) {
    bool changed = true;
    while (changed) {
      changed = false;
      for (int u = k - 1; u >= 2; u--) {
        if (!alive[u]) continue;
        int nb = 0, par = -1;
        for (int w = 0; w < k; w++)
          if (w != u && alive[w] && p.adj[u][w] != 0) {
            nb++;
            par = w;
          }
        // every earlier neighbour of a live vertex is live, so the single
        // live neighbour of a leaf is earlier than it
        if (nb == 1 && par < u) {
          alive[u] = false;
          p.fold_v[p.nfold] = u;
          p.fold_parent[p.nfold] = par;
          p.nfold++;
          p.fold_children[par] |= 1u << u;
          changed = true;
        }
      }
    }
  }
  p.kc = 0;
  for (int u = 0; u < k; u++)
    if (alive[u]) p.core_v[p.kc++] = u;
  return build_core(&p, err);
}

// Core arrays and slot operations for the core order in p.core_v.
int build_core(Plan* pp, std::string* err) {
  Plan& p = *pp;
  for (int i = 0; i < p.kc; i++) {
    const int u = p.core_v[i];
    p.clabel[i] = p.vlabel[u];
    p.cA[i] = p.cX[i] = p.cB[i] = 0;
    for (int j = 0; j < i; j++) {
      const int w = p.core_v[j];
      if (p.adj[u][w] != 0) p.cA[i] |= 1u << j;
      else p.cX[i] |= 1u << j;
      if (p.B[u] >> w & 1u) p.cB[i] |= 1u << j;
    }
  }

  // ---- slots over the core ----
  const int kc = p.kc;
  for (int d = 0; d < kMaxK; d++) {
    p.ops[d].clear();
    for (int u = 0; u < kMaxK; u++) p.slot_of[d][u] = -1;
  }
  p.nslots = 0;
  p.n_materialised = 0;
  for (int d = 0; d <= kc - 2; d++) {
    const uint32_t upto = (2u << d) - 1u;
    for (int u = std::max(d + 1, 2); u < kc; u++) {
      const uint32_t M = p.cA[u] & upto;
      if (M == 0) continue;
      if (!(M >> d & 1u)) {
        p.slot_of[d][u] = p.slot_of[d - 1][u];
        continue;
      }
      int found = -1;
      for (int w = d + 1; w < u; w++)
        if ((p.cA[w] & upto) == M && p.slot_of[d][w] >= 0) found = p.slot_of[d][w];
      if (found >= 0) {
        p.slot_of[d][u] = found;
        continue;
      }
      if (p.nslots >= kMaxSlots) {
        *err = "too many slots";
        return PM_ERR_PATTERN;
      }
      SlotOp op;
      std::memset(&op, 0, sizeof(op));
      op.slot = (int8_t)p.nslots++;
      if (M == (1u << d)) {
        op.kind = 0;
        op.parent = -1;
      } else {
        op.kind = 1;
        op.parent = (int8_t)p.slot_of[d - 1][u];
        if (d < kc - 2) p.n_materialised++;
      }
      op.labmask = 0;
      op.ncons = 0;
      for (int w = std::max(d + 1, 2); w < kc; w++) {
        if ((p.cA[w] & upto) != M) continue;
        op.labmask |= 1u << p.clabel[w];
        op.cons_b[op.ncons++] = (uint8_t)(p.cB[w] & upto);
      }
      p.slot_of[d][u] = op.slot;
      p.ops[d].push_back(op);
    }
  }
  return PM_OK;
}

// Labeled plans: the order of core vertices 2..kc-1.  A labeled count is
// the number of homomorphisms with phi(0) = r, phi(1) = c summed over the
// jobs, so it does not depend on the order in which vertices 2.. are
// matched (the job filter keeps the file-order degrees, see deg[]); only the
// work does.  Candidate orders are those in which every vertex has an
// earlier neighbour.  The cost model estimates the partial matches at the
// enumerated depths 2..kc-2 (the last vertex is counted, not enumerated): a
// vertex with b earlier neighbours and label l has about
// dnb * f(l) * q^(b-1) candidates, dnb the mean neighbour degree, f(l) the
// label's share and q the chance that a candidate also lies in one more
// neighbour list.  Returns up to `max` orders (core vertex ids at positions
// 2..kc-1), cheapest first; the file order is always included.
std::vector<std::vector<int>> rank_core_orders(const Plan& p, const OrderStats& st, int max) {
  std::vector<std::vector<int>> out;
  const int kc = p.kc;
  std::vector<int> file(p.core_v + 2, p.core_v + std::max(kc, 2));
  if (!p.labeled || kc < 4) {
    out.push_back(file);
    return out;
  }
  const double n = std::max(st.n, 1.0);
  const double dnb = std::max(st.dnb, 1.0);
  const double q = std::min(0.5, std::max(0.02, 4.0 * dnb / n));
  auto share = [&](int l) {
    return (l >= 0 && l < (int)st.label_share.size()) ? std::max(st.label_share[l], 1e-6) : 1e-6;
  };
  auto cost_of = [&](const std::vector<int>& ord, double* out_cost) {
    std::vector<int> seq = {p.core_v[0], p.core_v[1]};
    seq.insert(seq.end(), ord.begin(), ord.end());
    double prefixes = 1.0, cost = 0.0;
    for (int i = 2; i < kc; i++) {
      const int u = seq[i];
      int b = 0;
      for (int j = 0; j < i; j++) b += p.adj[u][seq[j]] != 0;
      if (b == 0) return false;
      if (i <= kc - 2) {
        prefixes *= dnb * share(p.vlabel[u]) * std::pow(q, b - 1);
        cost += prefixes;
      }
    }
    *out_cost = cost;
    return true;
  };
  std::vector<std::pair<double, std::vector<int>>> all;
  double file_cost = 0;
  if (!cost_of(file, &file_cost)) file_cost = 1e300;
  std::vector<int> rest = file;
  std::sort(rest.begin(), rest.end());
  do {
    double c;
    if (rest != file && cost_of(rest, &c)) all.push_back({c, rest});
  } while (std::next_permutation(rest.begin(), rest.end()));
  // ties keep file order first
  all.push_back({file_cost * 0.999, file});
  std::stable_sort(all.begin(), all.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  max = std::max(max, 1);
  bool have_file = false;
  for (size_t i = 0; i < all.size() && (int)out.size() < max; i++) {
    out.push_back(all[i].second);
    have_file |= all[i].second == file;
  }
  if (!have_file) out.back() = file;
  return out;
}

// Rebuild p's core for the order `rest` (core vertex ids at positions
// 2..kc-1); false (p unchanged) when that order needs too many slots.
bool apply_core_order(Plan* pp, const std::vector<int>& rest) {
  Plan& p = *pp;
  if ((int)rest.size() != std::max(p.kc - 2, 0)) return false;
  const Plan before = p;
  for (int i = 2; i < p.kc; i++) p.core_v[i] = rest[i - 2];
  std::string err;
  if (build_core(&p, &err) != PM_OK) {
    p = before;
    return false;
  }
  return true;
}

void reorder_core(Plan* p, const OrderStats& st) {
  if (!p->labeled || p->kc < 4) return;
  for (const auto& o : rank_core_orders(*p, st, 2))
    if (apply_core_order(p, o)) return;
}

OrderStats order_stats(const HostGraph& h) {
  OrderStats st;
  st.n = (double)h.n;
  st.m = (double)h.m;
  double s2 = 0;
  for (uint64_t v = 0; v < h.n; v++) {
    const double d = (double)(h.rowptr[v + 1] - h.rowptr[v]);
    s2 += d * d;
  }
  st.dnb = h.m ? s2 / (double)h.m : 0.0;
  if (!h.labels.empty()) {
    uint32_t mx = 0;
    for (uint32_t l : h.labels) mx = std::max(mx, l);
    st.label_share.assign((size_t)std::min<uint32_t>(mx, 1u << 20) + 1, 0.0);
    for (uint32_t l : h.labels)
      if (l < st.label_share.size()) st.label_share[l] += 1.0;
    for (double& x : st.label_share) x /= std::max(st.n, 1.0);
  }
  return st;
}

}  // namespace pm
