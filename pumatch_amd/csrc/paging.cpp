// Out-of-core package layout and cache replacement policies (host side).
//
// The reference packs vertices >= idx-1 in id order into packages of at
// most MAX_NODE_NUM_PER_PACKAGE vertices / MAX_PACKAGE_SIZE neighbours
// (create_buffer_from_graph, src/buffer.cu:986-1060) and keeps a bucketed
// LRU over the device slots (lruAccess_bucket / find_lowest_priority_*,
// src/buffer.cu:581-796).  Here a package is a contiguous vertex range whose
// rows fit one fixed-size slot (2 MB by default); every vertex belongs to
// exactly one package (rows longer than a slot get a package of their own
// spanning several consecutive slots).  Replacement decisions are made on
// the host between matching rounds from per-slot use stamps the kernels
// record, so the device never takes a lock.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "paging.hpp"

namespace pm {

int build_packages(const HostGraph& g, uint32_t slot_words, PackageLayout* out) {
  if (slot_words == 0) return PM_ERR_ARG;
  out->slot_words = slot_words;
  out->first_vertex.clear();
  out->edge_base.clear();
  out->words.clear();
  out->nslots.clear();
  out->pkg_of.assign(g.n, 0);
  uint64_t v = 0;
  while (v < g.n) {
    const uint64_t start = v;
    const uint64_t base = g.rowptr[v];
    uint64_t words = 0;
    while (v < g.n) {
      const uint64_t d = g.rowptr[v + 1] - g.rowptr[v];
      if (v > start && words + d > slot_words) break;
      words += d;
      v++;
      if (words >= slot_words) break;
    }
    const uint32_t p = (uint32_t)out->first_vertex.size();
    for (uint64_t u = start; u < v; u++) out->pkg_of[u] = p;
    out->first_vertex.push_back((uint32_t)start);
    out->edge_base.push_back(base);
    out->words.push_back(words);
    out->nslots.push_back((uint32_t)std::max<uint64_t>(1, (words + slot_words - 1) / slot_words));
  }
  return PM_OK;
}

void CachePolicy::reset(uint32_t nslots, int kind) {
  kind_ = kind;
  stamp_.assign(nslots, 0);
  loaded_.assign(nslots, 0);
  owner_.assign(nslots, -1);
  clock_ = 1;
}

void CachePolicy::touched(const uint8_t* used, uint32_t nslots) {
  clock_++;
  for (uint32_t s = 0; s < nslots && s < stamp_.size(); s++)
    if (used[s]) stamp_[s] = clock_;
}

void CachePolicy::set_scores(const std::vector<float>& package_score) { score_ = package_score; }

// Choose `need` contiguous free-or-evictable runs of `width` slots for each
// request; returns the first slot of the run or -1.  Slots whose package is
// in `pinned` are never chosen.
int64_t CachePolicy::victim(uint32_t width, const std::vector<uint8_t>& pinned) {
  const uint32_t n = (uint32_t)stamp_.size();
  int64_t best = -1;
  double best_key = 0;
  for (uint32_t s = 0; s + width <= n; s++) {
    bool ok = true;
    double key = 0;
    for (uint32_t w = 0; w < width && ok; w++) {
      const uint32_t t = s + w;
      if (pinned[t]) ok = false;
      if (owner_[t] < 0) continue;  // free slot: best possible
      double k;
      if (kind_ == PM_POLICY_FIFO) k = (double)loaded_[t];
      else if (kind_ == PM_POLICY_LEARNED && !score_.empty())
        // low predicted reuse first, then least recently used
        k = (double)score_[owner_[t]] * 1e12 + (double)stamp_[t];
      else k = (double)stamp_[t];
      key = std::max(key, k + 1.0);
    }
    if (!ok) continue;
    if (best < 0 || key < best_key) {
      best = s;
      best_key = key;
      if (key == 0) break;
    }
  }
  return best;
}

void CachePolicy::assign(uint32_t first, uint32_t width, int32_t pkg) {
  for (uint32_t w = 0; w < width; w++) {
    owner_[first + w] = pkg;
    loaded_[first + w] = clock_;
    stamp_[first + w] = clock_;
  }
}

}  // namespace pm
