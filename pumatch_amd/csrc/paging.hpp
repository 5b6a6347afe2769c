// Out-of-core package layout and cache replacement policies (host side).
#pragma once
#include <cstdint>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

// Package p holds the rows of vertices [first_vertex[p], first_vertex[p+1])
// = edges [edge_base[p], edge_base[p] + words[p]); it occupies nslots[p]
// consecutive cache slots of slot_words u32 each.
struct PackageLayout {
  uint32_t slot_words = 0;
  std::vector<uint32_t> first_vertex;
  std::vector<uint64_t> edge_base;
  std::vector<uint64_t> words;
  std::vector<uint32_t> nslots;
  std::vector<uint32_t> pkg_of;  // vertex -> package
  uint32_t count() const { return (uint32_t)first_vertex.size(); }
};

int build_packages(const HostGraph& g, uint32_t slot_words, PackageLayout* out);

// Victim selection over the slots of the HBM package cache.
class CachePolicy {
 public:
  void reset(uint32_t nslots, int kind);
  // slots used by the last round (one byte per slot, nonzero = used)
  void touched(const uint8_t* used, uint32_t nslots);
  // learned policy: predicted reuse score per package (higher = keep)
  void set_scores(const std::vector<float>& package_score);
  // first slot of the best run of `width` slots avoiding pinned slots, or -1
  int64_t victim(uint32_t width, const std::vector<uint8_t>& pinned);
  void assign(uint32_t first, uint32_t width, int32_t pkg);
  int32_t owner(uint32_t slot) const { return owner_[slot]; }
  void release(uint32_t slot) { owner_[slot] = -1; }
  int kind() const { return kind_; }

 private:
  int kind_ = 0;
  uint64_t clock_ = 1;
  std::vector<uint64_t> stamp_;   // last round the slot was used
  std::vector<uint64_t> loaded_;  // round the slot was filled
  std::vector<int32_t> owner_;    // package in the slot (-1 free)
  std::vector<float> score_;
};

}  // namespace pm
