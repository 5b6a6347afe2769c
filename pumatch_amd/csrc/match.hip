// Subgraph-matching kernels for gfx950 (CDNA4, wave64) and their host driver.
//
// One persistent wavefront runs one depth-first search at a time (the
// reference: one warp per DFS, _parallel_match / match / extend,
// src/gpu_match.cu:855-1178).
//
// Work distribution (replaces the spin-locked job array of
// src/gpu_match.cu:321-384 and the host job scan src/job_queue.h:39-89):
//   phase 1  persistent waves claim chunks of CSR edges with one global
//            atomic; the job filter of src/job_queue.h:44-61 is evaluated
//            lane-parallel; a job whose vertex-2 candidates carry more than
//            `split` list entries of estimated work keeps its first range
//            and writes the rest as range tasks;
//   phase 2  persistent waves claim range tasks, rebuild the depth-0/1 sets
//            of the task's job and finish the range.  This is the heavy-root
//            splitting the reference approximates with stealing (trans_layer,
//            src/gpu_match.cu:21-109, disabled by default).
//
// Set operations (src/gpu_match.cu:238-319): sorted-list intersections,
// each lane taking one element of the shorter list and searching the longer
// one inside a monotonically advancing window; hits are compacted in order
// with a 64-bit ballot + popcount.  Intermediate sets live in a per-wave HBM
// arena sized so no set is ever truncated (the reference caps sets at 4096,
// src/gpu_match.cu:302,316).
//
// The last two core vertices are fused (the reference enumerates vertex k-2
// one candidate at a time and counts vertex k-1, :1018-1041): candidates of
// vertex k-2 are taken 64 at a time and the matches of vertex k-1 for each
// are counted lane-parallel; when that count is an intersection with a
// shared parent set P, the candidates' neighbour lists are walked
// edge-parallel and membership in P is an LDS hash probe.
//
// Labeled (homomorphism) counting folds pendant trees of the pattern into
// per-vertex weight tables built by an edge-parallel sparse pass (k_fold);
// the DFS runs over the remaining core and multiplies weights.
//
// Out of core (the package buffer of src/buffer.cu and the missing-branch
// tree of src/missing_tree.cu): rows stay in pinned host memory, cut into
// packages that are paged into a fixed HBM cache between matching rounds.
// A wave that needs a row whose package is not resident records the package
// (deduplicated bitmap + list), writes a resume descriptor for exactly the
// point it reached (job, candidate index per depth) and returns with only
// completed subtrees counted.  The host loads the missed packages
// (hipMemcpyAsync on a side stream into policy-chosen victims) and the next
// round resumes every descriptor.  Nothing is dropped, so counts are exact
// (the reference replays abandoned branches at most REPEATED_CNT times from
// a 1024-node tree, src/gpu_match.cu:926,969, src/missing_tree.cu:46).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "paging.hpp"
#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr int kWave = 64;
constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr int kHashBits = 9;
constexpr int kHashCap = 1 << kHashBits;  // LDS hash of the fused parent set
// build-time tunables: fused-walk entries in flight per lane, and the
// minimum waves per SIMD the resident kernels are compiled for (the LDS
// state allows 5-6)
#ifndef PM_FUSED_WALK
#define PM_FUSED_WALK 4
#endif
#ifndef PM_RESIDENT_WAVES
#define PM_RESIDENT_WAVES 5
#endif
// sorted-set intersection: two 64-element batches per step (1) or one (0)
#ifndef PM_ISECT_DUAL
#define PM_ISECT_DUAL 0
#endif
constexpr int kWalk = PM_FUSED_WALK;
constexpr uint32_t kResumeClaim = 16;  // deferred pieces claimed per atomic
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
// heavy-root split cost of a vertex-2 candidate (build-time switch):
// 0: 1 + its list length; 1: 1 + length^2 / 64 for cores of 5+ vertices;
// 2: 1 + length^(k-3) / 64^(k-4) for cores of k >= 4 vertices
#ifndef PM_SPLIT_QUAD
#define PM_SPLIT_QUAD 2
#endif
// range-task buffer: 2^PM_TASK_CAP_LOG2 tasks of 32 B (build-time tunable)
#ifndef PM_TASK_CAP_LOG2
#define PM_TASK_CAP_LOG2 22
#endif
constexpr int kMaxDepth = 6;

struct KOp {
  int8_t kind;  // 0 VIEW, 1 INTERSECT
  int8_t slot;
  int8_t parent;
  int8_t ncons;
  uint32_t labmask;
  uint8_t cons_b[7];
  int8_t seg;  // labeled, one label: read only that label's part of the row
};

struct KPlan {
  int32_t k;  // core size
  int32_t labeled, rc_order, nops;
  uint32_t deg0, deg1, deg2, lab0, lab1;  // job filter (whole pattern)
  uint32_t lab[8];                          // core vertex labels
  uint32_t B[8], X[8];                      // core symmetry / non-neighbour masks
  int8_t op_begin[8], op_end[8];
  int8_t cand_slot[8];   // [d]: slot holding candidates of core vertex d+1
  int8_t cand_check[8];  // [d]: labeled, candidates still need the label test
  int8_t region[kMaxSlots];
  int8_t last_kind;   // 0 reuse slot last_src, 1 view N(phi(k-2)), 2 intersect
  int8_t last_src;
  int8_t last_check;  // labeled: label test needed on the last vertex
  int8_t pair;        // fused tail over unordered adjacent pairs (see fused_tail)
  int8_t segs;        // label-grouped rows available (labeled, resident)
  int8_t induced;     // vertex-induced (resident only; no fused tail)
  int8_t pad[2];
  uint32_t split;       // phase-1 split threshold (list entries of work)
  uint32_t hub_factor;  // fused: owners with deg > hub_factor*|P| go one by one
  KOp ops[kMaxSlots];
  const unsigned long long* wt[8];  // weight table per core vertex (nullptr = 1)
};

struct KGraph {
  uint64_t n, m;
  const uint64_t* rowptr;
  const uint32_t* colidx;  // rows in HBM (nullptr when paged)
  const uint32_t* jcol;    // rows for the edge scan (HBM, or mapped host memory)
  const uint8_t* lab;
  const uint32_t* maxnbr;
  const uint32_t* lcol;  // rows grouped by neighbour label
  const uint32_t* loff;
  uint32_t nl;
  const uint32_t* src64;  // source vertex of every 64th edge
  // out of core (pkg_slot == nullptr: rows resident)
  const int32_t* pkg_slot;
  const uint32_t* pkg_of;
  const uint64_t* pkg_base;
  const uint32_t* cache;
  uint64_t slot_words;
  unsigned int* miss_bits;
  uint32_t* miss_list;
  unsigned long long* miss_count;
  uint32_t miss_cap;
  uint8_t* used;  // per cache slot: read during this round
  uint32_t direct_max;  // a missing row of <= this many entries is read from jcol
};

struct Task {
  unsigned long long edge;
  uint32_t lo, hi;
  uint32_t r, c;  // the job's endpoints
  uint32_t pat;   // pattern of a batched launch
  uint32_t pad;
};

// A deferred piece of search.  depth 0: the whole job (edge, r, c) over its
// vertex-2 range [lo, hi).  depth D >= 1: the subtrees of candidates
// [idx[D], end) at depth D below the path given by the candidate indices
// idx[1..D-1]; phi(0..D) are that path's vertices (their rows are needed to
// rebuild it) and blk is the package whose absence deferred the piece.
struct Resume {
  unsigned long long edge;
  uint32_t r, c, lo, hi;
  int32_t depth;
  uint32_t end;
  uint32_t idx[kMaxDepth];
  uint32_t phi[kMaxDepth];
  uint32_t blk;
  uint32_t pat;  // pattern of a batched launch
};

struct DeferOut {
  Resume* buf;
  unsigned long long* count;
  uint64_t cap;
  uint64_t soft;  // beyond this many pieces, searches stop at their first miss
  uint64_t budget;  // resident: wall-clock ticks after which a search re-splits (0 = never)
};

struct WaveState {
  const uint32_t* sptr[kMaxSlots];
  uint32_t slen[kMaxSlots];
  uint32_t phi[8];
  unsigned long long wacc[8];
  uint32_t cnext[8];
  uint32_t clen[8];
  uint32_t cbase[8];
  uint32_t cidx[8];  // index of the candidate being expanded at each depth
  uint64_t cmask[8];
  uint32_t cval[kMaxDepth][kWave];  // depths 1..k-3 <= 4
  uint32_t jr[kWave];
  uint32_t jc[kWave];
  unsigned long long jedge[kWave];  // CSR edge of each lane's job
  // resumed group of one-candidate pieces on one search path: their
  // candidate indices at the fused level and the vertices there
  uint32_t lcount;                  // 0: not a group
  uint32_t lidx[kWave];
  uint32_t ltmp[kWave];
  unsigned long long alg0;          // work counter at the start of the job (trace)
  // fused step: per-owner data of the current batch
  uint32_t own_off[kWave + 1];
  uint32_t own_ub[kWave];
  unsigned long long own_w[kWave];
  unsigned long long own_b[kWave];
  uint64_t own_ptr[kWave];  // row address of each owner (as an integer)
  uint32_t hash[kHashCap];
  unsigned long long je;  // the job being searched: edge and vertex-2 range
  uint32_t jlo, jhi;
  uint32_t pat;           // its pattern (batched launches)
  unsigned long long acc_count, acc_jobs;  // phase 1: totals of pattern `pat`
  unsigned long long t0;  // wall clock at the start of the current search
  int32_t rs_depth, rs_stop;  // pending re-split (depth, first open depth); -1: none
  int32_t rs_from;            // its first candidate at rs_depth (-1: after cidx)
  uint32_t blk;     // package of the last row found missing
  int32_t rdepth;   // resume input: depth, candidate indices, range end
  uint32_t ridx[kMaxDepth];
  uint32_t rend;
};

struct Counters {
  uint64_t alg = 0, nops = 0;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
template <typename T>
__device__ __forceinline__ T* uniptr(T* p) {
  return (T*)uni64((uint64_t)p);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// number of elements of a[0..n) strictly below x (src/gpu_match.cu:185-202)
__device__ __forceinline__ uint32_t below(const uint32_t* a, uint32_t n, uint32_t x) {
  uint32_t lo = 0;
  while (n > 0) {
    uint32_t half = n >> 1;
    if (a[lo + half] < x) {
      lo += half + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

__device__ __forceinline__ bool member(const uint32_t* a, uint32_t n, uint32_t x) {
  uint32_t p = below(a, n, x);
  return p < n && a[p] == x;
}

__device__ __forceinline__ uint32_t degree(const KGraph& g, uint32_t v) {
  return (uint32_t)(g.rowptr[v + 1] - g.rowptr[v]);
}

// Length of N(v) (or its label-`seg` part) without touching the row itself.
__device__ __forceinline__ uint32_t row_len(const KGraph& g, uint32_t v, int seg) {
  if (seg < 0) return degree(g, v);
  if ((uint32_t)seg + 1 >= g.nl) return 0;
  const uint32_t* o = g.loff + (size_t)v * g.nl + seg;
  return o[1] - o[0];
}

// N(v), or only its neighbours of label `seg` when seg >= 0 (ascending ids
// either way).  Out of core, a row whose package is not resident is recorded
// as a miss and false is returned.  Lane-local.
template <bool PG>
__device__ __forceinline__ bool rows(const KGraph& g, uint32_t v, int seg, const uint32_t** p,
                                     uint32_t* len, uint32_t* missing = nullptr) {
  const uint64_t r0 = g.rowptr[v];
  if (!PG) {
    if (seg < 0) {
      *p = g.colidx + r0;
      *len = (uint32_t)(g.rowptr[v + 1] - r0);
      return true;
    }
    if ((uint32_t)seg + 1 >= g.nl) {
      *p = g.lcol;
      *len = 0;
      return true;
    }
    const uint32_t* o = g.loff + (size_t)v * g.nl + seg;
    const uint32_t b = o[0];
    *p = g.lcol + r0 + b;
    *len = o[1] - b;
    return true;
  }
  const uint32_t pk = g.pkg_of[v];
  const int32_t s = g.pkg_slot[pk];
  if (s < 0) {
    // short rows come straight from pinned host memory (the rows of the
    // edge scan) instead of deferring the search
    const uint32_t n = (uint32_t)(g.rowptr[v + 1] - r0);
    if (n <= g.direct_max) {
      *p = g.jcol + r0;
      *len = n;
      return true;
    }
    if (missing) *missing = pk;
    const unsigned int bit = 1u << (pk & 31);
    if (!(atomicOr(&g.miss_bits[pk >> 5], bit) & bit)) {
      const unsigned long long i = atomicAdd(g.miss_count, 1ull);
      if (i < g.miss_cap) g.miss_list[i] = pk;
    }
    return false;
  }
  g.used[s] = 1;
  *p = g.cache + (uint64_t)s * g.slot_words + (r0 - g.pkg_base[pk]);
  *len = (uint32_t)(g.rowptr[v + 1] - r0);
  return true;
}

__device__ __forceinline__ unsigned long long wt_of(const KPlan& P, int u, uint32_t x) {
  const unsigned long long* t = P.wt[u];
  return t ? t[x] : 1ull;
}

__device__ __forceinline__ uint32_t hslot(uint32_t y) {
  return (y * 0x9E3779B1u) >> (32 - kHashBits);
}

// Wave-cooperative intersection of two sorted lists (already truncated).
// WRITE: store the hits in order at out.  LABEL: keep y only when bit lab[y]
// of labmask is set.  WEIGHT: also sum wt[y] (1 when wt is null) over the
// hits into *wsum.
// resident kernels interleave two batches; the out-of-core ones, which run
// closer to their register limit, do not
template <bool PG>
constexpr bool kDual = !PG && PM_ISECT_DUAL;

template <bool WRITE, bool LABEL, bool WEIGHT, bool DUAL = false>
__device__ uint32_t wave_isect(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb,
                               uint32_t labmask, const uint8_t* __restrict__ lab,
                               const unsigned long long* __restrict__ wt, uint32_t* out,
                               uint32_t lane, Counters& ct, unsigned long long* wsum) {
  if (na > nb) {
    const uint32_t* t = a;
    a = b;
    b = t;
    uint32_t tn = na;
    na = nb;
    nb = tn;
  }
  ct.alg += 4ull * (na + nb);
  uint32_t cnt = 0, lo = 0;
  unsigned long long ws = 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  if (DUAL) {
  // two batches of 64 per step: their binary searches (both from the
  // window start lo) are interleaved so their loads are in flight together
  for (uint32_t base = 0; base < na && lo < nb; base += 2 * kWave) {
    const uint32_t i0 = base + lane, i1 = base + kWave + lane;
    uint32_t x0 = 0, x1 = 0, p0 = lo, p1 = lo, n0 = 0, n1 = 0;
    if (i0 < na) {
      x0 = a[i0];
      n0 = nb - lo;
    }
    if (i1 < na) {
      x1 = a[i1];
      n1 = nb - lo;
    }
    while (n0 | n1) {
      const uint32_t h0 = n0 >> 1, h1 = n1 >> 1;
      const uint32_t v0 = n0 ? b[p0 + h0] : 0u, v1 = n1 ? b[p1 + h1] : 0u;
      if (n0) {
        if (v0 < x0) {
          p0 += h0 + 1;
          n0 -= h0 + 1;
        } else {
          n0 = h0;
        }
      }
      if (n1) {
        if (v1 < x1) {
          p1 += h1 + 1;
          n1 -= h1 + 1;
        } else {
          n1 = h1;
        }
      }
    }
    bool hit0 = i0 < na && p0 < nb && b[p0] == x0;
    bool hit1 = i1 < na && p1 < nb && b[p1] == x1;
    if (LABEL && hit0) hit0 = (labmask >> lab[x0]) & 1u;
    if (LABEL && hit1) hit1 = (labmask >> lab[x1]) & 1u;
    if (WEIGHT && hit0) ws += wt ? wt[x0] : 1ull;
    if (WEIGHT && hit1) ws += wt ? wt[x1] : 1ull;
    const uint64_t m0 = __ballot(hit0), m1 = __ballot(hit1);
    if (WRITE && hit0) out[cnt + __popcll(m0 & lt)] = x0;
    cnt += (uint32_t)__popcll(m0);
    if (WRITE && hit1) out[cnt + __popcll(m1 & lt)] = x1;
    cnt += (uint32_t)__popcll(m1);
    // the last valid element of this step bounds the next window
    const uint32_t left = na - base;
    const uint32_t last = (left > kWave ? min(left - kWave, (uint32_t)kWave) : left) - 1u;
    const uint32_t pl = left > kWave ? p1 : p0;
    lo = uni((uint32_t)__shfl((int)(i0 < na || i1 < na ? pl : nb), (int)last));
  }
  } else {
  for (uint32_t base = 0; base < na && lo < nb; base += kWave) {
    const uint32_t i = base + lane;
    bool hit = false;
    uint32_t x = 0, pos = nb;
    if (i < na) {
      x = a[i];
      pos = lo + below(b + lo, nb - lo, x);
      hit = pos < nb && b[pos] == x;
      if (LABEL && hit) hit = (labmask >> lab[x]) & 1u;
      if (WEIGHT && hit) ws += wt ? wt[x] : 1ull;
    }
    const uint64_t m = __ballot(hit);
    if (WRITE && hit) out[cnt + __popcll(m & lt)] = x;
    cnt += (uint32_t)__popcll(m);
    const uint32_t last = min(na - base, (uint32_t)kWave) - 1u;
    lo = uni((uint32_t)__shfl((int)pos, (int)last));
  }
  }
  if (WRITE) ct.alg += 4ull * cnt;
  if (WEIGHT) *wsum = wave_sum(ws);
  return cnt;
}

// Sum over y in a[0..n) with lab[y] == want (all y when !check) of wt[y]
// (1 when wt is null): wave-parallel, uniform result.
__device__ unsigned long long wave_label_sum(const uint32_t* a, uint32_t n, uint32_t want,
                                             bool check, const uint8_t* __restrict__ lab,
                                             const unsigned long long* __restrict__ wt,
                                             uint32_t lane, Counters& ct) {
  unsigned long long s = 0;
  for (uint32_t base = 0; base < n; base += kWave) {
    const uint32_t i = base + lane;
    if (i < n) {
      const uint32_t y = a[i];
      if (!check || lab[y] == want) s += wt ? wt[y] : 1ull;
    }
  }
  ct.alg += (4ull + (check ? 1 : 0) + (wt ? 8 : 0)) * n;
  return wave_sum(s);
}

// slot truncation: max over consumers of min over their symmetry partners
// matched so far; kInf when any consumer is unconstrained
__device__ __forceinline__ uint32_t slot_bound(const KOp& op, const WaveState& ws) {
  uint32_t ub = 0;
  for (int c = 0; c < op.ncons; c++) {
    uint32_t b = op.cons_b[c];
    if (b == 0) return kInf;
    uint32_t mn = kInf;
    while (b) {
      int j = __builtin_ctz(b);
      b &= b - 1;
      mn = min(mn, ws.phi[j]);
    }
    ub = max(ub, mn);
  }
  return ub;
}

// bound of core vertex u from its symmetry partners among 0..upto
__device__ __forceinline__ uint32_t vertex_bound(const KPlan& P, const WaveState& ws, int u,
                                                 int upto) {
  uint32_t b = P.B[u] & ((2u << upto) - 1u), ub = kInf;
  while (b) {
    int j = __builtin_ctz(b);
    b &= b - 1;
    ub = min(ub, ws.phi[j]);
  }
  return ub;
}

// The slot operations of depth d (vertex d just matched).  False when a row
// is not resident (out of core); the caller defers.
template <bool PG>
__device__ bool run_ops(const KGraph& g, const KPlan& P, WaveState& ws, int d, uint32_t* arena,
                        uint32_t cap, uint32_t lane, Counters& ct) {
  const uint32_t v = uni(ws.phi[d]);
  for (int o = P.op_begin[d]; o < P.op_end[d]; o++) {
    const KOp& op = P.ops[o];
    const uint32_t* nv;
    uint32_t dv, mp = 0;
    const bool ok = rows<PG>(g, v, P.segs ? op.seg : -1, &nv, &dv, &mp);
    if (!uni(ok)) {
      ws.blk = uni(mp);
      return false;
    }
    nv = uniptr(nv);
    dv = uni(dv);
    const uint32_t ub = P.labeled ? kInf : slot_bound(op, ws);
    const uint32_t nb = (ub == kInf) ? dv : uni(below(nv, dv, ub));
    if (op.kind == 0) {
      ws.sptr[op.slot] = nv;
      ws.slen[op.slot] = nb;
    } else {
      const uint32_t* a = uniptr(ws.sptr[op.parent]);
      uint32_t na = uni(ws.slen[op.parent]);
      if (ub != kInf) na = uni(below(a, na, ub));
      uint32_t* out = arena + (size_t)P.region[op.slot] * cap;
      uint32_t cnt;
      if (P.labeled && !(P.segs && op.seg >= 0))
        cnt = wave_isect<true, true, false, kDual<PG>>(a, na, nv, nb, op.labmask, g.lab,
                                                       nullptr, out, lane, ct, nullptr);
      else
        cnt = wave_isect<true, false, false, kDual<PG>>(a, na, nv, nb, 0, g.lab, nullptr, out,
                                                        lane, ct, nullptr);
      ws.sptr[op.slot] = out;
      ws.slen[op.slot] = uni(cnt);
      ct.nops++;
    }
  }
  return true;
}

// candidates of core vertex d+1 (truncated by its full symmetry bound)
__device__ __forceinline__ void setup_cand(const KPlan& P, WaveState& ws, int d) {
  const int s = P.cand_slot[d];
  uint32_t n = uni(ws.slen[s]);
  if (!P.labeled) {
    const uint32_t ub = vertex_bound(P, ws, d + 1, d);
    if (ub != kInf) n = uni(below(uniptr(ws.sptr[s]), n, ub));
  }
  ws.clen[d] = n;
  ws.cnext[d] = 0;
  ws.cmask[d] = 0;
}

// label / injectivity test of candidate x for core vertex u (lane-local)
__device__ __forceinline__ bool admit(const KGraph& g, const KPlan& P, const WaveState& ws, int u,
                                      bool check, uint32_t x) {
  if (P.labeled) return !check || g.lab[x] == P.lab[u];
  uint32_t xm = P.X[u];
  bool ok = true;
  while (xm) {
    int j = __builtin_ctz(xm);
    xm &= xm - 1;
    ok = ok && (x != ws.phi[j]);
  }
  return ok;
}

// Vertex-induced matching: x is adjacent to none of the images of the
// earlier non-neighbours of core vertex u (the DIFF passes against their
// rows, src/gpu_match.cu:570-587,734-738).  Resident rows only; with
// label-grouped rows only the label-lab[u] part of each row is searched
// (x has passed the label test).  Lane-local.
__device__ __forceinline__ bool outside(const KGraph& g, const KPlan& P, const WaveState& ws,
                                        int u, uint32_t x) {
  uint32_t xm = P.X[u];
  const int seg = (P.labeled && P.segs) ? (int)P.lab[u] : -1;
  bool ok = true;
  while (xm && ok) {
    const int j = __builtin_ctz(xm);
    xm &= xm - 1;
    const uint32_t* a;
    uint32_t na;
    rows<false>(g, ws.phi[j], seg, &a, &na);
    ok = !member(a, na, x);
  }
  return ok;
}

// next admitted candidate of core vertex d+1 and its index; false when done
template <bool PG>
__device__ __forceinline__ bool next_cand(const KGraph& g, const KPlan& P, WaveState& ws, int d,
                                          uint32_t lane, uint32_t& x, uint32_t& ix, Counters& ct) {
  for (;;) {
    uint64_t m = uni64(ws.cmask[d]);
    if (m) {
      const int i = __builtin_ctzll(m);
      ws.cmask[d] = m & (m - 1);
      x = uni(ws.cval[d][i]);
      ix = uni(ws.cbase[d]) + i;
      return true;
    }
    const uint32_t base = uni(ws.cnext[d]);
    const uint32_t n = uni(ws.clen[d]);
    if (base >= n) return false;
    ws.cbase[d] = base;
    ws.cnext[d] = base + kWave;
    const uint32_t* c = uniptr(ws.sptr[P.cand_slot[d]]);
    const uint32_t idx = base + lane;
    bool ok = idx < n;
    uint32_t v = 0;
    if (ok) {
      v = c[idx];
      ok = admit(g, P, ws, d + 1, P.cand_check[d], v);
      if (!PG && P.induced && ok) ok = outside(g, P, ws, d + 1, v);
    }
    ws.cval[d][lane] = v;
    ct.alg += 4ull * min(n - base, (uint32_t)kWave);
    ws.cmask[d] = __ballot(ok);
    __syncthreads();
  }
}

// Weighted count of the last core vertex once 0..k-2 are fixed (k == 3).
// *ok = false when a row is not resident.
template <bool PG>
__device__ unsigned long long last_count(const KGraph& g, const KPlan& P, WaveState& ws,
                                         uint32_t lane, Counters& ct, bool* ok) {
  const int u = P.k - 1;
  const int dl = P.k - 2;
  const uint32_t ub = P.labeled ? kInf : vertex_bound(P, ws, u, dl);
  const unsigned long long* w2 = P.wt[u];
  const bool segs = P.labeled && P.segs;
  const uint32_t* a;
  uint32_t na;
  *ok = true;
  if (P.last_kind == 0) {
    a = uniptr(ws.sptr[P.last_src]);
    na = uni(ws.slen[P.last_src]);
  } else {
    uint32_t mp = 0;
    const bool r = rows<PG>(g, uni(ws.phi[dl]), segs ? (int)P.lab[u] : -1, &a, &na, &mp);
    if (!uni(r)) {
      ws.blk = uni(mp);
      *ok = false;
      return 0;
    }
    a = uniptr(a);
    na = uni(na);
  }
  if (ub != kInf) na = uni(below(a, na, ub));
  if (P.last_kind != 2) {
    if (P.labeled) {
      const bool check = (P.last_kind == 0 && P.last_check) || (P.last_kind == 1 && !segs);
      if (!check && !w2) return na;
      ct.nops++;
      return wave_label_sum(a, na, P.lab[u], check, g.lab, w2, lane, ct);
    }
    const uint32_t xm = P.X[u];
    bool in = false;
    if (lane < (uint32_t)u && ((xm >> lane) & 1u)) in = member(a, na, ws.phi[lane]);
    return (unsigned long long)(na - (uint32_t)__popcll(__ballot(in)));
  }
  const uint32_t* p = uniptr(ws.sptr[P.last_src]);
  uint32_t np = uni(ws.slen[P.last_src]);
  if (ub != kInf) np = uni(below(p, np, ub));
  ct.nops++;
  if (P.labeled) {
    unsigned long long s = 0;
    uint32_t c;
    if (segs)
      c = wave_isect<false, false, true, kDual<PG>>(p, np, a, na, 0, g.lab, w2, nullptr, lane,
                                                    ct, &s);
    else
      c = wave_isect<false, true, true, kDual<PG>>(p, np, a, na, 1u << P.lab[u], g.lab, w2,
                                                   nullptr, lane, ct, &s);
    return w2 ? s : (unsigned long long)c;
  }
  uint32_t cnt = uni(wave_isect<false, false, false, kDual<PG>>(p, np, a, na, 0, g.lab, nullptr,
                                                                nullptr, lane, ct, nullptr));
  const uint32_t xm = P.X[u];
  bool in = false;
  if (lane < (uint32_t)u && ((xm >> lane) & 1u)) {
    const uint32_t y = ws.phi[lane];
    in = member(a, na, y) && member(p, np, y);
  }
  cnt -= (uint32_t)__popcll(__ballot(in));
  return cnt;
}

// Vertex-induced count of the last core vertex once 0..k-2 are fixed
// (resident rows).  Each element of its list — or of the shorter of the two
// lists it is the intersection of — is tested lane by lane: membership in
// the other list, label, injectivity (unlabeled) and non-adjacency to the
// images of its earlier non-neighbours.
__device__ unsigned long long last_count_induced(const KGraph& g, const KPlan& P, WaveState& ws,
                                                 uint32_t lane, Counters& ct) {
  const int u = P.k - 1;
  const int dl = P.k - 2;
  const uint32_t ub = P.labeled ? kInf : vertex_bound(P, ws, u, dl);
  const bool segs = P.labeled && P.segs;
  const uint32_t* a;
  uint32_t na;
  if (P.last_kind == 0) {
    a = uniptr(ws.sptr[P.last_src]);
    na = uni(ws.slen[P.last_src]);
  } else {
    rows<false>(g, uni(ws.phi[dl]), segs ? (int)P.lab[u] : -1, &a, &na);
    a = uniptr(a);
    na = uni(na);
  }
  if (ub != kInf) na = uni(below(a, na, ub));
  const uint32_t* b = nullptr;
  uint32_t nb = 0;
  if (P.last_kind == 2) {
    b = uniptr(ws.sptr[P.last_src]);
    nb = uni(ws.slen[P.last_src]);
    if (ub != kInf) nb = uni(below(b, nb, ub));
    if (nb < na) {
      const uint32_t* t = a;
      a = b;
      b = t;
      const uint32_t tn = na;
      na = nb;
      nb = tn;
    }
  }
  // label-exact lists need no label test (as in last_count)
  const bool check = P.labeled && (P.last_kind == 0 ? P.last_check != 0 : !segs);
  unsigned long long s = 0;
  for (uint32_t base = 0; base < na; base += kWave) {
    const uint32_t i = base + lane;
    if (i < na) {
      const uint32_t y = a[i];
      bool ok = !b || member(b, nb, y);
      ok = ok && admit(g, P, ws, u, check, y);
      ok = ok && outside(g, P, ws, u, y);
      s += ok ? 1ull : 0ull;
    }
  }
  ct.alg += 4ull * ((uint64_t)na + nb);
  ct.nops++;
  return wave_sum(s);
}

// Write one deferred piece (uniform call): depth and the candidate range
// [a, b) at that depth below the current path.
__device__ void emit_one(const DeferOut& O, const WaveState& ws, int depth, uint32_t a,
                         uint32_t b, uint32_t blk, uint32_t lane) {
  unsigned long long slot = 0;
  if (lane == 0) slot = atomicAdd(O.count, 1ull);
  slot = uni64(__shfl(slot, 0));
  if (lane == 0 && slot < O.cap) {
    Resume& R = O.buf[slot];
    R.edge = ws.je;
    R.r = ws.phi[0];
    R.c = ws.phi[1];
    R.lo = ws.jlo;
    R.hi = ws.jhi;
    R.depth = depth;
    R.end = b;
    for (int d = 1; d < depth && d < kMaxDepth; d++) R.idx[d] = ws.cidx[d];
    if (depth >= 1 && depth < kMaxDepth) R.idx[depth] = a;
    for (int d = 0; d <= depth && d < kMaxDepth; d++) R.phi[d] = ws.phi[d];
    R.blk = blk;
    R.pat = ws.pat;
  }
}

// the re-split budget of the current search has run out and the piece
// buffer has room (uniform)
__device__ __forceinline__ bool resplit_due(const DeferOut& O, const WaveState& ws) {
  return O.budget && uni((uint32_t)(wall_clock64() - uni64(ws.t0) > O.budget)) &&
         uni64(*(volatile unsigned long long*)O.count) < O.soft;
}

// A resident search stopped by its re-split budget (run_job recorded
// rs_depth / rs_stop, right after the subtree of candidate cidx[rs_depth]
// was finished): the rest of that depth goes out as up to 64 pieces, the
// rest of each shallower open depth as one piece.  Called by the kernels
// after run_job returns, outside the search loop's register pressure.
template <bool PG, bool RS>
__device__ __forceinline__ void resplit_flush(const DeferOut& O, WaveState& ws, uint32_t lane) {
  if (PG || !RS) return;
  const int dd = (int)uni((uint32_t)ws.rs_depth);
  if (dd < 0) return;
  const int from = (int)uni((uint32_t)ws.rs_from);
  const uint32_t a0 = from >= 0 ? (uint32_t)from : uni(ws.cidx[dd]) + 1, b0 = uni(ws.clen[dd]);
  if (a0 < b0) {
    const uint32_t np = min(b0 - a0, 64u), per = (b0 - a0 + np - 1) / np;
    for (uint32_t a = a0; a < b0; a += per) emit_one(O, ws, dd, a, min(a + per, b0), 0, lane);
  }
  for (int d = (int)uni((uint32_t)ws.rs_stop); d < dd; d++) {
    const uint32_t s0 = uni(ws.cidx[d]) + 1, s1 = uni(ws.clen[d]);
    if (s0 < s1) emit_one(O, ws, d, s0, s1, 0, lane);
  }
  if (lane == 0) ws.rs_depth = -1;
  __syncthreads();
}

// Lane-parallel emission: every lane with `mine` writes the piece "candidate
// i at this depth" (range [i, i+1)), stopped by package blk.
__device__ void emit_lanes(const DeferOut& O, const WaveState& ws, int depth, bool mine,
                           uint32_t i, uint32_t blk, uint32_t lane) {
  const uint64_t mm = __ballot(mine);
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(O.count, (unsigned long long)__popcll(mm));
  base = uni64(__shfl(base, 0));
  if (!mine) return;
  const unsigned long long slot = base + __popcll(mm & ((1ull << lane) - 1ull));
  if (slot >= O.cap) return;
  Resume& R = O.buf[slot];
  R.edge = ws.je;
  R.r = ws.phi[0];
  R.c = ws.phi[1];
  R.lo = ws.jlo;
  R.hi = ws.jhi;
  R.depth = depth;
  R.end = i + 1;
  for (int d = 1; d < depth && d < kMaxDepth; d++) R.idx[d] = ws.cidx[d];
  if (depth < kMaxDepth) R.idx[depth] = i;
  for (int d = 0; d <= depth && d < kMaxDepth; d++) R.phi[d] = ws.phi[d];
  R.blk = blk;
  R.pat = ws.pat;
}

// The last two core vertices u1 = d+1, u2 = d+2 (d = k-3), fused: weighted
// number of matches of (u1, u2) over the candidates of u1 in
// [cnext[d], clen[d]).  Batches of 64 candidates are all-or-nothing: when a
// needed row is not resident, *fail = the batch's first index and the sum of
// the finished batches is returned.  Uniform result.
template <bool PG, bool RS = false>
__device__ unsigned long long fused_tail(const KGraph& g, const KPlan& P, WaveState& ws, int d,
                                         uint32_t lane, Counters& ct, const DeferOut& O,
                                         bool* stopped) {
  const int u1 = d + 1, u2 = d + 2;
  const bool lab = P.labeled;
  const unsigned long long wd = uni64(ws.wacc[d]);
  const unsigned long long* w1 = P.wt[u1];
  const unsigned long long* w2 = P.wt[u2];
  // a resumed group walks its listed candidates (ltmp, indices lidx)
  const bool listed = uni(ws.lcount) != 0;
  const uint32_t* cand = listed ? (const uint32_t*)ws.ltmp : uniptr(ws.sptr[P.cand_slot[d]]);
  const uint32_t cbeg = uni(ws.cnext[d]), cend = uni(ws.clen[d]);
  // u2's bound from vertices 0..d, and whether phi(u2) < phi(u1)
  const uint32_t ub_fixed = lab ? kInf : vertex_bound(P, ws, u2, d);
  const bool sym_x = !lab && ((P.B[u2] >> u1) & 1u);
  const uint32_t dropm = lab ? 0u : (P.X[u2] & ((2u << d) - 1u));
  const bool x_drop = !lab && ((P.X[u2] >> u1) & 1u);
  unsigned long long acc = 0, dropc = 0;

  if (P.last_kind == 0) {
    // u2's set does not depend on u1's image; no row of x is read
    const uint32_t* S = uniptr(ws.sptr[P.last_src]);
    const uint32_t nS = uni(ws.slen[P.last_src]);
    unsigned long long Lc = 0;
    uint32_t inS = 0;  // bit j: phi(j) in S (j in dropm)
    if (lab) {
      Lc = (!P.last_check && !w2)
               ? (unsigned long long)nS
               : wave_label_sum(S, nS, P.lab[u2], P.last_check, g.lab, w2, lane, ct);
    } else {
      bool in = false;
      if (lane <= (uint32_t)d && ((dropm >> lane) & 1u)) in = member(S, nS, ws.phi[lane]);
      inS = (uint32_t)__ballot(in);
    }
    for (uint32_t base = cbeg; base < cend; base += kWave) {
      const uint32_t i = base + lane;
      if (i < cend) {
        const uint32_t x = cand[i];
        if (admit(g, P, ws, u1, P.cand_check[d], x)) {
          if (lab) {
            acc += wd * (w1 ? w1[x] : 1ull) * Lc;
          } else {
            const uint32_t ub = sym_x ? min(ub_fixed, x) : ub_fixed;
            const uint32_t c = below(S, nS, ub);
            uint32_t dr = 0, b = inS;
            while (b) {
              int j = __builtin_ctz(b);
              b &= b - 1;
              dr += ws.phi[j] < ub;
            }
            if (x_drop && member(S, c, x)) dr++;
            acc += c - dr;
          }
        }
      }
    }
    ct.alg += 8ull * (cend - cbeg);
    ct.nops++;
    return wave_sum(acc);
  }

  if (P.last_kind == 1 && !lab) {
    // u2's set is N(x): per-lane binary searches
    for (uint32_t base = cbeg; base < cend; base += kWave) {
      if (RS && !PG && base > cbeg && !listed && resplit_due(O, ws)) {
        // re-split: the candidates from this batch on become pieces
        if (lane == 0) {
          ws.rs_depth = d;
          ws.rs_from = (int32_t)base;
        }
        *stopped = true;
        break;
      }
      const uint32_t i = base + lane;
      const uint32_t* nx = nullptr;
      uint32_t dx = 0, x = 0, mp = 0;
      bool ok = false, miss = false;
      if (i < cend) {
        x = cand[i];
        ok = admit(g, P, ws, u1, P.cand_check[d], x);
        if (ok) miss = !rows<PG>(g, x, -1, &nx, &dx, &mp);
      }
      bool stop_here = false;
      uint32_t rest_from = 0;
      const uint64_t mm = __ballot(miss);
      if (mm) {
        if (!listed && uni64(*(volatile unsigned long long*)O.count) > O.soft) {
          // too many pieces already: finish the owners before the first
          // miss; the rest of the range becomes one piece
          const int f = __builtin_ctzll(mm);
          ws.blk = uni((uint32_t)__shfl((int)mp, f));
          rest_from = base + f;
          stop_here = true;
          if ((int)lane >= f) ok = false;
        } else {
          // owners whose row is not resident become deferred pieces
          emit_lanes(O, ws, d, miss, listed && i < cend ? ws.lidx[i] : i, mp, lane);
          ok = ok && !miss;
        }
      }
      if (ok) {
        const uint32_t ub = sym_x ? min(ub_fixed, x) : ub_fixed;
        const uint32_t c = below(nx, dx, ub);
        uint32_t dr = 0, b = dropm;
        while (b) {
          int j = __builtin_ctz(b);
          b &= b - 1;
          const uint32_t y = ws.phi[j];
          dr += (y < ub) && member(nx, c, y);
        }
        acc += c - dr;
        ct.alg += 20ull;
      }
      if (stop_here) {
        emit_one(O, ws, d, rest_from, cend, uni(ws.blk), lane);
        *stopped = true;
        break;
      }
    }
    ct.nops++;
    return wave_sum(acc);
  }

  // Edge-parallel pass over the candidates' neighbour lists.  kind 1
  // (labeled): every y in N(x) counts; kind 2: y must lie in the parent set
  // P (LDS hash when it fits, else binary search in P).
  const bool isect = P.last_kind == 2;
  const uint32_t* Pset = isect ? uniptr(ws.sptr[P.last_src]) : nullptr;
  uint32_t nP = isect ? uni(ws.slen[P.last_src]) : 0;
  if (isect && ub_fixed != kInf) nP = uni(below(Pset, nP, ub_fixed));
  const bool use_hash = isect && nP <= (uint32_t)(kHashCap / 2);
  if (use_hash) {
    for (int i = lane; i < kHashCap; i += kWave) ws.hash[i] = kEmpty;
    __syncthreads();
    for (uint32_t i = lane; i < nP; i += kWave) {
      const uint32_t y = Pset[i];
      uint32_t h = hslot(y);
      while (atomicCAS(&ws.hash[h], kEmpty, y) != kEmpty) h = (h + 1) & (kHashCap - 1);
    }
    __syncthreads();
    ct.alg += 4ull * nP;
  }
  uint32_t inP = 0;  // bit j: phi(j) in P (unlabeled injectivity)
  if (isect && !lab) {
    bool in = false;
    if (lane <= (uint32_t)d && ((dropm >> lane) & 1u)) in = member(Pset, nP, ws.phi[lane]);
    inP = (uint32_t)__ballot(in);
  }
  const bool segs = lab && P.segs;
  const int seg2 = segs ? (int)P.lab[u2] : -1;  // labeled: u2's label part only
  const bool check2 = lab && !segs;
  const bool pair = P.pair;
  const uint64_t hub_lim = isect ? (uint64_t)P.hub_factor * (uint64_t)max(nP, 1u) : ~0ull;
  for (uint32_t base = cbeg; base < cend; base += kWave) {
    if (RS && !PG && base > cbeg && !listed && resplit_due(O, ws)) {
      // re-split: the candidates from this batch on become pieces
      if (lane == 0) {
        ws.rs_depth = d;
        ws.rs_from = (int32_t)base;
      }
      *stopped = true;
      break;
    }
    const uint32_t i = base + lane;
    uint32_t x = 0, dx = 0, ub = ub_fixed;
    unsigned long long wx = 0, wb = 0;
    const uint32_t* nx = nullptr;
    uint32_t mp = 0;
    bool ok = false, hub = false, miss = false;
    if (i < cend) {
      x = cand[i];
      ok = admit(g, P, ws, u1, P.cand_check[d], x);
      if (ok) miss = !rows<PG>(g, x, seg2, &nx, &dx, &mp);
    }
    bool stop_here = false;
    uint32_t rest_from = 0;
    const uint64_t mm = __ballot(miss);
    if (mm) {
      if (!listed && uni64(*(volatile unsigned long long*)O.count) > O.soft) {
        // too many pieces already: finish the owners before the first miss;
        // the rest of the range becomes one piece
        const int f = __builtin_ctzll(mm);
        ws.blk = uni((uint32_t)__shfl((int)mp, f));
        rest_from = base + f;
        stop_here = true;
        if ((int)lane >= f) ok = false;
      } else {
        // owners whose row is not resident become deferred pieces
        emit_lanes(O, ws, d, miss, listed && i < cend ? ws.lidx[i] : i, mp, lane);
        ok = ok && !miss;
      }
    }
    if (ok) {
      wx = wd * (w1 ? w1[x] : 1ull);
      if (pair) {
        // only partners y > x: the pair {x,y} is counted once, in both roles
        const uint32_t s0 = below(nx, dx, x + 1);
        nx += s0;
        dx -= s0;
        wb = wd * (w2 ? w2[x] : 1ull);
      }
      if (sym_x) ub = min(ub, x);
      if (!lab) {
        uint32_t b = inP;
        while (b) {
          int j = __builtin_ctz(b);
          b &= b - 1;
          const uint32_t y = ws.phi[j];
          dropc += (y < ub) && member(nx, dx, y);
        }
      }
      hub = (uint64_t)dx > hub_lim;
    }
    // exclusive scan of the owners' list lengths
    const uint32_t len = (ok && !hub) ? dx : 0;
    uint32_t incl = len;
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)incl, off);
      if ((int)lane >= off) incl += t;
    }
    ws.own_off[lane] = incl - len;
    if (lane == kWave - 1) ws.own_off[kWave] = incl;
    ws.own_ub[lane] = ub;
    ws.own_w[lane] = (ok && !hub) ? wx : 0ull;
    ws.own_b[lane] = (ok && !hub) ? wb : 0ull;
    ws.own_ptr[lane] = (uint64_t)nx;
    __syncthreads();
    const uint32_t total = uni(ws.own_off[kWave]);
    ct.alg += 4ull * total + 16ull * kWave;
    // kWalk entries per lane per step: their row loads, then their weight
    // gathers, are issued together so the latencies overlap
    for (uint32_t t0 = lane; t0 < total; t0 += kWalk * kWave) {
      uint32_t yv[kWalk], ov[kWalk];
#pragma unroll
      for (int q = 0; q < kWalk; q++) {
        const uint32_t t = t0 + q * kWave;
        // owner: last o with own_off[o] <= t
        uint32_t lo = 0, hi = kWave;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (ws.own_off[mid] <= t) lo = mid; else hi = mid;
        }
        ov[q] = lo;
        yv[q] = t < total ? ((const uint32_t*)ws.own_ptr[lo])[t - ws.own_off[lo]] : kInf;
      }
      unsigned long long wv2[kWalk], wv1[kWalk];
#pragma unroll
      for (int q = 0; q < kWalk; q++) {
        const uint32_t y = yv[q];
        bool hit = y < ws.own_ub[ov[q]];  // also false past the end (y = kInf)
        if (hit && isect) {
          if (use_hash) {
            uint32_t h = hslot(y);
            uint32_t e;
            while ((e = ws.hash[h]) != kEmpty && e != y) h = (h + 1) & (kHashCap - 1);
            hit = e == y;
          } else {
            hit = member(Pset, nP, y);
          }
        }
        if (hit && check2) hit = g.lab[y] == P.lab[u2];
        wv2[q] = hit ? (w2 ? w2[y] : 1ull) : 0ull;
        wv1[q] = (hit && pair) ? (w1 ? w1[y] : 1ull) : 0ull;
      }
#pragma unroll
      for (int q = 0; q < kWalk; q++) {
        acc += ws.own_w[ov[q]] * wv2[q];
        if (pair) acc += ws.own_b[ov[q]] * wv1[q];
      }
    }
    __syncthreads();
    // owners whose list is far longer than P: one by one, iterating P and
    // searching N(x)
    uint64_t hubs = __ballot(hub);
    while (hubs) {
      const int o = __builtin_ctzll(hubs);
      hubs &= hubs - 1;
      const uint32_t xo = uni((uint32_t)__shfl((int)x, o));
      const uint32_t ubo = uni((uint32_t)__shfl((int)ub, o));
      const unsigned long long wo = uni64(__shfl(wx, o));
      const unsigned long long wbo = uni64(__shfl(wb, o));
      const uint32_t* nxo = (const uint32_t*)uni64(ws.own_ptr[o]);
      uint32_t nxn = uni((uint32_t)__shfl((int)dx, o));
      if (ubo != kInf) nxn = uni(below(nxo, nxn, ubo));
      const uint32_t* Pp = Pset;
      uint32_t nPo = ubo != kInf ? uni(below(Pset, nP, ubo)) : nP;
      if (pair) {  // partners y > x only (nxo already starts past x)
        const uint32_t s0 = uni(below(Pset, nPo, xo + 1));
        Pp = Pset + s0;
        nPo -= s0;
      }
      unsigned long long s = 0, s1 = 0;
      uint32_t c;
      if (lab && !segs)
        c = wave_isect<false, true, true, kDual<PG>>(Pp, nPo, nxo, nxn, 1u << P.lab[u2], g.lab,
                                                     w2, nullptr, lane, ct, &s);
      else if (lab)
        c = wave_isect<false, false, true, kDual<PG>>(Pp, nPo, nxo, nxn, 0, g.lab, w2, nullptr,
                                                      lane, ct, &s);
      else
        c = wave_isect<false, false, false, kDual<PG>>(Pp, nPo, nxo, nxn, 0, g.lab, nullptr,
                                                       nullptr, lane, ct, nullptr);
      if (pair) {
        s1 = c;
        if (w1)
          wave_isect<false, false, true, kDual<PG>>(Pp, nPo, nxo, nxn, 0, g.lab, w1, nullptr,
                                                    lane, ct, &s1);
      }
      if (lane == 0) acc += wo * (lab && w2 ? s : (unsigned long long)c) + (pair ? wbo * s1 : 0ull);
    }
    if (stop_here) {
      emit_one(O, ws, d, rest_from, cend, uni(ws.blk), lane);
      *stopped = true;
      break;
    }
  }
  ct.nops++;
  return wave_sum(acc) - wave_sum(dropc);
}

// DFS over one job (edge e = (r,c)) restricted to vertex-2 candidates
// [lo, hi).  With `resume`, the deferred piece in ws.rdepth/ridx/rend: rebuild
// its path and search only its candidate range at its depth.  Out of core, a
// subtree that needs a non-resident row is written to O as a deferred piece
// and the search goes on with the next candidate, so each piece is searched
// exactly once.  Phase 1 (may_split) may hand a tail of the range to the
// task buffer.  RS: resident re-split compiled in (only kernels of batches
// that can use it carry its registers).
template <bool PG, bool RS = false>
__device__ unsigned long long run_job(const KGraph& g, const KPlan& P, WaveState& ws,
                                      uint32_t* arena, uint32_t cap, uint32_t lane, uint32_t r,
                                      uint32_t c, uint64_t e, uint32_t lo, uint32_t hi,
                                      bool resume, bool may_split, Task* tasks,
                                      unsigned long long* ntasks, uint64_t task_cap,
                                      const DeferOut& O, Counters& ct) {
  const int k = P.k;
  // resident re-split: a search running past its budget hands its untried
  // candidates back as pieces (below)
  if (RS && !PG && O.budget && lane == 0) ws.t0 = wall_clock64();
  ws.je = e;
  ws.jlo = lo;
  ws.jhi = hi;
  ws.phi[0] = r;
  ws.phi[1] = c;
  const unsigned long long w01 = uni64(wt_of(P, 0, r) * wt_of(P, 1, c));
  ws.wacc[1] = w01;
  if (k == 2) return w01;
  if (!run_ops<PG>(g, P, ws, 0, arena, cap, lane, ct)) {
    emit_one(O, ws, 0, 0, 0, uni(ws.blk), lane);
    return 0;
  }
  // vertex-induced plans (resident) enumerate up to vertex k-2 and count the
  // last vertex with last_count_induced; they never take the fused tail
  const bool induced = !PG && P.induced;
  if (k == 3) {
    bool ok = true;
    const unsigned long long v = induced ? last_count_induced(g, P, ws, lane, ct)
                                         : last_count<PG>(g, P, ws, lane, ct, &ok);
    if (!ok) {
      emit_one(O, ws, 0, 0, 0, uni(ws.blk), lane);
      return 0;
    }
    return w01 * v;
  }
  if (!run_ops<PG>(g, P, ws, 1, arena, cap, lane, ct)) {
    emit_one(O, ws, 0, 0, 0, uni(ws.blk), lane);
    return 0;
  }
  setup_cand(P, ws, 1);
  const uint32_t n2 = uni(ws.clen[1]);
  const uint32_t b = min(lo, n2);
  uint32_t en = min(hi, n2);
  const uint32_t W = P.split;
  if (may_split && W > 0 && en > b + 1) {
    // Work-aware split of the vertex-2 candidates.  Candidate x costs
    // 1 + |list the next level reads from x|; a new range starts at every
    // candidate whose exclusive running cost crosses a multiple of W.  The
    // wave keeps the first range and writes the others as tasks (hi = 0:
    // up to the next task's lo; hi = ~0: to the end of the list).
    const uint32_t* cand = uniptr(ws.sptr[P.cand_slot[1]]);
    const int cseg = (P.labeled && P.segs && k == 4) ? (int)P.lab[3] : -1;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t nstart = 0;
    uint32_t keep_end = en;
    unsigned long long tb = 0;
    bool write = false;
    for (int pass = 0; pass < 2; pass++) {
      uint64_t run = 0, prev_c = 0, rank0 = 0;
      for (uint32_t base = b; base < en; base += kWave) {
        const uint32_t i = base + lane;
#if PM_SPLIT_QUAD
        // deeper cores: a candidate's subtree grows with a power of its list
        // (scaled so that a 64-entry list costs 64 at every depth)
        uint64_t cost = 0;
        if (i < en) {
          const uint64_t ln = row_len(g, cand[i], cseg);
          uint64_t cw = ln;
#if PM_SPLIT_QUAD >= 2
          const int extra = k - 4;  // len^(k-3) / 64^(k-4)
#else
          const int extra = k >= 5 ? 1 : 0;  // len^2 / 64
#endif
          // capped at 2^40 per candidate (no overflow: cw * ln <= 2^46 here)
          for (int q = 0; q < extra; q++)
            cw = (ln && cw > (1ull << 46) / ln) ? (1ull << 40) : min(cw * ln / 64, 1ull << 40);
          cost = 1ull + cw;
        }
#else
        const uint64_t cost = i < en ? 1ull + row_len(g, cand[i], cseg) : 0;
#endif
        uint64_t incl = cost;
        for (int o = 1; o < kWave; o <<= 1) {
          const uint64_t t = __shfl_up(incl, o);
          if ((int)lane >= o) incl += t;
        }
        const uint64_t before = run + incl - cost;
        const uint64_t up = __shfl_up(cost, 1);
        const uint64_t cprev = lane > 0 ? up : prev_c;
        const bool starts = i < en && i > b && (before / W) != ((before - cprev) / W);
        const uint64_t sm = __ballot(starts);
        if (pass == 1 && write && starts) {
          const uint64_t rank = rank0 + __popcll(sm & lt);
          Task tk;
          tk.edge = e;
          tk.lo = i;
          tk.hi = rank + 1 == nstart ? 0xFFFFFFFFu : 0u;
          tk.r = r;
          tk.c = c;
          tk.pat = ws.pat;
          tasks[tb + rank] = tk;
        }
        if (pass == 1 && sm && rank0 == 0 && keep_end == en)
          keep_end = uni((uint32_t)__shfl((int)i, __builtin_ctzll(sm)));
        rank0 += __popcll(sm);
        run += uni64(wave_sum(cost));
        const uint32_t ll = min(en - base, (uint32_t)kWave) - 1u;
        prev_c = uni64(__shfl(cost, ll));
      }
      if (pass == 0) {
        nstart = rank0;
        if (nstart == 0) break;
        if (lane == 0) tb = atomicAdd(ntasks, (unsigned long long)nstart);
        tb = uni64(__shfl(tb, 0));
        write = tb + nstart <= task_cap;
        if (!write) {
          // buffer full: blank the reserved slots inside it, keep all work
          for (uint64_t t = tb + lane; t < min(tb + nstart, task_cap); t += kWave) {
            Task tk;
            tk.edge = e;
            tk.lo = 1;
            tk.hi = 1;
            tk.r = r;
            tk.c = c;
            tk.pat = ws.pat;
            tasks[t] = tk;
          }
          break;
        }
      }
    }
    if (write) en = keep_end;
  }
  ws.cnext[1] = b;
  ws.clen[1] = en;
  ws.cmask[1] = 0;
  int d = 1, stop = 1;  // stop: depth whose exhaustion ends this search
  unsigned long long cnt = 0;
  bool fused_now = (k == 4) && !induced;  // the fused level is depth k-3
  if (resume && uni(ws.rdepth) >= 1) {
    // rebuild the path phi(2..D) from the recorded candidate indices
    const int D = (int)uni(ws.rdepth);
    for (int dd = 1; dd < D; dd++) ws.cidx[dd] = uni(ws.ridx[dd]);
    for (int dd = 1; dd < D; dd++) {
      const uint32_t ix = uni(ws.ridx[dd]);
      const uint32_t x = uni(uniptr(ws.sptr[P.cand_slot[dd]])[ix]);
      ws.phi[dd + 1] = x;
      ws.wacc[dd + 1] = uni64(ws.wacc[dd]) * wt_of(P, dd + 1, x);
      if (!run_ops<PG>(g, P, ws, dd + 1, arena, cap, lane, ct)) {
        const uint32_t nl = uni(ws.lcount);
        if (nl == 0) emit_one(O, ws, D, uni(ws.ridx[D]), uni(ws.rend), uni(ws.blk), lane);
        for (uint32_t q = 0; q < nl; q++) {
          const uint32_t iq = uni(ws.lidx[q]);
          emit_one(O, ws, D, iq, iq + 1, uni(ws.blk), lane);
        }
        return 0;
      }
      setup_cand(P, ws, dd + 1);
    }
    const uint32_t nl = uni(ws.lcount);
    if (nl) {
      // a group of one-candidate pieces: their vertices in index order
      const uint32_t* cs = uniptr(ws.sptr[P.cand_slot[D]]);
      if (lane < nl) ws.ltmp[lane] = cs[ws.lidx[lane]];
      __syncthreads();
      ws.cnext[D] = 0;
      ws.clen[D] = nl;
    } else {
      ws.cnext[D] = uni(ws.ridx[D]);
      ws.clen[D] = min(uni(ws.clen[D]), uni(ws.rend));
    }
    ws.cmask[D] = 0;
    d = D;
    stop = D;
    fused_now = (d == k - 3) && !induced;
  }
  // the rest of the search above depth d: siblings after the current
  // candidate at every depth stop..d-1, one piece each
  auto emit_rest = [&](int upto) {
    for (int dd = stop; dd < upto; dd++) {
      const uint32_t a0 = uni(ws.cidx[dd]) + 1, b0 = uni(ws.clen[dd]);
      if (a0 < b0) emit_one(O, ws, dd, a0, b0, uni(ws.blk), lane);
    }
  };
  // Resident re-split (work stealing, src/gpu_match.cu:21-148 splits a
  // victim's shallow level the same way): called right after the subtree of
  // candidate cidx[d] at depth d is finished, so every search makes
  // progress before it splits.  Past the budget, the rest of depth d goes
  // out as up to 64 pieces and the rest of the shallower depths as one
  // piece each; later rounds resume them over all waves.
  auto resplit = [&](int dd) -> bool {
    if (PG || !RS || O.budget == 0) return false;
    if (!uni((uint32_t)(wall_clock64() - uni64(ws.t0) > O.budget))) return false;
    if (uni64(*(volatile unsigned long long*)O.count) >= O.soft) return false;
    // the pieces are written by resplit_flush once the search has returned
    if (lane == 0) {
      ws.rs_depth = dd;
      ws.rs_stop = stop;
      ws.rs_from = -1;
    }
    return true;
  };
  for (;;) {
    if (fused_now) {
      fused_now = false;
      bool stopped = false;
      cnt += fused_tail<PG, RS>(g, P, ws, d, lane, ct, O, &stopped);
      if (stopped) {
        if (RS && (int)uni((uint32_t)ws.rs_depth) >= 0) {
          // re-split inside the fused step: resplit_flush writes its rest
          // and the rest of the shallower depths
          if (lane == 0) ws.rs_stop = stop;
          return cnt;
        }
        emit_rest(d);
        return cnt;
      }
      if (d == stop) break;
      d--;
      if (resplit(d)) return cnt;
      continue;
    }
    uint32_t x, ix;
    if (!next_cand<PG>(g, P, ws, d, lane, x, ix, ct)) {
      if (d == stop) break;
      d--;
      if (resplit(d)) return cnt;
      continue;
    }
    ws.cidx[d] = ix;
    ws.phi[d + 1] = x;
    ws.wacc[d + 1] = uni64(ws.wacc[d]) * wt_of(P, d + 1, x);
    if (induced && d + 1 == k - 2) {
      cnt += uni64(ws.wacc[d + 1]) * last_count_induced(g, P, ws, lane, ct);
      continue;
    }
    if (!run_ops<PG>(g, P, ws, d + 1, arena, cap, lane, ct)) {
      if (uni64(*(volatile unsigned long long*)O.count) > O.soft) {
        // too many pieces already: stop, leaving the rest as a few pieces
        emit_one(O, ws, d, ix, uni(ws.clen[d]), uni(ws.blk), lane);
        emit_rest(d);
        return cnt;
      }
      // the subtree of candidate ix waits for its row; go on with the rest
      emit_one(O, ws, d, ix, ix + 1, uni(ws.blk), lane);
      continue;
    }
    setup_cand(P, ws, d + 1);
    d++;
    fused_now = (d == k - 3) && !induced;
  }
  return cnt;
}

// largest r >= lo with rowptr[r] <= e (rowptr[lo] <= e holds), by galloping
__device__ __forceinline__ uint32_t source_of(const KGraph& g, uint64_t e, uint32_t lo) {
  uint32_t step = 1;
  const uint32_t n = (uint32_t)g.n;
  while (lo + step < n && g.rowptr[lo + step] <= e) {
    lo += step;
    step <<= 1;
  }
  uint32_t hi = min(lo + step, n);
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (g.rowptr[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// source of edge e: start at the indexed source of edge 64*(e/64), gallop
// index i of the last starts[i] <= s, galloping up from lo
__device__ __forceinline__ uint32_t gallop(const uint64_t* starts, uint32_t n, uint64_t s,
                                           uint32_t lo) {
  uint32_t step = 1;
  while (lo + step < n && starts[lo + step] <= s) {
    lo += step;
    step <<= 1;
  }
  uint32_t hi = min(lo + step, n);
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= s) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t source_full(const KGraph& g, uint64_t e) {
  return source_of(g, e, g.src64[e >> 6]);
}

// One launch runs npat patterns (a batch; npat = 1 for a single pattern).
// Phase-1 work units are (pattern, chunk of edges) pairs numbered pattern-
// major and claimed from one cursor; tasks and deferred pieces carry their
// pattern.
constexpr uint64_t kSampleBlock = 4096;  // edges per block of a sampled run

struct Launch {
  KGraph g;
  const KPlan* plans;
  uint32_t npat;
  uint32_t chunk;
  uint64_t nchunk;             // edge chunks per pattern
  uint64_t npos;               // job positions per pattern (edges, or sampled edges)
  uint32_t sample;             // > 1: only every sample-th block of kSampleBlock edges
  uint32_t pad2;
  unsigned long long* cursor;  // [0] units [1] task claim [2] tasks [3] piece claim [4] pieces
  uint64_t edge_lo, edge_hi;
  uint32_t cap;
  uint32_t pad;
  uint32_t* arena;
  uint64_t arena_stride;
  unsigned long long* out;       // statistics: alg/ops per phase
  unsigned long long* pcount;    // per pattern: count, jobs
  unsigned long long* per_edge;  // optional (npat == 1), indexed e - edge_lo
  Task* tasks;
  uint64_t task_cap;
  const Resume* rin;  // deferred pieces resumed this round
  uint64_t nrin;
  const unsigned long long* nrin_dev;  // when set, the count is read from here
  Resume* rout;  // deferred pieces for the next round
  uint64_t rout_cap;
  uint64_t soft;  // beyond this many new pieces, searches stop at their first miss
  uint64_t budget;  // resident re-split budget in wall-clock ticks (0 = off)
  uint32_t claim;   // pieces claimed per atomic by k_resume (0 = kResumeClaim)
  uint32_t pad3;
  // learned job order: position s lies in the row of root_order[i] where
  // root_start[i] <= s < root_start[i+1] (null: CSR order)
  const uint32_t* root_order;
  const uint64_t* root_start;
  unsigned long long* root_work;  // trace: work of the jobs rooted at each vertex
};

// edge of job position s (positions enumerate the sampled edges in order)
__device__ __forceinline__ uint64_t edge_at(const Launch& L, uint64_t s) {
  if (L.sample <= 1) return L.edge_lo + s;
  return L.edge_lo + (s / kSampleBlock) * (kSampleBlock * L.sample) + s % kSampleBlock;
}

__device__ __forceinline__ DeferOut defer_out(const Launch& L) {
  DeferOut O;
  O.buf = L.rout;
  O.count = &L.cursor[4];
  O.cap = L.rout_cap;
  O.soft = L.soft;
  O.budget = L.budget;
  return O;
}

template <bool PG, bool RS = false>
__global__ __launch_bounds__(64, PG ? 4 : PM_RESIDENT_WAVES) void k_match_edges(Launch L) {
  __shared__ WaveState ws;
  const KGraph& g = L.g;
  const uint32_t lane = threadIdx.x;
  uint32_t* arena = L.arena + (size_t)blockIdx.x * L.arena_stride;
  Counters ct;
  // per-pattern totals accumulate in LDS (lane 0) and are flushed when the
  // wave moves on to another pattern's chunks
  if (lane == 0) {
    ws.pat = 0xFFFFFFFFu;
    ws.acc_count = ws.acc_jobs = 0;
    ws.lcount = 0;
    ws.rs_depth = -1;
  }
  __syncthreads();
  const DeferOut O = defer_out(L);
  const uint64_t units = (uint64_t)L.npat * L.nchunk;
  for (;;) {
    // out of core: stop taking new jobs once the piece buffer is half full;
    // a later round continues the scan from the cursor
    if (PG && uni64(*(volatile unsigned long long*)O.count) > O.soft) break;
    unsigned long long u = 0;
    if (lane == 0) u = atomicAdd(&L.cursor[0], 1ull);
    u = uni64(__shfl(u, 0));
    if (u >= units) break;
    const uint32_t p = uni((uint32_t)(u / L.nchunk));
    if (lane == 0 && p != ws.pat) {
      if (ws.pat != 0xFFFFFFFFu && (ws.acc_count | ws.acc_jobs)) {
        atomicAdd(&L.pcount[2 * ws.pat], ws.acc_count);
        atomicAdd(&L.pcount[2 * ws.pat + 1], ws.acc_jobs);
      }
      ws.acc_count = ws.acc_jobs = 0;
      ws.pat = p;
    }
    __syncthreads();
    const KPlan& P = L.plans[p];
    const uint64_t s0 = (u % L.nchunk) * L.chunk;
    const uint64_t s1 = min(s0 + L.chunk, L.npos);
    const bool lorder = L.root_order != nullptr;
    // rlo: source vertex (CSR order) or root index (learned order) of s0
    uint32_t rlo = lorder ? uni(gallop(L.root_start, (uint32_t)g.n, s0, 0))
                          : uni(source_full(g, edge_at(L, s0)));
    for (uint64_t sb = s0; sb < s1; sb += kWave) {
      bool ok = sb + lane < s1;
      uint64_t e = 0;
      uint32_t r = rlo, c = 0, ri = rlo;
      if (ok && lorder) {
        ri = gallop(L.root_start, (uint32_t)g.n, sb + lane, rlo);
        r = L.root_order[ri];
        e = g.rowptr[r] + (sb + lane - L.root_start[ri]);
      } else if (ok) {
        e = edge_at(L, sb + lane);
      }
      if (ok) {
        if (!lorder) r = source_of(g, e, rlo);
        c = g.jcol[e];
        if (!P.labeled && P.rc_order && !(r > c)) ok = false;
        if (ok && P.labeled) ok = g.lab[r] == P.lab0 && g.lab[c] == P.lab1;
        if (ok) ok = degree(g, r) >= P.deg0 && degree(g, c) >= P.deg1 && g.maxnbr[c] >= P.deg2;
      }
      const uint32_t lastl = (uint32_t)min((uint64_t)kWave, s1 - sb) - 1u;
      uint64_t jm = __ballot(ok);
      ws.jr[lane] = r;
      ws.jc[lane] = c;
      ws.jedge[lane] = e;
      __syncthreads();
      rlo = uni((uint32_t)__shfl((int)(lorder ? ri : r), (int)lastl));
      ct.alg += 24ull * min((uint64_t)kWave, s1 - sb);
      jm = uni64(jm);
      while (jm) {
        const int i = __builtin_ctzll(jm);
        jm &= jm - 1;
        const uint32_t jr = uni(ws.jr[i]), jc = uni(ws.jc[i]);
        const uint64_t ej = uni64(ws.jedge[i]);
        if (lane == 0) ws.alg0 = ct.alg;
        const unsigned long long cnt =
            run_job<PG, RS>(g, P, ws, arena, L.cap, lane, jr, jc, ej, 0, kInf, false,
                            L.tasks != nullptr, L.tasks, &L.cursor[2], L.task_cap, O, ct);
        resplit_flush<PG, RS>(O, ws, lane);
        if (L.root_work && lane == 0)
          atomicAdd(&L.root_work[jr], (unsigned long long)(ct.alg - ws.alg0));
        if (lane == 0) {
          ws.acc_count += cnt;
          ws.acc_jobs++;
          if (L.per_edge && cnt) atomicAdd(&L.per_edge[ej - L.edge_lo], cnt);
        }
      }
      __syncthreads();
    }
  }
  if (lane == 0) {
    if (ws.pat != 0xFFFFFFFFu && (ws.acc_count | ws.acc_jobs)) {
      atomicAdd(&L.pcount[2 * ws.pat], ws.acc_count);
      atomicAdd(&L.pcount[2 * ws.pat + 1], ws.acc_jobs);
    }
    atomicAdd(&L.out[2], (unsigned long long)ct.alg);
    atomicAdd(&L.out[3], (unsigned long long)ct.nops);
  }
}

// Range tasks (phase 2).
template <bool PG, bool RS = false>
__global__ __launch_bounds__(64, PG ? 4 : PM_RESIDENT_WAVES) void k_match_tasks(Launch L) {
  __shared__ WaveState ws;
  const KGraph& g = L.g;
  const uint32_t lane = threadIdx.x;
  uint32_t* arena = L.arena + (size_t)blockIdx.x * L.arena_stride;
  Counters ct;
  const DeferOut O = defer_out(L);
  if (lane == 0) {
    ws.lcount = 0;
    ws.rs_depth = -1;
  }
  __syncthreads();
  const uint64_t ntask = L.tasks ? min((uint64_t)L.cursor[2], L.task_cap) : 0;
  for (;;) {
    // out of core: likewise for tasks (claims then stay below the count)
    if (PG && uni64(*(volatile unsigned long long*)O.count) > O.soft) break;
    unsigned long long t = 0;
    if (lane == 0) t = atomicAdd(&L.cursor[1], 1ull);
    t = uni64(__shfl(t, 0));
    if (t >= ntask) break;
    const uint64_t e = uni64(L.tasks[t].edge);
    const uint32_t tlo = uni(L.tasks[t].lo);
    uint32_t thi = uni(L.tasks[t].hi);
    if (thi == 0) thi = uni(L.tasks[t + 1].lo);  // same job: slots are contiguous
    const uint32_t r = uni(L.tasks[t].r);
    const uint32_t c = uni(L.tasks[t].c);
    const uint32_t p = uni(L.tasks[t].pat);
    ws.pat = p;
    if (lane == 0) ws.alg0 = ct.alg;
    const unsigned long long cnt =
        run_job<PG, RS>(g, L.plans[p], ws, arena, L.cap, lane, r, c, e, tlo, thi, false, false,
                        nullptr, nullptr, 0, O, ct);
    resplit_flush<PG, RS>(O, ws, lane);
    if (L.root_work && lane == 0)
      atomicAdd(&L.root_work[r], (unsigned long long)(ct.alg - ws.alg0));
    if (lane == 0 && cnt) atomicAdd(&L.pcount[2 * p], cnt);
    if (L.per_edge && lane == 0 && cnt) atomicAdd(&L.per_edge[e - L.edge_lo], cnt);
  }
  if (lane == 0) {
    atomicAdd(&L.out[4], (unsigned long long)ct.alg);
    atomicAdd(&L.out[5], (unsigned long long)ct.nops);
  }
}

// Resume round: the admitted deferred pieces (a depth-0 piece may split).
template <bool PG, bool RS = false>
__global__ __launch_bounds__(64, PG ? 4 : PM_RESIDENT_WAVES) void k_resume(Launch L) {
  __shared__ WaveState ws;
  const KGraph& g = L.g;
  const uint32_t lane = threadIdx.x;
  uint32_t* arena = L.arena + (size_t)blockIdx.x * L.arena_stride;
  Counters ct;
  const DeferOut O = defer_out(L);
  if (lane == 0) {
    ws.lcount = 0;
    ws.rs_depth = -1;
  }
  __syncthreads();
  const uint64_t nrin = L.nrin_dev ? uni64(*L.nrin_dev) : L.nrin;
  // Pieces are claimed kResumeClaim at a time.  Consecutive one-candidate
  // pieces at the fused level with the same search path (the ones a single
  // batch of candidates deferred) run as one group: the path is rebuilt
  // once and the fused step walks all their candidates together.
  for (;;) {
    const uint32_t claim = L.claim ? L.claim : kResumeClaim;
    unsigned long long t0 = 0;
    if (lane == 0) t0 = atomicAdd(&L.cursor[3], (unsigned long long)claim);
    t0 = uni64(__shfl(t0, 0));
    if (t0 >= nrin) break;
    const uint64_t t1 = min(t0 + claim, nrin);
    for (uint64_t t = t0; t < t1;) {
      const Resume& R0 = L.rin[t];
      const uint64_t e = uni64(R0.edge);
      const uint32_t r = uni(R0.r), c = uni(R0.c);
      const uint32_t rlo = uni(R0.lo), rhi = uni(R0.hi);
      const int depth = (int)uni((uint32_t)R0.depth);
      const uint32_t p = uni(R0.pat);
      const KPlan& P = L.plans[p];
      uint64_t t2 = t + 1;
      if (depth >= 1 && depth < kMaxDepth && depth == P.k - 3 &&
          uni(R0.end) == uni(R0.idx[depth]) + 1) {
        while (t2 < t1 && t2 - t < (uint64_t)kWave) {
          const Resume& R = L.rin[t2];
          bool same = uni64(R.edge) == e && uni(R.pat) == p && uni((uint32_t)R.depth) == (uint32_t)depth &&
                      uni(R.lo) == rlo && uni(R.hi) == rhi && uni(R.end) == uni(R.idx[depth]) + 1;
          for (int d = 1; d < depth && same; d++) same = uni(R.idx[d]) == uni(R0.idx[d]);
          if (!same) break;
          t2++;
        }
      }
      ws.pat = p;
      ws.rdepth = depth;
      for (int d = 1; d <= depth && d < kMaxDepth; d++) ws.ridx[d] = uni(R0.idx[d]);
      for (int d = 0; d <= depth && d < kMaxDepth; d++) ws.phi[d] = uni(R0.phi[d]);
      ws.rend = uni(R0.end);
      const uint32_t ng = (uint32_t)(t2 - t);
      if (ng > 1 && lane < ng) ws.lidx[lane] = L.rin[t + lane].idx[depth];
      if (lane == 0) {
        ws.lcount = ng > 1 ? ng : 0;
        ws.alg0 = ct.alg;
      }
      __syncthreads();
      const unsigned long long cnt =
          run_job<PG, RS>(g, P, ws, arena, L.cap, lane, r, c, e, rlo, rhi, true,
                          depth == 0 && L.tasks != nullptr, L.tasks, &L.cursor[2], L.task_cap, O,
                          ct);
      resplit_flush<PG, RS>(O, ws, lane);
      if (lane == 0) ws.lcount = 0;
      __syncthreads();
      if (L.root_work && lane == 0)
        atomicAdd(&L.root_work[r], (unsigned long long)(ct.alg - ws.alg0));
      if (lane == 0 && cnt) atomicAdd(&L.pcount[2 * p], cnt);
      if (L.per_edge && lane == 0 && cnt) atomicAdd(&L.per_edge[e - L.edge_lo], cnt);
      t = t2;
    }
  }
  if (lane == 0) {
    atomicAdd(&L.out[2], (unsigned long long)ct.alg);
    atomicAdd(&L.out[3], (unsigned long long)ct.nops);
  }
}

// F(v) = sum over x in N(v) with lab[x] == want of prod_c child_c[x].
// Edge-balanced: each wave walks 512 consecutive CSR edges per claim, 64 at
// a time; lanes of the same row are summed with a segmented wave scan and
// the last lane of each row segment adds into out[row] (zeroed before), so
// a hub row is spread over many waves instead of serialising one.  Out of
// core the rows are read from mapped host memory (streaming, coalesced).
struct FoldArgs {
  KGraph g;
  uint32_t want;
  int nchild;
  const unsigned long long* child[kMaxK];
  unsigned long long* out;
  unsigned long long* bytes;  // algorithmic bytes accumulator
};

// edges per wave of a fold pass (build-time tunable)
#ifndef PM_FOLD_CHUNK
#define PM_FOLD_CHUNK 1024
#endif
constexpr uint32_t kFoldChunk = PM_FOLD_CHUNK;

__global__ __launch_bounds__(64) void k_fold(FoldArgs A) {
  const KGraph& g = A.g;
  const uint32_t lane = threadIdx.x;
  unsigned long long hits = 0, edges = 0;
  for (uint64_t c0 = (uint64_t)blockIdx.x * kFoldChunk; c0 < g.m;
       c0 += (uint64_t)gridDim.x * kFoldChunk) {
    const uint64_t c1 = min(c0 + kFoldChunk, g.m);
    uint32_t r = uni(source_full(g, c0));
    for (uint64_t eb = c0; eb < c1; eb += kWave) {
      const uint64_t e = eb + lane;
      uint32_t row = 0xFFFFFFFFu;
      unsigned long long val = 0;
      if (e < c1) {
        row = source_of(g, e, r);
        const uint32_t x = g.jcol[e];
        if (g.lab[x] == A.want) {
          val = 1;
          for (int c = 0; c < A.nchild; c++) val *= A.child[c][x];
          hits++;
        }
      }
      // rows are non-decreasing across lanes: segmented inclusive scan
      for (int o = 1; o < kWave; o <<= 1) {
        const unsigned long long vv = __shfl_up(val, o);
        const uint32_t rr = (uint32_t)__shfl_up((int)row, o);
        if ((int)lane >= o && rr == row) val += vv;
      }
      const uint32_t next = (uint32_t)__shfl_down((int)row, 1);
      if (row != 0xFFFFFFFFu && (lane == kWave - 1 || next != row) && val)
        atomicAdd(&A.out[row], val);
      const uint32_t lastl = (uint32_t)min((uint64_t)kWave, c1 - eb) - 1u;
      r = uni((uint32_t)__shfl((int)row, (int)lastl));
    }
    edges += c1 - c0;
  }
  // colidx 4 B + label 1 B per edge, 8 B per child value gathered
  hits = wave_sum(hits);
  if (lane == 0) atomicAdd(A.bytes, 5ull * edges + 8ull * A.nchild * hits);
}

// A leaf of a pendant tree over the label-grouped rows: F(v) is the size of
// the label-`want` part of N(v), read from the row's label offsets.
__global__ void k_fold_leaf(FoldArgs A) {
  const KGraph& g = A.g;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < g.n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t* o = g.loff + v * g.nl + A.want;
    A.out[v] = o[1] - o[0];
  }
  // two label offsets per vertex
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(A.bytes, 8ull * g.n);
}

// out[v] = prod_c child_c[v]
__global__ void k_product(uint64_t n, FoldArgs A) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long p = 1;
    for (int c = 0; c < A.nchild; c++) p *= A.child[c][v];
    A.out[v] = p;
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// One round's package loads in one launch: wave w copies job w's rows from
// pinned host memory (read by the GPU over the link) into its cache slots.
struct CopyJob {
  uint64_t src;  // word offset in the host rows
  uint64_t dst;  // word offset in the cache
  uint64_t words;
};

__global__ __launch_bounds__(256) void k_copy_packages(const uint32_t* __restrict__ host,
                                                       uint32_t* __restrict__ cache,
                                                       const CopyJob* __restrict__ jobs,
                                                       uint32_t njobs) {
  const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  for (uint32_t j = blockIdx.x * (blockDim.x / kWave) + wave; j < njobs;
       j += gridDim.x * (blockDim.x / kWave)) {
    const CopyJob c = jobs[j];
    const uint32_t* src = host + c.src;
    uint32_t* dst = cache + c.dst;
    uint64_t i = lane;
    for (; i + 3 * kWave < c.words; i += 4 * kWave) {
      const uint32_t a = src[i], b = src[i + kWave], d = src[i + 2 * kWave], e = src[i + 3 * kWave];
      dst[i] = a;
      dst[i + kWave] = b;
      dst[i + 2 * kWave] = d;
      dst[i + 3 * kWave] = e;
    }
    for (; i < c.words; i += kWave) dst[i] = src[i];
  }
}

// Pending pieces stay in HBM between rounds.  Needs of a piece: the
// packages of its path rows phi[0..depth] (rows read directly from host
// memory, of <= direct_max entries, need none) and the package that
// stopped it.
constexpr int kNeed = kMaxDepth + 2;  // count + packages

struct NeedArgs {
  const uint32_t* pkg_of;
  const uint64_t* rowptr;
  uint32_t direct_max;
};

__device__ __forceinline__ int piece_needs(const Resume& d, const NeedArgs& A, uint32_t* q) {
  int c = 0;
  for (int j = 0; j <= d.depth && j < kMaxDepth; j++) {
    const uint32_t v = d.phi[j];
    if (A.rowptr[v + 1] - A.rowptr[v] > A.direct_max) q[c++] = A.pkg_of[v];
  }
  q[c++] = d.blk;
  return c;
}

// needs of pieces [0, n) for the host's admission (row i: count, packages)
__global__ void k_piece_needs(const Resume* __restrict__ P, uint64_t n, NeedArgs A,
                              uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t q[kNeed];
    const int c = piece_needs(P[i], A, q);
    out[i * kNeed] = (uint32_t)c;
    for (int j = 0; j < c; j++) out[i * kNeed + 1 + j] = q[j];
  }
}

// Split the pending pieces: those whose needs are all resident go to `run`
// (resumed this round), the others to `keep`.  Counts in cnt[0] (run) and
// cnt[1] (keep); one atomic per wave for each.  Pieces keep their relative
// order within a wave only; the rounds' progress does not depend on it.
__global__ __launch_bounds__(256) void k_piece_select(const Resume* __restrict__ P, uint64_t n,
                                                      NeedArgs A,
                                                      const int32_t* __restrict__ pkg_slot,
                                                      Resume* __restrict__ run,
                                                      Resume* __restrict__ keep,
                                                      unsigned long long* cnt) {
  const uint32_t lane = threadIdx.x % kWave;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x - lane); b < n; b += stride) {
    const uint64_t i = b + lane;
    bool live = i < n, here = false;
    Resume d;
    if (live) {
      d = P[i];
      uint32_t q[kNeed];
      const int c = piece_needs(d, A, q);
      here = true;
      for (int j = 0; j < c; j++) here = here && pkg_slot[q[j]] >= 0;
    }
    const uint64_t mr = __ballot(live && here), mk = __ballot(live && !here);
    unsigned long long br = 0, bk = 0;
    if (lane == 0) {
      if (mr) br = atomicAdd(&cnt[0], (unsigned long long)__popcll(mr));
      if (mk) bk = atomicAdd(&cnt[1], (unsigned long long)__popcll(mk));
    }
    br = __shfl(br, 0);
    bk = __shfl(bk, 0);
    if (live && here) run[br + __popcll(mr & lt)] = d;
    if (live && !here) keep[bk + __popcll(mk & lt)] = d;
  }
}

#define HIP_OK(call)                                                \
  do {                                                              \
    hipError_t _e = (call);                                         \
    if (_e != hipSuccess) {                                         \
      set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return PM_ERR_DEVICE;                                         \
    }                                                               \
  } while (0)

// Per-process work areas, grown on demand and reused across calls.
struct Workspace {
  int device = -1;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;  // host->HBM package copies
  void* arena = nullptr;
  size_t arena_bytes = 0;
  // One block for everything a launch reports, so that it is cleared by
  // one fill and read back by one copy: cursor (8 words), out (8 words),
  // pcount (count, jobs per pattern).
  void* resblk = nullptr;
  size_t resblk_bytes = 0;
  unsigned long long* cursor = nullptr;  // views into resblk
  unsigned long long* out = nullptr;
  unsigned long long* pcount = nullptr;
  unsigned long long* h_res = nullptr;   // pinned copy of resblk
  size_t h_res_cap = 0;                  // words
  void* plans = nullptr;                 // KPlan per pattern of the launch
  size_t plans_bytes = 0;
  Task* tasks = nullptr;
  uint64_t task_cap = 0;
  void* tables = nullptr;  // fold / weight tables
  size_t table_bytes = 0;
  void* resume[2] = {nullptr, nullptr};
  size_t resume_bytes[2] = {0, 0};
  void* pend[2] = {nullptr, nullptr};  // pending pieces between rounds
  size_t pend_bytes[2] = {0, 0};
  void* d_needs = nullptr;             // needs of the admission window
  size_t d_needs_bytes = 0;
  uint32_t* h_needs = nullptr;         // pinned copy of the same
  size_t h_needs_cap = 0;
  unsigned long long* h_stat = nullptr;  // pinned round status
  uint8_t* h_used = nullptr;             // pinned per-slot use of a round
  size_t h_used_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evf = nullptr, eve = nullptr, evc = nullptr;
  // captured launch sequence of the last graph-mode run and the bytes of
  // every argument it was captured with
  hipGraphExec_t gexec = nullptr;
  std::string gsig;
  int wall_khz = 0;  // rate of wall_clock64()
  int cus = 0;
  int resident = 0;
};
std::mutex g_ws_mu;
Workspace g_ws;

int ensure_workspace(int device) {
  if (g_ws.device == device) return PM_OK;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&g_ws.stream, hipStreamNonBlocking));
  HIP_OK(hipStreamCreateWithFlags(&g_ws.copy_stream, hipStreamNonBlocking));
  HIP_OK(hipEventCreate(&g_ws.ev0));
  HIP_OK(hipEventCreate(&g_ws.ev1));
  HIP_OK(hipEventCreate(&g_ws.evf));
  HIP_OK(hipEventCreate(&g_ws.eve));
  HIP_OK(hipEventCreate(&g_ws.evc));
  HIP_OK(hipDeviceGetAttribute(&g_ws.cus, hipDeviceAttributeMultiprocessorCount, device));
  HIP_OK(hipDeviceGetAttribute(&g_ws.wall_khz, hipDeviceAttributeWallClockRate, device));
  g_ws.task_cap = 1ull << PM_TASK_CAP_LOG2;  // range tasks (32 B each)
  HIP_OK(hipMalloc(&g_ws.tasks, sizeof(Task) * g_ws.task_cap));
  g_ws.device = device;
  return PM_OK;
}

// the result block for npat patterns (device and pinned host copy)
int grow_results(int npat) {
  const size_t words = 16 + 2 * (size_t)std::max(npat, 1);
  const size_t need = sizeof(unsigned long long) * words;
  if (need > g_ws.resblk_bytes) {
    if (g_ws.resblk) HIP_OK(hipFree(g_ws.resblk));
    g_ws.resblk = nullptr;
    g_ws.resblk_bytes = 0;
    HIP_OK(hipMalloc(&g_ws.resblk, need));
    g_ws.resblk_bytes = need;
  }
  if (words > g_ws.h_res_cap) {
    if (g_ws.h_res) HIP_OK(hipHostFree(g_ws.h_res));
    g_ws.h_res = nullptr;
    g_ws.h_res_cap = 0;
    HIP_OK(hipHostMalloc((void**)&g_ws.h_res, need, hipHostMallocDefault));
    g_ws.h_res_cap = words;
  }
  g_ws.cursor = (unsigned long long*)g_ws.resblk;
  g_ws.out = g_ws.cursor + 8;
  g_ws.pcount = g_ws.cursor + 16;
  return PM_OK;
}

int grow_pinned_u32(uint32_t** p, size_t* have, size_t need) {
  if (need <= *have) return PM_OK;
  if (*p) HIP_OK(hipHostFree(*p));
  *p = nullptr;
  *have = 0;
  HIP_OK(hipHostMalloc((void**)p, sizeof(uint32_t) * need, hipHostMallocDefault));
  *have = need;
  return PM_OK;
}

// grow a device buffer to `need` bytes, keeping its first `keep` bytes
int grow_keep(void** p, size_t* have, size_t need, size_t keep, hipStream_t st) {
  if (need <= *have) return PM_OK;
  const size_t n = std::max(need, *have * 2);
  void* q = nullptr;
  HIP_OK(hipMalloc(&q, n));
  if (keep) HIP_OK(hipMemcpyAsync(q, *p, keep, hipMemcpyDeviceToDevice, st));
  HIP_OK(hipStreamSynchronize(st));
  if (*p) HIP_OK(hipFree(*p));
  *p = q;
  *have = n;
  return PM_OK;
}

int grow(void** p, size_t* have, size_t need) {
  if (need <= *have) return PM_OK;
  if (*p) HIP_OK(hipFree(*p));
  *p = nullptr;
  *have = 0;
  HIP_OK(hipMalloc(p, need));
  *have = need;
  return PM_OK;
}

void build_kplan(const Plan& p, bool segs, KPlan* kp, int* nregions) {
  std::memset(kp, 0, sizeof(*kp));
  const int kc = p.kc;
  kp->k = kc;
  kp->labeled = p.labeled;
  kp->rc_order = p.rc_order;
  kp->segs = (p.labeled && segs) ? 1 : 0;
  kp->induced = p.induced ? 1 : 0;
  kp->deg0 = p.deg[0];
  kp->deg1 = p.deg[1];
  kp->deg2 = p.deg[2];
  kp->lab0 = p.vlabel[0];
  kp->lab1 = p.vlabel[1];
  for (int u = 0; u < kc; u++) {
    kp->lab[u] = p.clabel[u];
    kp->B[u] = p.cB[u];
    kp->X[u] = p.cX[u];
  }
  for (int s = 0; s < kMaxSlots; s++) kp->region[s] = -1;
  int nops = 0, nreg = 0;
  for (int d = 0; d <= kc - 2; d++) {
    kp->op_begin[d] = (int8_t)nops;
    if (d < kc - 2) {
      for (const SlotOp& op : p.ops[d]) {
        KOp& ko = kp->ops[nops++];
        ko.kind = op.kind;
        ko.slot = op.slot;
        ko.parent = op.parent;
        ko.ncons = (int8_t)op.ncons;
        ko.labmask = op.labmask;
        for (int c = 0; c < op.ncons; c++) ko.cons_b[c] = op.cons_b[c];
        ko.seg = (kp->segs && __builtin_popcount(op.labmask) == 1)
                     ? (int8_t)__builtin_ctz(op.labmask)
                     : (int8_t)-1;
        if (op.kind == 1) kp->region[op.slot] = (int8_t)nreg++;
      }
    }
    kp->op_end[d] = (int8_t)nops;
  }
  kp->nops = nops;
  // a slot holds only label-`label` vertices when its label filter is that
  // one label (INTERSECT filters; VIEW reads that label's part of the row,
  // which needs the label-grouped rows)
  auto exact_slot = [&](int s, int label, int upto) {
    if (!p.labeled) return false;
    for (int dd = 0; dd <= upto; dd++)
      for (const SlotOp& op : p.ops[dd])
        if (op.slot == s && op.labmask == (1u << label) && (op.kind == 1 || kp->segs))
          return true;
    return false;
  };
  for (int d = 1; d <= kc - 3; d++) {
    const int s = p.slot_of[d][d + 1];
    kp->cand_slot[d] = (int8_t)s;
    kp->cand_check[d] = exact_slot(s, p.clabel[d + 1], d) ? 0 : 1;
  }
  if (kc >= 3) {
    const int dl = kc - 2;
    if (p.ops[dl].empty()) {
      kp->last_kind = 0;
      kp->last_src = (int8_t)p.slot_of[dl][kc - 1];
      kp->last_check = exact_slot(kp->last_src, p.clabel[kc - 1], dl - 1) ? 0 : 1;
    } else {
      const SlotOp& op = p.ops[dl][0];
      kp->last_kind = op.kind == 0 ? 1 : 2;
      kp->last_src = op.parent;
      kp->last_check = kp->segs ? 0 : 1;
    }
    // Labeled tail whose two vertices are interchangeable: u1's candidates
    // are the set P that u2 intersects with N(x), both label-exact with one
    // label.  Then the ordered pairs (x,y) are the adjacent pairs inside P,
    // and each unordered pair can be visited once (y > x) and credited in
    // both roles.
    if (p.labeled && !p.induced && kp->segs && kc >= 4 && kp->last_kind == 2 &&
        kp->last_src == kp->cand_slot[kc - 3] && p.clabel[kc - 2] == p.clabel[kc - 1] &&
        kp->cand_check[kc - 3] == 0)
      kp->pair = 1;
  }
  *nregions = nreg;
}

}  // namespace

// ---------------------------------------------------------------------------
// device graph
// ---------------------------------------------------------------------------

// (Re)build the label-grouped copy of the rows when the graph has labels and
// its rows are resident.
int upload_label_grouped(pm_graph* g) {
  DeviceGraph& d = g->d;
  if (d.lcol) HIP_OK(hipFree(d.lcol));
  if (d.loff) HIP_OK(hipFree(d.loff));
  d.lcol = nullptr;
  d.loff = nullptr;
  d.nl = 0;
  const HostGraph& h = g->h;
  if (h.labels.empty() || h.n == 0 || !d.colidx) return PM_OK;
  std::vector<uint32_t> lcol, loff;
  uint32_t nl = 0;
  build_label_grouped(h, &lcol, &loff, &nl);
  HIP_OK(hipMalloc(&d.lcol, sizeof(uint32_t) * std::max<uint64_t>(h.m, 1)));
  HIP_OK(hipMalloc(&d.loff, sizeof(uint32_t) * loff.size()));
  if (h.m) HIP_OK(hipMemcpy(d.lcol, lcol.data(), sizeof(uint32_t) * h.m, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(d.loff, loff.data(), sizeof(uint32_t) * loff.size(), hipMemcpyHostToDevice));
  d.nl = nl;
  return PM_OK;
}

// Per-vertex metadata always goes to HBM; the rows (colidx) only when
// with_rows (out of core they stay in pinned host memory).
int device_upload_rows(pm_graph* g, bool with_rows) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  int rc = ensure_workspace(dev);
  if (rc) return rc;
  const HostGraph& h = g->h;
  DeviceGraph& d = g->d;
  if (!g->on_device) {
    std::vector<uint8_t> lab8(h.n, 0);
    for (uint64_t v = 0; v < h.n && !h.labels.empty(); v++) lab8[v] = (uint8_t)h.labels[v];
    std::vector<uint32_t> maxnbr;
    compute_max_neighbour_degree(h, &maxnbr);
    HIP_OK(hipMalloc(&d.rowptr, sizeof(uint64_t) * (h.n + 1)));
    HIP_OK(hipMalloc(&d.label8, std::max<uint64_t>(h.n, 1)));
    HIP_OK(hipMalloc(&d.maxnbr, sizeof(uint32_t) * std::max<uint64_t>(h.n, 1)));
    HIP_OK(hipMemcpy(d.rowptr, h.rowptr.data(), sizeof(uint64_t) * (h.n + 1), hipMemcpyHostToDevice));
    if (h.n) {
      HIP_OK(hipMemcpy(d.label8, lab8.data(), h.n, hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d.maxnbr, maxnbr.data(), sizeof(uint32_t) * h.n, hipMemcpyHostToDevice));
    }
    // source of every 64th edge: edge -> source lookups start here
    const uint64_t ns = (h.m + 63) / 64 + 1;
    std::vector<uint32_t> src64(ns, 0);
    uint64_t v = 0;
    for (uint64_t i = 0; i * 64 < h.m; i++) {
      while (h.rowptr[v + 1] <= i * 64) v++;
      src64[i] = (uint32_t)v;
    }
    HIP_OK(hipMalloc(&d.src64, sizeof(uint32_t) * ns));
    HIP_OK(hipMemcpy(d.src64, src64.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
    d.bytes = sizeof(uint64_t) * (h.n + 1) + h.n + 4 * h.n + 4 * ns;
    g->on_device = true;
    g->device = dev;
  }
  if (with_rows && !d.colidx) {
    HIP_OK(hipMalloc(&d.colidx, sizeof(uint32_t) * std::max<uint64_t>(h.m, 1)));
    if (h.m)
      HIP_OK(hipMemcpy(d.colidx, h.colidx.data(), sizeof(uint32_t) * h.m, hipMemcpyHostToDevice));
    d.bytes += sizeof(uint32_t) * h.m;
    rc = upload_label_grouped(g);
    if (rc) return rc;
  }
  return PM_OK;
}

int device_upload(pm_graph* g) { return device_upload_rows(g, true); }

int device_update_labels(pm_graph* g) {
  if (!g->on_device) return PM_OK;
  const HostGraph& h = g->h;
  std::vector<uint8_t> lab8(h.n, 0);
  for (uint64_t v = 0; v < h.n && !h.labels.empty(); v++) lab8[v] = (uint8_t)h.labels[v];
  if (h.n) HIP_OK(hipMemcpy(g->d.label8, lab8.data(), h.n, hipMemcpyHostToDevice));
  return upload_label_grouped(g);
}

void device_release_paging(pm_graph* g);

void device_release(pm_graph* g) {
  device_release_paging(g);
  if (!g->on_device) return;
  (void)hipFree(g->d.rowptr);
  if (g->d.colidx) (void)hipFree(g->d.colidx);
  (void)hipFree(g->d.label8);
  (void)hipFree(g->d.maxnbr);
  if (g->d.src64) (void)hipFree(g->d.src64);
  if (g->d.lcol) (void)hipFree(g->d.lcol);
  if (g->d.loff) (void)hipFree(g->d.loff);
  if (g->d_root_order) (void)hipFree(g->d_root_order);
  if (g->d_root_start) (void)hipFree(g->d_root_start);
  g->d_root_order = nullptr;
  g->d_root_start = nullptr;
  g->d = DeviceGraph();
  g->on_device = false;
}

// ---------------------------------------------------------------------------
// out-of-core state of one graph
// ---------------------------------------------------------------------------

struct PagedGraph {
  uint32_t slot_words = 0;
  uint64_t nslots = 0;
  PackageLayout layout;
  uint32_t* host_rows = nullptr;  // pinned host copy of colidx
  uint32_t* host_rows_dev = nullptr;  // its device address
  int32_t* d_pkg_slot = nullptr;
  uint32_t* d_pkg_of = nullptr;
  uint64_t* d_pkg_base = nullptr;
  uint32_t* d_cache = nullptr;
  unsigned int* d_miss_bits = nullptr;
  uint32_t* d_miss_list = nullptr;
  unsigned long long* d_miss_count = nullptr;
  uint32_t miss_cap = 0;
  uint8_t* d_used = nullptr;
  CopyJob* d_jobs = nullptr;       // package copies of one call (<= nslots)
  std::vector<CopyJob> jobs;       // host staging of the same
  std::vector<int32_t> pkg_slot;  // host mirror
  CachePolicy policy;
  // per package: entries of its rows longer than `long_for` (the rows that
  // can miss when shorter ones are read directly); computed on demand
  std::vector<uint64_t> long_words;
  uint32_t long_for = ~0u;
};

std::mutex g_paged_mu;
std::vector<std::pair<const pm_graph*, std::unique_ptr<PagedGraph>>> g_paged;

PagedGraph* find_paged(const pm_graph* g) {
  for (auto& e : g_paged)
    if (e.first == g) return e.second.get();
  return nullptr;
}

void device_release_paging(pm_graph* g) {
  std::lock_guard<std::mutex> lk(g_paged_mu);
  for (size_t i = 0; i < g_paged.size(); i++) {
    if (g_paged[i].first != g) continue;
    PagedGraph* pg = g_paged[i].second.get();
    if (pg->host_rows) (void)hipHostFree(pg->host_rows);
    (void)hipFree(pg->d_pkg_slot);
    (void)hipFree(pg->d_pkg_of);
    (void)hipFree(pg->d_pkg_base);
    (void)hipFree(pg->d_cache);
    (void)hipFree(pg->d_miss_bits);
    (void)hipFree(pg->d_miss_list);
    (void)hipFree(pg->d_miss_count);
    (void)hipFree(pg->d_used);
    (void)hipFree(pg->d_jobs);
    g_paged.erase(g_paged.begin() + i);
    return;
  }
}

// Set up (or reuse) the package cache of `cache_bytes` over slots of
// `package_bytes`.
int ensure_paged(pm_graph* g, uint64_t cache_bytes, uint32_t package_bytes, PagedGraph** out) {
  const uint32_t words = std::max<uint32_t>(package_bytes, 4096) / 4;
  const uint64_t nslots = cache_bytes / ((uint64_t)words * 4);
  std::lock_guard<std::mutex> lk(g_paged_mu);
  PagedGraph* pg = find_paged(g);
  if (pg && pg->slot_words == words && pg->nslots == nslots) {
    *out = pg;
    return PM_OK;
  }
  if (pg) {
    // different geometry: rebuild
    for (size_t i = 0; i < g_paged.size(); i++)
      if (g_paged[i].first == g) {
        PagedGraph* q = g_paged[i].second.get();
        if (q->host_rows) (void)hipHostFree(q->host_rows);
        (void)hipFree(q->d_pkg_slot);
        (void)hipFree(q->d_pkg_of);
        (void)hipFree(q->d_pkg_base);
        (void)hipFree(q->d_cache);
        (void)hipFree(q->d_miss_bits);
        (void)hipFree(q->d_miss_list);
        (void)hipFree(q->d_miss_count);
        (void)hipFree(q->d_used);
        (void)hipFree(q->d_jobs);
        g_paged.erase(g_paged.begin() + i);
        break;
      }
  }
  const HostGraph& h = g->h;
  auto npg = std::make_unique<PagedGraph>();
  pg = npg.get();
  pg->slot_words = words;
  pg->nslots = nslots;
  int rc = build_packages(h, words, &pg->layout);
  if (rc) return rc;
  uint32_t widest = 1;
  for (uint32_t w : pg->layout.nslots) widest = std::max(widest, w);
  if (nslots < (uint64_t)widest + 72) {
    set_error("out-of-core cache too small: it must hold the widest package plus 72 slots");
    return PM_ERR_ARG;
  }
  const uint32_t np = pg->layout.count();
  HIP_OK(hipHostMalloc((void**)&pg->host_rows, sizeof(uint32_t) * std::max<uint64_t>(h.m, 1),
                       hipHostMallocMapped));
  if (h.m) std::memcpy(pg->host_rows, h.colidx.data(), sizeof(uint32_t) * h.m);
  HIP_OK(hipHostGetDevicePointer((void**)&pg->host_rows_dev, pg->host_rows, 0));
  HIP_OK(hipMalloc(&pg->d_pkg_slot, sizeof(int32_t) * std::max(np, 1u)));
  HIP_OK(hipMalloc(&pg->d_pkg_of, sizeof(uint32_t) * std::max<uint64_t>(h.n, 1)));
  HIP_OK(hipMalloc(&pg->d_pkg_base, sizeof(uint64_t) * std::max(np, 1u)));
  HIP_OK(hipMalloc(&pg->d_cache, (size_t)nslots * words * 4));
  pg->miss_cap = std::max(np, 1u);
  HIP_OK(hipMalloc(&pg->d_miss_bits, sizeof(unsigned int) * ((np + 31) / 32 + 1)));
  HIP_OK(hipMalloc(&pg->d_miss_list, sizeof(uint32_t) * pg->miss_cap));
  HIP_OK(hipMalloc(&pg->d_miss_count, sizeof(unsigned long long)));
  HIP_OK(hipMalloc(&pg->d_used, nslots));
  HIP_OK(hipMalloc(&pg->d_jobs, sizeof(CopyJob) * std::max<uint64_t>(nslots, 1)));
  if (h.n)
    HIP_OK(hipMemcpy(pg->d_pkg_of, pg->layout.pkg_of.data(), sizeof(uint32_t) * h.n,
                     hipMemcpyHostToDevice));
  if (np)
    HIP_OK(hipMemcpy(pg->d_pkg_base, pg->layout.edge_base.data(), sizeof(uint64_t) * np,
                     hipMemcpyHostToDevice));
  g_paged.emplace_back(g, std::move(npg));
  *out = pg;
  return PM_OK;
}

// Load `want` (package ids) into victims chosen by the policy; packages in
// `keep` are not evicted.  Copies go on the copy stream; returns bytes.
int load_packages(PagedGraph* pg, const std::vector<uint32_t>& want, std::vector<uint8_t>& pinned,
                  hipStream_t cs, uint64_t* bytes) {
  *bytes = 0;
  const PackageLayout& L = pg->layout;
  for (uint32_t p : want) {
    if (pg->pkg_slot[p] >= 0) continue;
    const uint32_t width = L.nslots[p];
    const int64_t s = pg->policy.victim(width, pinned);
    if (s < 0) break;  // everything left is pinned: the rest waits a round
    for (uint32_t w = 0; w < width; w++) {
      const int32_t old = pg->policy.owner((uint32_t)s + w);
      if (old >= 0) {
        // evict the whole package that owned the slot
        const int32_t os = pg->pkg_slot[old];
        for (uint32_t k = 0; k < L.nslots[old]; k++) pg->policy.release((uint32_t)os + k);
        pg->pkg_slot[old] = -1;
      }
    }
    pg->policy.assign((uint32_t)s, width, (int32_t)p);
    for (uint32_t w = 0; w < width; w++) pinned[s + w] = 1;
    pg->pkg_slot[p] = (int32_t)s;
    if (L.words[p]) pg->jobs.push_back({L.edge_base[p], (uint64_t)s * pg->slot_words, L.words[p]});
    *bytes += L.words[p] * 4;
  }
  return PM_OK;
}

// Issue the package copies gathered by load_packages (one launch on cs).
int flush_packages(PagedGraph* pg, hipStream_t cs, bool wait = true) {
  if (pg->jobs.empty()) return PM_OK;
  HIP_OK(hipMemcpyAsync(pg->d_jobs, pg->jobs.data(), sizeof(CopyJob) * pg->jobs.size(),
                        hipMemcpyHostToDevice, cs));
  const uint32_t nj = (uint32_t)pg->jobs.size();
  hipLaunchKernelGGL(k_copy_packages, dim3(std::min<uint32_t>((nj + 3) / 4, 4096)), dim3(256), 0, cs,
                     pg->host_rows_dev, pg->d_cache, pg->d_jobs, nj);
  HIP_OK(hipGetLastError());
  // the staging vector is reused after the stream has drained (without
  // `wait`, the caller waits and clears it)
  if (!wait) return PM_OK;
  HIP_OK(hipStreamSynchronize(cs));
  pg->jobs.clear();
  return PM_OK;
}

// Learned job order for model m: roots by predicted work (output 0),
// heaviest first (bucketed, ties by id), and the first job position of each.
int learned_order(pm_graph* g, const pm_model* m, const std::vector<float>& pred) {
  if (g->d_root_order && g->order_version == m->version) return PM_OK;
  const HostGraph& h = g->h;
  const uint64_t n = h.n;
  float lo = 0, hi = 0;
  for (uint64_t v = 0; v < n; v++) {
    const float y = pred[2 * v];
    if (v == 0 || y < lo) lo = y;
    if (v == 0 || y > hi) hi = y;
  }
  constexpr int kBuckets = 4096;
  const float scale = hi > lo ? (kBuckets - 1) / (hi - lo) : 0.0f;
  std::vector<uint32_t> bucket(n), cnt(kBuckets + 1, 0);
  for (uint64_t v = 0; v < n; v++) {
    const int b = (kBuckets - 1) - std::min(kBuckets - 1, std::max(0, (int)((pred[2 * v] - lo) * scale)));
    bucket[v] = (uint32_t)b;
    cnt[b + 1]++;
  }
  for (int b = 0; b < kBuckets; b++) cnt[b + 1] += cnt[b];
  std::vector<uint32_t> order(n);
  for (uint64_t v = 0; v < n; v++) order[cnt[bucket[v]]++] = (uint32_t)v;
  std::vector<uint64_t> start(n + 1, 0);
  for (uint64_t i = 0; i < n; i++)
    start[i + 1] = start[i] + (h.rowptr[order[i] + 1] - h.rowptr[order[i]]);
  if (!g->d_root_order) {
    HIP_OK(hipMalloc(&g->d_root_order, sizeof(uint32_t) * std::max<uint64_t>(n, 1)));
    HIP_OK(hipMalloc(&g->d_root_start, sizeof(uint64_t) * (n + 1)));
  }
  if (n) HIP_OK(hipMemcpy(g->d_root_order, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(g->d_root_start, start.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice));
  g->order_version = m->version;
  return PM_OK;
}

// Out-of-core rounds after the first launch.  Deferred searches wait in
// HBM; each round admits them oldest first while the union of what they
// need -- the packages of their path rows and of the package that stopped
// them -- fits the cache, loads exactly that set (pinned against eviction)
// and resumes every piece whose needs are resident.  Every admitted search
// then passes its stopping point, so the rounds terminate; nothing is
// dropped, so counts are exact.
struct RoundStats {
  uint64_t misses = 0, deferred = 0, link_bytes = 0;
  double link_ms = 0, kernel_ms = 0;
  int rounds = 1;
};

int run_rounds(PagedGraph* pg, Launch& L, hipStream_t st, int waves, uint64_t span, uint64_t rcap,
               const pm_opts* opts, const std::vector<uint32_t>& by_score,
               std::vector<uint64_t>& use_acc, RoundStats& rs) {
  const uint32_t np = pg->layout.count();
  int rc = PM_OK;
  const PackageLayout& PL = pg->layout;
  // Pending pieces stay in HBM (pend[pc][0..npend)); each round the
  // host reads the needs of the first kWindow of them, admits those
  // oldest first while the union of their needs fits the cache, loads
  // that set, and the device moves every pending piece whose needs are
  // resident to the resume buffer.
  constexpr uint64_t kWindow = 1u << 16;
  constexpr uint64_t kPendMax = 1ull << 26;  // pieces held between rounds
  uint64_t npend = 0;
  int pc = 0;
  rc = grow_pinned_u32(&g_ws.h_needs, &g_ws.h_needs_cap, kWindow * kNeed);
  if (rc) return rc;
  rc = grow(&g_ws.d_needs, &g_ws.d_needs_bytes, sizeof(uint32_t) * kWindow * kNeed);
  if (rc) return rc;
  // Round status (cursors, miss count, selection counts, slot use) comes
  // back in one batch of copies into pinned memory after each round's
  // kernels: [0..7] cursors, [8] misses.
  if (!g_ws.h_stat) HIP_OK(hipHostMalloc((void**)&g_ws.h_stat, sizeof(unsigned long long) * 16,
                                         hipHostMallocDefault));
  if (g_ws.h_used_cap < pg->nslots) {
    if (g_ws.h_used) HIP_OK(hipHostFree(g_ws.h_used));
    g_ws.h_used = nullptr;
    g_ws.h_used_cap = 0;
    HIP_OK(hipHostMalloc((void**)&g_ws.h_used, pg->nslots, hipHostMallocDefault));
    g_ws.h_used_cap = pg->nslots;
  }
  auto fetch_status = [&]() -> int {
    HIP_OK(hipMemcpyAsync(g_ws.h_stat, g_ws.cursor, sizeof(unsigned long long) * 8,
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(g_ws.h_stat + 8, pg->d_miss_count, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(g_ws.h_used, pg->d_used, pg->nslots, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return PM_OK;
  };
  rc = fetch_status();
  if (rc) return rc;
  bool selected = false;  // the last round ran the selection (its keep count is pending)
  const NeedArgs need_args = {L.g.pkg_of, L.g.rowptr, L.g.direct_max};
  for (;;) {
    unsigned long long cur8[8];
    std::memcpy(cur8, g_ws.h_stat, sizeof(cur8));
    const unsigned long long mc = g_ws.h_stat[8];
    if (selected) npend = cur8[7];  // pieces the selection kept
    const uint64_t nd = cur8[4];
    if (nd > rcap) {
      set_error("out-of-core: more deferred pieces in one round than the buffer holds");
      return PM_ERR_OVERFLOW;
    }
    rs.misses += mc;
    rs.deferred += nd;
    if (nd) {
      if (npend + nd > kPendMax) {
        set_error("out-of-core: too many deferred pieces pending");
        return PM_ERR_OVERFLOW;
      }
      // pending buffers hold npend + nd; the run buffer may take all
      for (int i = 0; i < 2; i++) {
        rc = grow_keep(&g_ws.pend[i], &g_ws.pend_bytes[i], sizeof(Resume) * (npend + nd),
                       i == pc ? sizeof(Resume) * npend : 0, st);
        if (rc) return rc;
      }
      rc = grow_keep(&g_ws.resume[0], &g_ws.resume_bytes[0], sizeof(Resume) * (npend + nd), 0,
                     st);
      if (rc) return rc;
      HIP_OK(hipMemcpyAsync((Resume*)g_ws.pend[pc] + npend, g_ws.resume[1],
                            sizeof(Resume) * nd, hipMemcpyDeviceToDevice, st));
      npend += nd;
    }
    const bool edges_left = cur8[0] < span;
    // tasks past the buffer's end were never written (their jobs kept the
    // work, see run_job), so only the first task_cap exist
    const uint64_t tasks_made = std::min<uint64_t>(cur8[2], g_ws.task_cap);
    const uint64_t tasks_done = std::min<uint64_t>(cur8[1], tasks_made);
    const bool tasks_left = L.tasks && tasks_done < tasks_made;
    const uint8_t* used = g_ws.h_used;
    if (!use_acc.empty()) {
      // trace: each package used this round counts once
      std::vector<uint8_t> seen(np, 0);
      for (uint64_t sl = 0; sl < pg->nslots; sl++) {
        const int32_t own = pg->policy.owner((uint32_t)sl);
        if (used[sl] && own >= 0 && !seen[own]) {
          seen[own] = 1;
          use_acc[own]++;
        }
      }
    }
    if (npend == 0 && !edges_left && !tasks_left) break;
    pg->policy.touched(used, (uint32_t)pg->nslots);
    // admission: the oldest pieces of the window while their needs fit
    const uint64_t nw = std::min(npend, kWindow);
    if (nw) {
      hipLaunchKernelGGL(k_piece_needs, dim3((unsigned)std::min<uint64_t>((nw + 255) / 256, 1024)),
                         dim3(256), 0, st, (const Resume*)g_ws.pend[pc], nw, need_args,
                         (uint32_t*)g_ws.d_needs);
      HIP_OK(hipMemcpyAsync(g_ws.h_needs, g_ws.d_needs, sizeof(uint32_t) * nw * kNeed,
                            hipMemcpyDeviceToHost, st));
      HIP_OK(hipStreamSynchronize(st));
    }
    std::vector<uint8_t> in_set(np, 0);
    std::vector<uint32_t> want;
    uint64_t set_slots = 0;
    size_t nadmit = 0;
    for (uint64_t i = 0; i < nw; i++) {
      const uint32_t* need = &g_ws.h_needs[i * kNeed + 1];
      const int nneed = (int)g_ws.h_needs[i * kNeed];
      uint64_t extra = 0;
      for (int q = 0; q < nneed; q++) {
        bool dup = in_set[need[q]] != 0;
        for (int z = 0; z < q && !dup; z++) dup = need[z] == need[q];
        if (!dup) extra += PL.nslots[need[q]];
      }
      if (set_slots + extra > pg->nslots) break;
      for (int q = 0; q < nneed; q++)
        if (!in_set[need[q]]) {
          in_set[need[q]] = 1;
          want.push_back(need[q]);
        }
      set_slots += extra;
      nadmit++;
    }
    if (nadmit == 0 && npend && !edges_left && !tasks_left) {
      set_error("out-of-core: a deferred search needs more packages than the cache holds");
      return PM_ERR_OVERFLOW;
    }
    std::vector<uint8_t> pinned(pg->nslots, 0);
    for (uint32_t q : want)
      if (pg->pkg_slot[q] >= 0)
        for (uint32_t w = 0; w < PL.nslots[q]; w++) pinned[pg->pkg_slot[q] + w] = 1;
    const auto c0 = std::chrono::steady_clock::now();
    uint64_t b = 0;
    rc = load_packages(pg, want, pinned, g_ws.copy_stream, &b);
    if (rc) return rc;
    rs.link_bytes += b;
    if (opts->prefetch && !by_score.empty()) {
      // learned prefetch: free slots take the packages of highest
      // predicted use that are not resident (nothing is evicted for it)
      uint64_t free_slots = 0;
      std::vector<uint8_t> keep(pg->nslots, 1);
      for (uint64_t sl = 0; sl < pg->nslots; sl++)
        if (pg->policy.owner((uint32_t)sl) < 0) {
          free_slots++;
          keep[sl] = 0;
        }
      std::vector<uint32_t> pre;
      for (uint32_t q : by_score) {
        if (free_slots == 0) break;
        if (pg->pkg_slot[q] >= 0 || PL.nslots[q] > free_slots) continue;
        pre.push_back(q);
        free_slots -= PL.nslots[q];
      }
      if (!pre.empty()) {
        rc = load_packages(pg, pre, keep, g_ws.copy_stream, &b);
        if (rc) return rc;
        rs.link_bytes += b;
      }
    }
    // one copy launch for the round's loads, the new slot map, one wait
    rc = flush_packages(pg, g_ws.copy_stream, false);
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(pg->d_pkg_slot, pg->pkg_slot.data(), sizeof(int32_t) * np,
                          hipMemcpyHostToDevice, g_ws.copy_stream));
    HIP_OK(hipStreamSynchronize(g_ws.copy_stream));
    pg->jobs.clear();
    rs.link_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0)
                   .count();
    // every pending piece whose needs are now resident runs this round;
    // the run count stays on the device (cursor[6]) for k_resume, the kept
    // count (cursor[7]) comes back with the next status
    const bool run_pieces = npend > 0;
    selected = run_pieces;
    if (run_pieces) {
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, g_ws.cursor + 6, 2, 0ull);
      hipLaunchKernelGGL(k_piece_select,
                         dim3((unsigned)std::min<uint64_t>((npend + 255) / 256, 16384)),
                         dim3(256), 0, st, (const Resume*)g_ws.pend[pc], npend, need_args,
                         pg->d_pkg_slot, (Resume*)g_ws.resume[0], (Resume*)g_ws.pend[1 - pc],
                         g_ws.cursor + 6);
      pc = 1 - pc;
    }
    // next round: continue the edge scan and the task queue, resume the
    // selected pieces (they may split and defer again)
    unsigned long long nc[5] = {cur8[0], tasks_left ? tasks_done : 0ull,
                                tasks_left ? cur8[2] : 0ull, 0ull, 0ull};
    HIP_OK(hipMemcpyAsync(g_ws.cursor, nc, sizeof(nc), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, st, (uint32_t*)pg->d_miss_bits,
                       (uint64_t)((np + 31) / 32 + 1), 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, pg->d_miss_count, 1, 0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, st, (uint32_t*)pg->d_used,
                       (pg->nslots + 3) / 4, 0u);
    L.rin = (const Resume*)g_ws.resume[0];
    L.nrin = run_pieces ? ~0ull : 0;  // the count is read from cursor[6]
    L.nrin_dev = run_pieces ? g_ws.cursor + 6 : nullptr;
    L.rout = (Resume*)g_ws.resume[1];
    const auto k0 = std::chrono::steady_clock::now();
    if (edges_left) hipLaunchKernelGGL(k_match_edges<true>, dim3(waves), dim3(kWave), 0, st, L);
    if (L.nrin) hipLaunchKernelGGL(k_resume<true>, dim3(waves), dim3(kWave), 0, st, L);
    // tasks left over, or written this round by the edge scan or by
    // resumed depth-0 pieces
    if (L.tasks && (tasks_left || run_pieces || edges_left))
      hipLaunchKernelGGL(k_match_tasks<true>, dim3(waves), dim3(kWave), 0, st, L);
    HIP_OK(hipGetLastError());
    rc = fetch_status();  // waits for the round's kernels
    if (rc) return rc;
    rs.kernel_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - k0).count();
    rs.rounds++;
    if (rs.rounds > 1000000) {
      set_error("out-of-core: no progress");
      return PM_ERR_OVERFLOW;
    }
  }
  L.rin = nullptr;
  L.nrin = 0;
  return PM_OK;
}

// Resident runs with re-split on: resume the pieces the last launch wrote
// (in resume[1]) over all waves, round after round, until a round writes
// none.  One wait per round, for the piece count.
int resident_rounds(Launch& L, hipStream_t st, int waves, uint64_t rcap, int* rounds) {
  int b = 1;
  for (;;) {
    HIP_OK(hipMemcpyAsync(g_ws.h_res, g_ws.cursor, sizeof(unsigned long long) * 8,
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const uint64_t n = g_ws.h_res[4];
    if (n == 0) break;
    if (n > rcap) {
      set_error("re-split: more pieces than the piece buffer holds");
      return PM_ERR_OVERFLOW;
    }
    L.rin = (const Resume*)g_ws.resume[b];
    L.nrin = n;
    L.nrin_dev = nullptr;
    L.rout = (Resume*)g_ws.resume[1 - b];
    L.claim = 1;  // few, heavy pieces: one per wave
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, g_ws.cursor + 3, 2, 0ull);
    hipLaunchKernelGGL((k_resume<false, true>), dim3(waves), dim3(kWave), 0, st, L);
    HIP_OK(hipGetLastError());
    b = 1 - b;
    if (++*rounds > 100000) {
      set_error("re-split: no progress");
      return PM_ERR_DEVICE;
    }
  }
  L.rin = nullptr;
  L.nrin = 0;
  L.rout = (Resume*)g_ws.resume[1];
  L.claim = 0;
  return PM_OK;
}

int device_match(pm_graph* g, const Plan& plan, const pm_opts* opts, pm_result* res) {
  const Plan* one = &plan;
  return device_match_batch(g, &one, 1, opts, res);
}

// Identity of a pendant-tree table: F_u depends only on u's label and the
// tables of its children, so equal subtrees of different patterns (or of one
// pattern) share one table.
static std::string fold_signature(const Plan& p, int u, bool product) {
  std::vector<std::string> kids;
  for (int c = 0; c < p.k; c++)
    if (p.fold_children[u] >> c & 1u) kids.push_back(fold_signature(p, c, false));
  std::sort(kids.begin(), kids.end());
  std::string s = product ? "P(" : ("L" + std::to_string(p.vlabel[u]) + "(");
  for (const std::string& k : kids) s += k + ",";
  return s + ")";
}

int device_match_batch(pm_graph* g, const Plan* const* plans, int npat, const pm_opts* opts,
                       pm_result* res) {
  if (npat <= 0) return PM_OK;
  const bool paged = opts->cache_bytes > 0;
  if (paged)
    for (int q = 0; q < npat; q++)
      if (plans[q]->induced) {
        set_error("vertex-induced matching needs the rows resident (cache_bytes = 0)");
        return PM_ERR_ARG;
      }
  int rc = device_upload_rows(g, !paged);
  if (rc) return rc;
  // learned parts: model predictions per vertex (cached per graph and model)
  const pm_model* model = opts->model;
  const std::vector<float>* pred = nullptr;
  if (model && (opts->learned & (PM_LEARN_ORDER | PM_LEARN_PAGING))) {
    rc = graph_predictions(g, model, &pred);
    if (rc) return rc;
  }
  PagedGraph* pg = nullptr;
  if (paged) {
    rc = ensure_paged(g, opts->cache_bytes, opts->package_bytes ? opts->package_bytes : (2u << 20),
                      &pg);
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> lk(g_ws_mu);
  rc = ensure_workspace(g->device);
  if (rc) return rc;
  const HostGraph& h = g->h;
  for (int q = 0; q < npat; q++)
    if (plans[q]->labeled && h.labels.empty()) {
      set_error("labeled matching needs vertex labels (open the graph with PM_GRAPH_LABELS)");
      return PM_ERR_ARG;
    }
  if (opts->per_edge && npat != 1) {
    set_error("per-edge counts are available for a single pattern only");
    return PM_ERR_ARG;
  }
  // labeled plans run in the cheapest core order (the count does not
  // depend on it); `plans` then points at the reordered copies
  if (!g->stats_ok) {
    g->stats = order_stats(h);
    g->stats_ok = true;
  }
  std::vector<Plan> reordered;
  std::vector<const Plan*> use(plans, plans + npat);
  if (opts->order == 0) {
    reordered.reserve(npat);
    for (int q = 0; q < npat; q++)
      if (plans[q]->labeled && plans[q]->kc >= 4) {
        reordered.push_back(*plans[q]);
        Plan& pl = reordered.back();
        if (pl.tuned_graph == (const void*)g)
          (void)apply_core_order(&pl, std::vector<int>(pl.tuned_core + 2, pl.tuned_core + pl.kc));
        else
          reorder_core(&pl, g->stats);
        use[q] = &pl;
      }
  }
  plans = use.data();
  std::vector<KPlan> kps(npat);
  int nreg = 0;
  bool any_split = false;
  for (int q = 0; q < npat; q++) {
    int nr = 0;
    build_kplan(*plans[q], !paged && g->d.nl > 0, &kps[q], &nr);
    nreg = std::max(nreg, nr);
    any_split |= plans[q]->kc >= 4;
  }
  hipStream_t st = opts->stream ? (hipStream_t)opts->stream : g_ws.stream;
  // default: every wave the chip can hold at once (persistent grid)
  if (g_ws.resident == 0) {
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_match_edges<false>, kWave, 0));
    g_ws.resident = std::max(1, per_cu) * g_ws.cus;
  }
  const int waves = opts->waves > 0 ? opts->waves : g_ws.resident;
  const uint64_t lo = opts->edge_lo, hi = opts->edge_hi ? opts->edge_hi : h.m;
  if (hi > h.m || lo > hi) {
    set_error("edge range outside the graph");
    return PM_ERR_ARG;
  }
  // sampled runs (autotuning): every sample-th block of kSampleBlock edges
  const uint32_t sample = opts->sample > 1 ? (uint32_t)opts->sample : 1u;
  uint64_t npos = hi - lo;
  if (sample > 1) {
    const uint64_t stride = kSampleBlock * sample;
    const uint64_t full = (hi - lo) / stride;
    npos = full * kSampleBlock + std::min<uint64_t>(kSampleBlock, (hi - lo) - full * stride);
  }
  const uint32_t chunk =
      opts->chunk > 0
          ? (uint32_t)opts->chunk
          : (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(64, npos * npat / ((uint64_t)waves * 8)));
  const uint64_t nchunk = (npos + chunk - 1) / chunk;
  // learned order applies to full-graph, unsampled runs
  const bool lorder = pred && (opts->learned & PM_LEARN_ORDER) && sample == 1 && lo == 0 && hi == h.m;
  if (lorder) {
    rc = learned_order(g, model, *pred);
    if (rc) return rc;
  }
  for (KPlan& kp : kps) {
    kp.split = opts->split > 0 ? (uint32_t)opts->split : 8192u;
    kp.hub_factor = opts->hub_factor > 0 ? (uint32_t)opts->hub_factor : 8u;
  }

  // fold / weight tables (u64[n] each): one per distinct pendant subtree and
  // per distinct weighted core vertex, over all patterns of the launch
  struct TableJob {
    int plan, u;
    bool product;
    std::vector<int> child;  // table ids
  };
  const uint64_t nn = std::max<uint64_t>(h.n, 1);
  std::map<std::string, int> table_id;
  std::vector<TableJob> table_jobs;
  std::vector<std::array<int, kMaxK>> wt_tab(npat);
  for (int q = 0; q < npat; q++) {
    const Plan& pl = *plans[q];
    int table_of[kMaxK];
    for (int u = 0; u < kMaxK; u++) table_of[u] = -1;
    auto make = [&](int u, bool product) {
      const std::string sig = fold_signature(pl, u, product);
      auto it = table_id.find(sig);
      if (it != table_id.end()) return it->second;
      TableJob j;
      j.plan = q;
      j.u = u;
      j.product = product;
      for (int c = 0; c < pl.k; c++)
        if (pl.fold_children[u] >> c & 1u) j.child.push_back(table_of[c]);
      const int id = (int)table_jobs.size();
      table_jobs.push_back(j);
      table_id.emplace(sig, id);
      return id;
    };
    for (int i = 0; i < pl.nfold; i++) table_of[pl.fold_v[i]] = make(pl.fold_v[i], false);
    wt_tab[q].fill(-1);
    for (int i = 0; i < pl.kc; i++)
      if (pl.fold_children[pl.core_v[i]]) wt_tab[q][i] = make(pl.core_v[i], true);
  }
  const int ntab = (int)table_jobs.size();
  if (ntab) {
    rc = grow(&g_ws.tables, &g_ws.table_bytes, sizeof(unsigned long long) * ntab * nn);
    if (rc) return rc;
  }
  auto table = [&](int t) { return (unsigned long long*)g_ws.tables + (size_t)t * nn; };
  for (int q = 0; q < npat; q++)
    for (int i = 0; i < plans[q]->kc; i++)
      kps[q].wt[i] = wt_tab[q][i] >= 0 ? table(wt_tab[q][i]) : nullptr;

  const uint32_t cap = std::max<uint32_t>(h.max_degree, 1);
  const uint64_t stride = (uint64_t)std::max(nreg, 1) * cap;
  rc = grow(&g_ws.arena, &g_ws.arena_bytes, (size_t)waves * stride * sizeof(uint32_t));
  if (rc) return rc;
  // resident re-split budget (µs -> wall-clock ticks)
  const int resplit_us = opts->resplit_us > 0 ? opts->resplit_us : 0;
  // a search can re-split at depths 1..k-3 (the fused step included), so
  // batches whose cores all have <= 3 vertices run without it
  bool can_resplit = false;
  for (int q = 0; q < npat; q++) can_resplit |= plans[q]->kc >= 4;
  const uint64_t budget =
      (paged || !can_resplit) ? 0
                              : (uint64_t)resplit_us * (uint64_t)std::max(g_ws.wall_khz, 1) / 1000u;
  const bool rs = budget != 0;
  // deferred pieces per round (out of core) or re-split pieces per round
  const uint64_t rcap = paged ? (1ull << 22)
                        : budget ? std::max<uint64_t>(1ull << 20, (uint64_t)waves * 128)
                                 : 1;
  for (int i = 0; i < 2; i++) {
    rc = grow(&g_ws.resume[i], &g_ws.resume_bytes[i], sizeof(Resume) * rcap);
    if (rc) return rc;
  }
  rc = grow(&g_ws.plans, &g_ws.plans_bytes, sizeof(KPlan) * npat);
  if (rc) return rc;
  rc = grow_results(npat);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(g_ws.plans, kps.data(), sizeof(KPlan) * npat, hipMemcpyHostToDevice, st));
  unsigned long long* d_per_edge = nullptr;
  if (opts->per_edge && hi > lo)
    HIP_OK(hipMallocAsync((void**)&d_per_edge, sizeof(unsigned long long) * (hi - lo), st));

  KGraph kg;
  std::memset(&kg, 0, sizeof(kg));
  kg.n = h.n;
  kg.m = h.m;
  kg.rowptr = g->d.rowptr;
  kg.colidx = g->d.colidx;
  kg.jcol = paged ? pg->host_rows_dev : g->d.colidx;
  kg.lab = g->d.label8;
  kg.maxnbr = g->d.maxnbr;
  kg.lcol = g->d.lcol;
  kg.loff = g->d.loff;
  kg.nl = paged ? 0 : g->d.nl;
  kg.src64 = g->d.src64;
  if (paged) {
    kg.pkg_slot = pg->d_pkg_slot;
    kg.pkg_of = pg->d_pkg_of;
    kg.pkg_base = pg->d_pkg_base;
    kg.cache = pg->d_cache;
    kg.slot_words = pg->slot_words;
    kg.miss_bits = pg->d_miss_bits;
    kg.miss_list = pg->d_miss_list;
    kg.miss_count = pg->d_miss_count;
    kg.miss_cap = pg->miss_cap;
    kg.used = pg->d_used;
    kg.direct_max = opts->direct_degree > 0    ? (uint32_t)opts->direct_degree
                    : opts->direct_degree == 0 ? 1024u
                                               : 0u;
  }
  Launch L;
  std::memset(&L, 0, sizeof(L));
  L.g = kg;
  L.plans = (const KPlan*)g_ws.plans;
  L.npat = (uint32_t)npat;
  L.nchunk = nchunk;
  L.npos = npos;
  L.sample = sample;
  L.pcount = g_ws.pcount;
  L.cursor = g_ws.cursor;
  L.edge_lo = lo;
  L.edge_hi = hi;
  L.chunk = chunk;
  L.cap = cap;
  L.arena = (uint32_t*)g_ws.arena;
  L.arena_stride = stride;
  L.out = g_ws.out;
  L.tasks = (any_split && opts->split >= 0) ? g_ws.tasks : nullptr;
  L.root_order = lorder ? g->d_root_order : nullptr;
  L.root_start = lorder ? g->d_root_start : nullptr;
  unsigned long long* d_root_work = nullptr;
  if (opts->root_work && h.n) {
    HIP_OK(hipMallocAsync((void**)&d_root_work, sizeof(unsigned long long) * h.n, st));
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, st, d_root_work, h.n, 0ull);
  }
  // learned paging: a package's score is the mean predicted use of its vertices
  std::vector<float> pscore;
  std::vector<uint32_t> by_score;  // packages, highest score first
  if (paged && pred && (opts->learned & PM_LEARN_PAGING)) {
    const PackageLayout& PLo = pg->layout;
    pscore.assign(PLo.count(), 0.0f);
    for (uint32_t q = 0; q < PLo.count(); q++) {
      const uint32_t v0 = PLo.first_vertex[q];
      const uint32_t v1 = q + 1 < PLo.count() ? PLo.first_vertex[q + 1] : (uint32_t)h.n;
      double sum = 0;
      for (uint32_t v = v0; v < v1; v++) sum += (*pred)[2 * v + 1];
      pscore[q] = v1 > v0 ? (float)(sum / (v1 - v0)) : 0.0f;
    }
    by_score.resize(PLo.count());
    std::iota(by_score.begin(), by_score.end(), 0u);
    std::stable_sort(by_score.begin(), by_score.end(),
                     [&](uint32_t a, uint32_t b) { return pscore[a] > pscore[b]; });
  }
  std::vector<uint64_t> use_acc;  // trace: rounds each package was used in
  if (paged && opts->package_use) use_acc.assign(pg->layout.count(), 0);
  L.task_cap = g_ws.task_cap;
  L.rout_cap = rcap;
  L.soft = rcap / 2;
  L.budget = budget;
  const int reps = opts->repeats > 0 ? opts->repeats : 1;
  double sum_ms = 0, min_ms = 1e30, phase[3] = {0, 0, 0}, link_ms = 0, round_kernel_ms = 0;
  uint64_t link_bytes = 0, misses = 0, deferred = 0;
  int rounds = 1;
  const int fold_blocks = (int)std::max<uint64_t>(
      1, std::min<uint64_t>((h.m + kFoldChunk - 1) / kFoldChunk,
                            (uint64_t)g_ws.cus * 16 * std::max<uint32_t>(1, 512 / kFoldChunk)));
  const uint32_t np = paged ? pg->layout.count() : 0;
  // pendant-tree tables, children first (table ids are in that order)
  struct FoldStep {
    int kind;  // 0 product, 1 leaf over label-grouped rows, 2 fill + fold
    FoldArgs A;
  };
  std::vector<FoldStep> folds;
  if (h.n)
    for (const TableJob& j : table_jobs) {
      const Plan& pl = *plans[j.plan];
      const int id = (int)(&j - table_jobs.data());
      // built in place after zeroing, so padding bytes are zero too (the
      // bytes take part in the graph signature)
      folds.emplace_back();
      FoldStep& f = folds.back();
      std::memset(&f, 0, sizeof(f));
      FoldArgs& A = f.A;
      A.g = kg;
      A.want = (uint32_t)pl.vlabel[j.u];
      for (int c : j.child) A.child[A.nchild++] = table(c);
      A.out = table(id);
      A.bytes = g_ws.out + 6;
      f.kind = j.product ? 0 : (A.nchild == 0 && !paged && kg.nl > 0 && A.want + 1 < kg.nl) ? 1 : 2;
    }
  auto launch_folds = [&]() {
    for (const FoldStep& f : folds) {
      if (f.kind == 0) {
        hipLaunchKernelGGL(k_product, dim3(1024), dim3(256), 0, st, h.n, f.A);
      } else if (f.kind == 1) {
        // a leaf with label-grouped rows: the label part's size, no row reads
        hipLaunchKernelGGL(k_fold_leaf, dim3((unsigned)std::min<uint64_t>((h.n + 255) / 256, 4096)),
                           dim3(256), 0, st, f.A);
      } else {
        hipLaunchKernelGGL(k_fill_u64, dim3(256), dim3(256), 0, st, f.A.out, h.n, 0ull);
        if (h.m) hipLaunchKernelGGL(k_fold, dim3(fold_blocks), dim3(kWave), 0, st, f.A);
      }
    }
  };
  // Graph mode (resident rows, no per-edge counts or traces): the whole
  // launch sequence of a run — result fill, tables, match kernels — is one
  // hipGraph, captured once and replayed while every argument is unchanged.
  const bool use_graph = opts->graph && !paged && !d_per_edge && !d_root_work;
  if (use_graph) {
    L.rout = (Resume*)g_ws.resume[1];
    std::string sig;
    auto put = [&](const void* p, size_t n) { sig.append((const char*)p, n); };
    put(&L, sizeof(L));
    for (const FoldStep& f : folds) put(&f, sizeof(f));
    put(&waves, sizeof(waves));
    put(&fold_blocks, sizeof(fold_blocks));
    put(&npat, sizeof(npat));
    put(&st, sizeof(st));
    if (!g_ws.gexec || sig != g_ws.gsig) {
      if (g_ws.gexec) HIP_OK(hipGraphExecDestroy(g_ws.gexec));
      g_ws.gexec = nullptr;
      g_ws.gsig.clear();
      HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, g_ws.cursor, 16 + 2ull * npat, 0ull);
      launch_folds();
      hipLaunchKernelGGL((rs ? k_match_edges<false, true> : k_match_edges<false, false>),
                         dim3(waves), dim3(kWave), 0, st, L);
      if (L.tasks)
        hipLaunchKernelGGL((rs ? k_match_tasks<false, true> : k_match_tasks<false, false>),
                           dim3(waves), dim3(kWave), 0, st, L);
      const hipError_t le = hipGetLastError();
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(st, &graph);
      if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        set_error(std::string("graph capture: ") +
                  hipGetErrorString(le != hipSuccess ? le : ce));
        return PM_ERR_DEVICE;
      }
      const hipError_t ie = hipGraphInstantiate(&g_ws.gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        g_ws.gexec = nullptr;
        set_error(std::string("graph instantiate: ") + hipGetErrorString(ie));
        return PM_ERR_DEVICE;
      }
      g_ws.gsig = sig;
    }
  }
  for (int r = 0; r < reps && use_graph; r++) {
    HIP_OK(hipEventRecord(g_ws.ev0, st));
    HIP_OK(hipGraphLaunch(g_ws.gexec, st));
    rounds = 1;
    if (budget) {
      rc = resident_rounds(L, st, waves, rcap, &rounds);
      if (rc) return rc;
    }
    HIP_OK(hipEventRecord(g_ws.ev1, st));
    HIP_OK(hipEventSynchronize(g_ws.ev1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, g_ws.ev0, g_ws.ev1));
    sum_ms += ms;
    min_ms = std::min(min_ms, (double)ms);
  }
  for (int r = 0; r < reps && !use_graph; r++) {
    // cursor, out and pcount in one fill
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, st, g_ws.cursor, 16 + 2ull * npat, 0ull);
    L.per_edge = (r == 0) ? d_per_edge : nullptr;
    L.root_work = (r == 0) ? d_root_work : nullptr;
    if (L.per_edge)
      hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, st, d_per_edge, hi - lo, 0ull);
    const auto t0 = std::chrono::steady_clock::now();
    link_bytes = misses = deferred = 0;
    link_ms = 0;
    round_kernel_ms = 0;
    rounds = 1;
    if (paged) {
      // empty cache, then the reference's initial preload (buffer_init_on_gpu,
      // src/buffer.cu:264-415): packages in id order until the cache is full
      pg->policy.reset((uint32_t)pg->nslots, opts->policy);
      if (!pscore.empty()) pg->policy.set_scores(pscore);
      pg->pkg_slot.assign(np, -1);
      std::vector<uint8_t> pinned(pg->nslots, 0);
      std::vector<uint32_t> pre;
      uint64_t used_slots = 0;
      const bool learned_pre = opts->prefetch && !by_score.empty();
      if (kg.direct_max > 0) {
        // Rows of <= direct_max entries are read directly and never miss,
        // so packages holding longer rows go first, longest first; the
        // remaining slots take the other packages in id order, or by
        // predicted use with learned prefetch (a cached short row is still
        // faster than a read over the link)
        if (pg->long_for != kg.direct_max) {
          pg->long_words.assign(np, 0);
          for (uint64_t v = 0; v < h.n; v++) {
            const uint64_t d = h.rowptr[v + 1] - h.rowptr[v];
            if (d > kg.direct_max) pg->long_words[pg->layout.pkg_of[v]] += d;
          }
          pg->long_for = kg.direct_max;
        }
        std::vector<uint32_t> order;
        for (uint32_t p = 0; p < np; p++)
          if (pg->long_words[p]) order.push_back(p);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
          return pg->long_words[a] > pg->long_words[b];
        });
        for (uint32_t p : order) {
          if (used_slots + pg->layout.nslots[p] > pg->nslots) continue;
          pre.push_back(p);
          used_slots += pg->layout.nslots[p];
        }
        for (uint32_t i = 0; i < np; i++) {
          const uint32_t p = learned_pre ? by_score[i] : i;
          if (pg->long_words[p]) continue;
          if (used_slots + pg->layout.nslots[p] > pg->nslots) {
            if (!learned_pre) break;
            continue;
          }
          pre.push_back(p);
          used_slots += pg->layout.nslots[p];
        }
      } else {
        // learned prefetch preloads the packages of highest predicted use;
        // otherwise packages in id order
        for (uint32_t i = 0; i < np; i++) {
          const uint32_t p = learned_pre ? by_score[i] : i;
          if (used_slots + pg->layout.nslots[p] > pg->nslots) {
            if (!learned_pre) break;
            continue;
          }
          pre.push_back(p);
          used_slots += pg->layout.nslots[p];
        }
      }
      HIP_OK(hipEventRecord(g_ws.evc, g_ws.copy_stream));
      uint64_t b = 0;
      const auto c0 = std::chrono::steady_clock::now();
      rc = load_packages(pg, pre, pinned, g_ws.copy_stream, &b);
      if (rc) return rc;
      rc = flush_packages(pg, g_ws.copy_stream);
      if (rc) return rc;
      link_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0)
                     .count();
      link_bytes += b;
      HIP_OK(hipMemcpyAsync(pg->d_pkg_slot, pg->pkg_slot.data(), sizeof(int32_t) * np,
                            hipMemcpyHostToDevice, g_ws.copy_stream));
      HIP_OK(hipEventRecord(g_ws.evc, g_ws.copy_stream));
      HIP_OK(hipStreamWaitEvent(st, g_ws.evc, 0));
      hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, st, (uint32_t*)pg->d_miss_bits,
                         (uint64_t)((np + 31) / 32 + 1), 0u);
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, pg->d_miss_count, 1, 0ull);
      hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, st, (uint32_t*)pg->d_used,
                         (pg->nslots + 3) / 4, 0u);
    }
    HIP_OK(hipEventRecord(g_ws.ev0, st));
    launch_folds();
    HIP_OK(hipEventRecord(g_ws.evf, st));
    L.rout = (Resume*)g_ws.resume[1];
    hipLaunchKernelGGL((paged ? k_match_edges<true, false>
                        : rs  ? k_match_edges<false, true>
                              : k_match_edges<false, false>),
                       dim3(waves), dim3(kWave), 0, st, L);
    HIP_OK(hipEventRecord(g_ws.eve, st));
    if (L.tasks)
      hipLaunchKernelGGL((paged ? k_match_tasks<true, false>
                          : rs  ? k_match_tasks<false, true>
                                : k_match_tasks<false, false>),
                         dim3(waves), dim3(kWave), 0, st, L);
    HIP_OK(hipGetLastError());
    if (budget) {
      rc = resident_rounds(L, st, waves, rcap, &rounds);
      if (rc) return rc;
    }
    HIP_OK(hipEventRecord(g_ws.ev1, st));
    HIP_OK(hipEventSynchronize(g_ws.ev1));
    float ms = 0, mf = 0, me = 0, mt = 0;
    HIP_OK(hipEventElapsedTime(&ms, g_ws.ev0, g_ws.ev1));
    HIP_OK(hipEventElapsedTime(&mf, g_ws.ev0, g_ws.evf));
    HIP_OK(hipEventElapsedTime(&me, g_ws.evf, g_ws.eve));
    HIP_OK(hipEventElapsedTime(&mt, g_ws.eve, g_ws.ev1));
    phase[0] = mf;
    phase[1] = me;
    phase[2] = mt;
    if (paged) {
      RoundStats rst;
      rc = run_rounds(pg, L, st, waves, (uint64_t)npat * nchunk, rcap, opts, by_score, use_acc, rst);
      if (rc) return rc;
      misses = rst.misses;
      deferred = rst.deferred;
      link_bytes += rst.link_bytes;
      link_ms += rst.link_ms;
      round_kernel_ms += rst.kernel_ms;
      rounds = rst.rounds;
    }
    const double wall =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const double t = paged ? wall : (double)ms;
    sum_ms += t;
    min_ms = std::min(min_ms, t);
  }
  // cursor, out and pcount in one copy into pinned memory
  HIP_OK(hipMemcpyAsync(g_ws.h_res, g_ws.resblk, sizeof(unsigned long long) * (16 + 2 * (size_t)npat),
                        hipMemcpyDeviceToHost, st));
  if (d_per_edge) {
    HIP_OK(hipMemcpyAsync(opts->per_edge + lo, d_per_edge, sizeof(unsigned long long) * (hi - lo),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipFreeAsync(d_per_edge, st));
  }
  if (d_root_work) {
    HIP_OK(hipMemcpyAsync(opts->root_work, d_root_work, sizeof(unsigned long long) * h.n,
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipFreeAsync(d_root_work, st));
  }
  HIP_OK(hipStreamSynchronize(st));
  const unsigned long long* cur = g_ws.h_res;
  const unsigned long long* out = g_ws.h_res + 8;
  const unsigned long long* pc = g_ws.h_res + 16;
  if (opts->package_use) {
    // per vertex: the rounds in which its package was used (0 when resident)
    for (uint64_t v = 0; v < h.n; v++)
      opts->package_use[v] =
          use_acc.empty() ? 0u
                          : (uint32_t)std::min<uint64_t>(use_acc[pg->layout.pkg_of[v]], 0xFFFFFFFFu);
  }
  if (!paged && cur[4] != 0) {
    set_error("internal: a resident-row run deferred work");
    return PM_ERR_DEVICE;
  }
  for (int q = 0; q < npat; q++) {
    pm_result* rq = &res[q];
    std::memset(rq, 0, sizeof(*rq));
    uint64_t total = pc[2 * q];
    if (!plans[q]->labeled) total *= (uint64_t)plans[q]->aut;  // test.cu:198
    rq->count = total;
    rq->jobs = pc[2 * q + 1];
    // launch-wide figures (the whole batch)
    rq->kernel_ms = sum_ms / reps;
    rq->min_ms = min_ms;
    rq->alg_bytes = out[2] + out[4] + out[6];
    rq->set_ops = out[3] + out[5];
    for (int i = 0; i < 3; i++) rq->phase_ms[i] = phase[i];
    rq->phase_alg_bytes[0] = out[6];
    rq->phase_alg_bytes[1] = out[2];
    rq->phase_alg_bytes[2] = out[4];
    rq->tasks = std::min<uint64_t>(cur[2], g_ws.task_cap);
    rq->waves = waves;
    rq->chunk = (int32_t)chunk;
    rq->link_bytes = link_bytes;
    rq->link_ms = link_ms;
    rq->round_kernel_ms = round_kernel_ms;
    rq->misses = misses;
    rq->deferred = deferred;
    rq->rounds = rounds;
    rq->packages = (int32_t)np;
  }
  return PM_OK;
}

}  // namespace pm
