// Subgraph-matching kernels for gfx950 (CDNA4, wave64).
//
// One persistent wavefront runs one depth-first search at a time
// (the reference: one warp per DFS, _parallel_match / match / extend,
// src/gpu_match.cu:855-1178).  Jobs are CSR edges (r,c) claimed in chunks
// with one global atomic (replaces the spin-locked job array of
// src/gpu_match.cu:321-384 and the host job scan src/job_queue.h:39-89; the
// job filter of src/job_queue.h:44-61 is evaluated lane-parallel on the
// device).  Set operations are wave-cooperative sorted-list intersections:
// each lane takes one element of the shorter list, searches the longer one
// inside a window that advances monotonically, and the hits are compacted in
// order with a 64-bit ballot + popcount (src/gpu_match.cu:238-319 does the
// same with 32-bit ballots and a binary search over the whole list).
// Candidate batches of 64 are staged in LDS together with the DFS stack.
// Intermediate sets live in a per-wave HBM arena sized so no set is ever
// truncated (the reference caps every set at 4096 entries,
// src/gpu_match.cu:302,316).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr int kWave = 64;
constexpr uint32_t kInf = 0xFFFFFFFFu;

struct KOp {
  int8_t kind;   // 0 VIEW, 1 INTERSECT
  int8_t slot;
  int8_t parent;
  int8_t ncons;
  uint32_t labmask;
  uint8_t cons_b[8];
};

struct KPlan {
  int32_t k, labeled, rc_order, nops;
  uint32_t deg0, deg1, deg2, lab0, lab1;
  uint32_t lab[8];        // label of pattern vertex u
  uint32_t B[8], X[8];
  int8_t op_begin[8], op_end[8];
  int8_t cand_slot[8];    // [d]: slot holding vertex d+1's candidates
  int8_t cand_check[8];   // [d]: labeled, candidates still need the label test
  int8_t region[32];      // slot -> arena region (-1: not materialised)
  int8_t last_kind;       // 0 reuse slot last_src, 1 view of N(phi(k-2)), 2 inline intersect
  int8_t last_src;
  int8_t last_check;
  int8_t pad;
  KOp ops[kMaxSlots];
};

struct KGraph {
  uint64_t n, m;
  const uint64_t* rowptr;
  const uint32_t* colidx;
  const uint8_t* lab;
  const uint32_t* maxnbr;
};

struct WaveState {
  const uint32_t* sptr[kMaxSlots];
  uint32_t slen[kMaxSlots];
  uint32_t phi[8];
  uint32_t cbase[8];
  uint32_t cnext[8];
  uint32_t clen[8];
  uint64_t cmask[8];
  uint32_t cval[8][kWave];
  uint32_t jr[kWave];
  uint32_t jc[kWave];
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

__device__ __forceinline__ uint32_t degree(const KGraph& g, uint32_t v) {
  return (uint32_t)(g.rowptr[v + 1] - g.rowptr[v]);
}

// Wave-cooperative intersection of two sorted lists (already truncated).
// WRITE: store the hits in order at out.  LABEL: keep x only when bit
// lab[x] of labmask is set.  Returns the number of kept elements.
template <bool WRITE, bool LABEL>
__device__ uint32_t wave_isect(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb,
                               uint32_t labmask, const uint8_t* __restrict__ lab, uint32_t* out,
                               uint32_t lane, uint64_t& alg) {
  if (na > nb) {
    const uint32_t* t = a;
    a = b;
    b = t;
    uint32_t tn = na;
    na = nb;
    nb = tn;
  }
  alg += 4ull * (na + nb);
  uint32_t cnt = 0, lo = 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (uint32_t base = 0; base < na && lo < nb; base += kWave) {
    const uint32_t i = base + lane;
    bool hit = false;
    uint32_t x = 0, pos = nb;
    if (i < na) {
      x = a[i];
      pos = lo + below(b + lo, nb - lo, x);
      hit = pos < nb && b[pos] == x;
      if (LABEL && hit) hit = (labmask >> lab[x]) & 1u;
    }
    const uint64_t m = __ballot(hit);
    if (WRITE && hit) out[cnt + __popcll(m & lt)] = x;
    cnt += (uint32_t)__popcll(m);
    const uint32_t last = min(na - base, (uint32_t)kWave) - 1u;
    lo = uni((uint32_t)__shfl((int)pos, (int)last));
  }
  if (WRITE) alg += 4ull * cnt;
  return cnt;
}

// Count x in a[0..n) with lab[x] == want (wave-parallel).
__device__ uint32_t wave_count_label(const uint32_t* a, uint32_t n, uint32_t want,
                                     const uint8_t* __restrict__ lab, uint32_t lane, uint64_t& alg) {
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < n; base += kWave) {
    const uint32_t i = base + lane;
    bool hit = i < n && lab[a[i]] == want;
    cnt += (uint32_t)__popcll(__ballot(hit));
  }
  alg += 5ull * n;
  return cnt;
}

// truncation bound of a slot: max over consumers of min over their
// symmetry partners matched so far; kInf when any consumer is unconstrained.
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

__device__ __forceinline__ uint32_t vertex_bound(const KPlan& P, const WaveState& ws, int u) {
  uint32_t b = P.B[u], ub = kInf;
  while (b) {
    int j = __builtin_ctz(b);
    b &= b - 1;
    ub = min(ub, ws.phi[j]);
  }
  return ub;
}

// run the slot operations of depth d (vertex d just matched)
__device__ void run_ops(const KGraph& g, const KPlan& P, WaveState& ws, int d, uint32_t* arena,
                        uint32_t cap, uint32_t lane, uint64_t& alg, uint64_t& nops) {
  const uint32_t v = uni(ws.phi[d]);
  const uint32_t* nv = uniptr(g.colidx + g.rowptr[v]);
  const uint32_t dv = uni(degree(g, v));
  for (int o = P.op_begin[d]; o < P.op_end[d]; o++) {
    const KOp& op = P.ops[o];
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
      if (P.labeled)
        cnt = wave_isect<true, true>(a, na, nv, nb, op.labmask, g.lab, out, lane, alg);
      else
        cnt = wave_isect<true, false>(a, na, nv, nb, 0, g.lab, out, lane, alg);
      ws.sptr[op.slot] = out;
      ws.slen[op.slot] = uni(cnt);
      nops++;
    }
  }
}

__device__ __forceinline__ void setup_cand(const KPlan& P, WaveState& ws, int d) {
  const int s = P.cand_slot[d];
  uint32_t n = uni(ws.slen[s]);
  if (!P.labeled) {
    const uint32_t ub = vertex_bound(P, ws, d + 1);
    if (ub != kInf) n = uni(below(uniptr(ws.sptr[s]), n, ub));
  }
  ws.clen[d] = n;
  ws.cnext[d] = 0;
  ws.cbase[d] = 0;
  ws.cmask[d] = 0;
}

// next admitted candidate of vertex d+1; false when exhausted
__device__ __forceinline__ bool next_cand(const KGraph& g, const KPlan& P, WaveState& ws, int d,
                                          uint32_t lane, uint32_t& x, uint64_t& alg) {
  for (;;) {
    uint64_t m = uni64(ws.cmask[d]);
    if (m) {
      const int i = __builtin_ctzll(m);
      ws.cmask[d] = m & (m - 1);
      x = uni(ws.cval[d][i]);
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
    const int u = d + 1;
    if (ok) {
      v = c[idx];
      if (P.labeled) {
        if (P.cand_check[d]) ok = g.lab[v] == P.lab[u];
      } else {
        uint32_t xm = P.X[u];
        while (xm) {
          int j = __builtin_ctz(xm);
          xm &= xm - 1;
          ok = ok && (v != ws.phi[j]);
        }
      }
    }
    ws.cval[d][lane] = v;
    alg += 4ull * min(n - base, (uint32_t)kWave);
    ws.cmask[d] = __ballot(ok);
    __syncthreads();
  }
}

// matches of the last pattern vertex k-1 once vertices 0..k-2 are fixed
__device__ uint64_t last_count(const KGraph& g, const KPlan& P, WaveState& ws, uint32_t lane,
                               uint64_t& alg, uint64_t& nops) {
  const int u = P.k - 1;
  const int dl = P.k - 2;
  const uint32_t ub = P.labeled ? kInf : vertex_bound(P, ws, u);
  const uint32_t* a;
  uint32_t na;
  if (P.last_kind == 0) {
    a = uniptr(ws.sptr[P.last_src]);
    na = uni(ws.slen[P.last_src]);
  } else {
    const uint32_t v = uni(ws.phi[dl]);
    a = uniptr(g.colidx + g.rowptr[v]);
    na = uni(degree(g, v));
  }
  if (ub != kInf) na = uni(below(a, na, ub));
  if (P.last_kind != 2) {
    if (P.labeled) {
      if (!P.last_check) return na;
      nops++;
      return uni(wave_count_label(a, na, P.lab[u], g.lab, lane, alg));
    }
    // injectivity: drop earlier non-neighbours that sit in the list
    uint32_t drop = 0;
    const uint32_t xm = P.X[u];
    bool in = false;
    if (lane < (uint32_t)u && ((xm >> lane) & 1u)) {
      const uint32_t y = ws.phi[lane];
      const uint32_t p = below(a, na, y);
      in = p < na && a[p] == y;
    }
    drop = (uint32_t)__popcll(__ballot(in));
    return (uint64_t)(na - drop);
  }
  // inline intersection of the parent slot with N(phi(k-2)), count only
  const uint32_t* p = uniptr(ws.sptr[P.last_src]);
  uint32_t np = uni(ws.slen[P.last_src]);
  if (ub != kInf) np = uni(below(p, np, ub));
  nops++;
  if (P.labeled) {
    return uni(wave_isect<false, true>(p, np, a, na, 1u << P.lab[u], g.lab, nullptr, lane, alg));
  }
  uint32_t cnt = uni(wave_isect<false, false>(p, np, a, na, 0, g.lab, nullptr, lane, alg));
  const uint32_t xm = P.X[u];
  bool in = false;
  if (lane < (uint32_t)u && ((xm >> lane) & 1u)) {
    const uint32_t y = ws.phi[lane];
    const uint32_t q1 = below(a, na, y);
    const uint32_t q2 = below(p, np, y);
    in = q1 < na && a[q1] == y && q2 < np && p[q2] == y;
  }
  cnt -= (uint32_t)__popcll(__ballot(in));
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
  // now rowptr[lo] <= e < rowptr[min(lo+step, n)]
  uint32_t hi = min(lo + step, n);
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (g.rowptr[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(64) void k_match(KGraph g, const KPlan* __restrict__ Pp,
                                              unsigned long long* cursor, uint64_t edge_lo,
                                              uint64_t edge_hi, uint32_t chunk, uint32_t* arena,
                                              uint64_t arena_stride, uint32_t cap,
                                              unsigned long long* out, unsigned long long* per_edge) {
  __shared__ WaveState ws;
  const KPlan& P = *Pp;
  const uint32_t lane = threadIdx.x;
  uint32_t* my_arena = arena + (size_t)blockIdx.x * arena_stride;
  uint64_t total = 0, jobs = 0, alg = 0, nops = 0;
  const uint64_t span = edge_hi - edge_lo;
  const int k = P.k;

  for (;;) {
    unsigned long long c0 = 0;
    if (lane == 0) c0 = atomicAdd(cursor, (unsigned long long)chunk);
    c0 = uni64(__shfl(c0, 0));
    if (c0 >= span) break;
    const uint64_t e0 = edge_lo + c0;
    const uint64_t e1 = min(e0 + chunk, edge_hi);
    // source of e0: binary search over all vertices (uniform)
    uint32_t rlo = 0;
    {
      uint32_t lo = 0, hi = (uint32_t)g.n;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (g.rowptr[mid] <= e0) lo = mid; else hi = mid;
      }
      rlo = uni(lo);
    }
    for (uint64_t eb = e0; eb < e1; eb += kWave) {
      const uint64_t e = eb + lane;
      bool ok = e < e1;
      uint32_t r = rlo, c = 0;
      if (ok) {
        r = source_of(g, e, rlo);
        c = g.colidx[e];
        if (!P.labeled && P.rc_order && !(r > c)) ok = false;
        if (ok && P.labeled) ok = g.lab[r] == P.lab0 && g.lab[c] == P.lab1;
        if (ok) ok = degree(g, r) >= P.deg0 && degree(g, c) >= P.deg1 && g.maxnbr[c] >= P.deg2;
      }
      const uint32_t lastl = (uint32_t)min((uint64_t)kWave, e1 - eb) - 1u;
      uint64_t jm = __ballot(ok);
      ws.jr[lane] = r;
      ws.jc[lane] = c;
      __syncthreads();
      rlo = uni((uint32_t)__shfl((int)r, (int)lastl));
      alg += 24ull * min((uint64_t)kWave, e1 - eb);
      jm = uni64(jm);
      while (jm) {
        const int i = __builtin_ctzll(jm);
        jm &= jm - 1;
        jobs++;
        ws.phi[0] = uni(ws.jr[i]);
        ws.phi[1] = uni(ws.jc[i]);
        uint64_t cnt = 0;
        run_ops(g, P, ws, 0, my_arena, cap, lane, alg, nops);
        if (k == 3) {
          cnt = last_count(g, P, ws, lane, alg, nops);
        } else {
          run_ops(g, P, ws, 1, my_arena, cap, lane, alg, nops);
          setup_cand(P, ws, 1);
          int d = 1;
          for (;;) {
            uint32_t x;
            if (!next_cand(g, P, ws, d, lane, x, alg)) {
              if (d == 1) break;
              d--;
              continue;
            }
            ws.phi[d + 1] = x;
            if (d + 1 == k - 2) {
              cnt += last_count(g, P, ws, lane, alg, nops);
              continue;
            }
            run_ops(g, P, ws, d + 1, my_arena, cap, lane, alg, nops);
            setup_cand(P, ws, d + 1);
            d++;
          }
        }
        total += cnt;
        if (per_edge && lane == 0) per_edge[eb + i - edge_lo] = cnt;
      }
      __syncthreads();
    }
  }
  if (lane == 0) {
    out[blockIdx.x * 4 + 0] = total;
    out[blockIdx.x * 4 + 1] = jobs;
    out[blockIdx.x * 4 + 2] = alg;
    out[blockIdx.x * 4 + 3] = nops;
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

#define HIP_OK(call)                                                              \
  do {                                                                            \
    hipError_t _e = (call);                                                       \
    if (_e != hipSuccess) {                                                       \
      set_error(std::string(#call) + ": " + hipGetErrorString(_e));               \
      return PM_ERR_DEVICE;                                                       \
    }                                                                             \
  } while (0)

// Per-process work areas, grown on demand and reused across calls.
struct Workspace {
  int device = -1;
  hipStream_t stream = nullptr;
  void* arena = nullptr;
  size_t arena_bytes = 0;
  unsigned long long* cursor = nullptr;
  unsigned long long* out = nullptr;
  size_t out_waves = 0;
  KPlan* plan = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 0;
};
std::mutex g_ws_mu;
Workspace g_ws;

int ensure_workspace(int device) {
  if (g_ws.device == device) return PM_OK;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&g_ws.stream, hipStreamNonBlocking));
  HIP_OK(hipMalloc(&g_ws.cursor, 64));
  HIP_OK(hipMalloc(&g_ws.plan, sizeof(KPlan)));
  HIP_OK(hipEventCreate(&g_ws.ev0));
  HIP_OK(hipEventCreate(&g_ws.ev1));
  HIP_OK(hipDeviceGetAttribute(&g_ws.cus, hipDeviceAttributeMultiprocessorCount, device));
  g_ws.device = device;
  return PM_OK;
}

void build_kplan(const Plan& p, KPlan* kp, int* nregions) {
  std::memset(kp, 0, sizeof(*kp));
  kp->k = p.k;
  kp->labeled = p.labeled;
  kp->rc_order = p.rc_order;
  kp->deg0 = p.deg[0];
  kp->deg1 = p.deg[1];
  kp->deg2 = p.deg[2];
  kp->lab0 = p.vlabel[0];
  kp->lab1 = p.vlabel[1];
  for (int u = 0; u < p.k; u++) {
    kp->lab[u] = p.vlabel[u];
    kp->B[u] = p.B[u];
    kp->X[u] = p.X[u];
  }
  for (int s = 0; s < kMaxSlots; s++) kp->region[s] = -1;
  int nops = 0, nreg = 0;
  for (int d = 0; d <= p.k - 2; d++) {
    kp->op_begin[d] = (int8_t)nops;
    if (d < p.k - 2) {
      for (const SlotOp& op : p.ops[d]) {
        KOp& ko = kp->ops[nops++];
        ko.kind = op.kind;
        ko.slot = op.slot;
        ko.parent = op.parent;
        ko.ncons = (int8_t)op.ncons;
        ko.labmask = op.labmask;
        for (int c = 0; c < op.ncons; c++) ko.cons_b[c] = op.cons_b[c];
        if (op.kind == 1) kp->region[op.slot] = (int8_t)nreg++;
      }
    }
    kp->op_end[d] = (int8_t)nops;
  }
  kp->nops = nops;
  for (int d = 1; d <= p.k - 3; d++) {
    const int s = p.slot_of[d][d + 1];
    kp->cand_slot[d] = (int8_t)s;
    // a materialised slot whose label filter is exactly this vertex's label
    // needs no further test
    bool exact = false;
    for (int dd = 0; dd <= d; dd++)
      for (const SlotOp& op : p.ops[dd])
        if (op.slot == s && op.kind == 1 && op.labmask == (1u << p.vlabel[d + 1])) exact = true;
    kp->cand_check[d] = exact ? 0 : 1;
  }
  const int dl = p.k - 2;
  if (p.ops[dl].empty()) {
    kp->last_kind = 0;
    kp->last_src = (int8_t)p.slot_of[dl][p.k - 1];
    bool exact = false;
    for (int dd = 0; dd < dl; dd++)
      for (const SlotOp& op : p.ops[dd])
        if (op.slot == kp->last_src && op.kind == 1 && op.labmask == (1u << p.vlabel[p.k - 1]))
          exact = true;
    kp->last_check = exact ? 0 : 1;
  } else {
    const SlotOp& op = p.ops[dl][0];
    kp->last_kind = op.kind == 0 ? 1 : 2;
    kp->last_src = op.parent;
    kp->last_check = 1;
  }
  *nregions = nreg;
}

}  // namespace

int device_upload(pm_graph* g) {
  if (g->on_device) return PM_OK;
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  int rc = ensure_workspace(dev);
  if (rc) return rc;
  const HostGraph& h = g->h;
  std::vector<uint8_t> lab8(h.n, 0);
  for (uint64_t v = 0; v < h.n && !h.labels.empty(); v++) lab8[v] = (uint8_t)h.labels[v];
  std::vector<uint32_t> maxnbr;
  compute_max_neighbour_degree(h, &maxnbr);
  DeviceGraph& d = g->d;
  HIP_OK(hipMalloc(&d.rowptr, sizeof(uint64_t) * (h.n + 1)));
  HIP_OK(hipMalloc(&d.colidx, sizeof(uint32_t) * std::max<uint64_t>(h.m, 1)));
  HIP_OK(hipMalloc(&d.label8, std::max<uint64_t>(h.n, 1)));
  HIP_OK(hipMalloc(&d.maxnbr, sizeof(uint32_t) * std::max<uint64_t>(h.n, 1)));
  HIP_OK(hipMemcpy(d.rowptr, h.rowptr.data(), sizeof(uint64_t) * (h.n + 1), hipMemcpyHostToDevice));
  if (h.m) HIP_OK(hipMemcpy(d.colidx, h.colidx.data(), sizeof(uint32_t) * h.m, hipMemcpyHostToDevice));
  if (h.n) {
    HIP_OK(hipMemcpy(d.label8, lab8.data(), h.n, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d.maxnbr, maxnbr.data(), sizeof(uint32_t) * h.n, hipMemcpyHostToDevice));
  }
  d.bytes = sizeof(uint64_t) * (h.n + 1) + sizeof(uint32_t) * h.m + h.n + 4 * h.n;
  g->on_device = true;
  g->device = dev;
  return PM_OK;
}

int device_update_labels(pm_graph* g) {
  if (!g->on_device) return PM_OK;
  const HostGraph& h = g->h;
  std::vector<uint8_t> lab8(h.n, 0);
  for (uint64_t v = 0; v < h.n && !h.labels.empty(); v++) lab8[v] = (uint8_t)h.labels[v];
  if (h.n) HIP_OK(hipMemcpy(g->d.label8, lab8.data(), h.n, hipMemcpyHostToDevice));
  return PM_OK;
}

void device_release(pm_graph* g) {
  if (!g->on_device) return;
  (void)hipFree(g->d.rowptr);
  (void)hipFree(g->d.colidx);
  (void)hipFree(g->d.label8);
  (void)hipFree(g->d.maxnbr);
  g->d = DeviceGraph();
  g->on_device = false;
}

int device_match(pm_graph* g, const Plan& plan, const pm_opts* opts, pm_result* res) {
  int rc = device_upload(g);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  rc = ensure_workspace(g->device);
  if (rc) return rc;
  const HostGraph& h = g->h;
  KPlan kp;
  int nreg = 0;
  build_kplan(plan, &kp, &nreg);
  if (plan.labeled) {
    if (h.labels.empty()) {
      set_error("labeled matching needs vertex labels (open the graph with PM_GRAPH_LABELS)");
      return PM_ERR_ARG;
    }
  }
  hipStream_t st = opts->stream ? (hipStream_t)opts->stream : g_ws.stream;
  int waves = opts->waves > 0 ? opts->waves : g_ws.cus * 16;
  uint32_t chunk = opts->chunk > 0 ? (uint32_t)opts->chunk : 64u;
  uint64_t lo = opts->edge_lo, hi = opts->edge_hi ? opts->edge_hi : h.m;
  if (hi > h.m || lo > hi) {
    set_error("edge range outside the graph");
    return PM_ERR_ARG;
  }
  const uint32_t cap = std::max<uint32_t>(h.max_degree, 1);
  const uint64_t stride = (uint64_t)std::max(nreg, 1) * cap;
  const size_t need = (size_t)waves * stride * sizeof(uint32_t);
  if (need > g_ws.arena_bytes) {
    if (g_ws.arena) HIP_OK(hipFree(g_ws.arena));
    g_ws.arena = nullptr;
    g_ws.arena_bytes = 0;
    HIP_OK(hipMalloc(&g_ws.arena, need));
    g_ws.arena_bytes = need;
  }
  if ((size_t)waves > g_ws.out_waves) {
    if (g_ws.out) HIP_OK(hipFree(g_ws.out));
    HIP_OK(hipMalloc(&g_ws.out, sizeof(unsigned long long) * 4 * waves));
    g_ws.out_waves = waves;
  }
  HIP_OK(hipMemcpyAsync(g_ws.plan, &kp, sizeof(kp), hipMemcpyHostToDevice, st));
  unsigned long long* d_per_edge = nullptr;
  if (opts->per_edge && hi > lo) {
    HIP_OK(hipMallocAsync((void**)&d_per_edge, sizeof(unsigned long long) * (hi - lo), st));
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, st, d_per_edge, hi - lo, 0ull);
  }
  KGraph kg{h.n, h.m, g->d.rowptr, g->d.colidx, g->d.label8, g->d.maxnbr};
  const int reps = opts->repeats > 0 ? opts->repeats : 1;
  double sum_ms = 0, min_ms = 1e30;
  std::vector<unsigned long long> out(4 * (size_t)waves);
  for (int r = 0; r < reps; r++) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, g_ws.cursor, 1, 0ull);
    HIP_OK(hipEventRecord(g_ws.ev0, st));
    hipLaunchKernelGGL(k_match, dim3(waves), dim3(kWave), 0, st, kg, g_ws.plan, g_ws.cursor, lo, hi,
                       chunk, (uint32_t*)g_ws.arena, stride, cap, g_ws.out,
                       (r == 0) ? d_per_edge : nullptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(g_ws.ev1, st));
    HIP_OK(hipEventSynchronize(g_ws.ev1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, g_ws.ev0, g_ws.ev1));
    sum_ms += ms;
    min_ms = std::min(min_ms, (double)ms);
  }
  HIP_OK(hipMemcpyAsync(out.data(), g_ws.out, sizeof(unsigned long long) * 4 * waves,
                        hipMemcpyDeviceToHost, st));
  if (d_per_edge) {
    HIP_OK(hipMemcpyAsync(opts->per_edge + lo, d_per_edge, sizeof(unsigned long long) * (hi - lo),
                          hipMemcpyDeviceToHost, st));
    HIP_OK(hipFreeAsync(d_per_edge, st));
  }
  HIP_OK(hipStreamSynchronize(st));
  uint64_t total = 0, jobs = 0, alg = 0, nops = 0;
  for (int w = 0; w < waves; w++) {
    total += out[4 * w];
    jobs += out[4 * w + 1];
    alg += out[4 * w + 2];
    nops += out[4 * w + 3];
  }
  if (!plan.labeled) total *= (uint64_t)plan.aut;  // test.cu:198
  std::memset(res, 0, sizeof(*res));
  res->kernel_ms = sum_ms / reps;
  res->min_ms = min_ms;
  res->count = total;
  res->jobs = jobs;
  res->alg_bytes = alg;
  res->set_ops = nops;
  res->waves = waves;
  res->chunk = (int32_t)chunk;
  return PM_OK;
}

}  // namespace pm
