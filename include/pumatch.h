/*
 * pumatch.h -- C ABI of the MI355X subgraph matcher (libpumatch.so).
 *
 * The reference (PUMatch) exposes no library: its boundary is the CLI
 * `test <graph-prefix> <pattern.g>` (test.cu:15-212), the binary CSR files
 * read by GraphPreprocessor (src/graph.h:228-254), the .g pattern files read
 * by PatternPreprocessor (src/pattern.h:67-100) and the compile-time switch
 * LABELED (src/config.h:14).  Each entry point below replaces one step of
 * that driver; the comment on each names the reference code it stands in for.
 *
 * Conventions: every function returns PM_OK (0) or a negative PM_ERR_*;
 * pm_last_error() returns a message for the calling thread.  Handles own
 * their host and HBM memory and are released by the matching *_destroy.
 * Calls on one handle are serial.  No torch types cross this boundary.
 */
#ifndef PUMATCH_H
#define PUMATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_MAX_PATTERN 7 /* PAT_SIZE, src/config.h:9 */

enum {
  PM_OK = 0,
  PM_ERR_IO = -1,       /* file missing / short read (src/graph.h:219-223 exits) */
  PM_ERR_FORMAT = -2,   /* malformed .meta.txt / .g / CSR */
  PM_ERR_PATTERN = -3,  /* pattern violates the matcher's preconditions */
  PM_ERR_DEVICE = -4,   /* HIP error / no device */
  PM_ERR_ARG = -5,      /* bad argument */
  PM_ERR_OVERFLOW = -6, /* a device work area would overflow */
  PM_ERR_NOMEM = -7
};

/* out-of-core cache replacement policies */
#define PM_POLICY_LRU 0     /* least recently used slot (lruAccess_bucket, src/buffer.cu:581-628) */
#define PM_POLICY_FIFO 1    /* oldest load first */
#define PM_POLICY_LEARNED 2 /* lowest predicted reuse first (pm_model scores), then LRU */

/* graph flags */
#define PM_GRAPH_LABELS 0x1    /* also read <prefix>.label.bin */
#define PM_GRAPH_HOST_ONLY 0x2 /* keep the CSR on the host (no HBM copy) */

typedef struct pm_graph pm_graph;
typedef struct pm_pattern pm_pattern;

typedef struct {
  int32_t waves;        /* persistent wavefronts (0 = auto: 16 per CU) */
  int32_t chunk;        /* jobs (CSR edges) claimed per atomic (0 = auto) */
  int32_t repeats;      /* timed launches; kernel_ms is their mean (0 = 1) */
  int32_t split;        /* heavy-root split: a job's vertex-2 candidates are
                           cut into range tasks of about `split` list
                           entries of work (0 = auto 8192, < 0 = off) */
  int32_t hub_factor;   /* fused tail: candidates whose list is longer than
                           hub_factor x the shared set go one by one (0 = 8) */
  int32_t policy;       /* out-of-core replacement policy (PM_POLICY_*) */
  uint64_t edge_lo;     /* restrict jobs to CSR edges [edge_lo, edge_hi); hi=0: all */
  uint64_t edge_hi;
  uint64_t* per_edge;   /* optional host u64[m]: count per job edge (before x|Aut|) */
  void* stream;         /* hipStream_t to launch on (NULL: the library's own) */
  uint64_t cache_bytes; /* out-of-core: rows stay in pinned host memory and are
                           paged into an HBM cache of this size (0 = rows
                           resident in HBM).  Replaces the package buffer,
                           src/buffer.cu:264-415,800-929 */
  uint32_t package_bytes; /* cache slot / package size (0 = 2 MiB) */
  int32_t prefetch;     /* out-of-core: load predicted packages before each round */
  int32_t order;        /* labeled patterns: 0 = cost-based order of pattern
                           vertices 2.. (same count, less work), 1 = file
                           order as the reference (src/pattern.h:67-100) */
  int32_t sample;       /* > 1: jobs only from every sample-th block of 4096
                           CSR edges (timing samples; the count is then that
                           of the sampled jobs) */
  int32_t learned;      /* PM_LEARN_* bits, used when model is set */
  const struct pm_model* model;
  uint64_t* root_work;  /* trace, optional host u64[n]: work units (algorithmic
                           bytes) of the jobs rooted at each vertex */
  uint32_t* package_use; /* trace, optional host u32[n]: out-of-core rounds in
                           which each vertex's package was used (0 resident) */
  int32_t direct_degree; /* out-of-core: a missing row of at most this many
                           entries is read straight from pinned host memory
                           instead of deferring its search, as the reference
                           reads rows of <= DEGREE_THRESHOLD entries through
                           managed memory (src/config.h:41,
                           src/gpu_match.cu:478,601,663,752).  0 = 1024 (the
                           reference's value), < 0 = off */
  int32_t graph;        /* 1: run the launch sequence (result fill, pendant
                           tables, match kernels) as one hipGraph, captured on
                           first use and replayed by later runs and calls
                           whose kernel arguments are all unchanged.  Rows
                           resident, no per_edge / traces (else ignored);
                           phase_ms stays 0 */
  int32_t resplit_us;   /* rows resident: a search that has run longer than
                           this many microseconds hands its untried
                           candidates back as pieces, which further rounds
                           spread over all waves (work stealing, as
                           src/gpu_match.cu:21-148,818-853).  <= 0 = off */
} pm_opts;

#define PM_LEARN_ORDER 0x1  /* phase-1 jobs heaviest predicted root first */
#define PM_LEARN_PAGING 0x2 /* predicted package use ranks eviction/prefetch */

typedef struct {
  double kernel_ms;     /* mean kernel time (hipEvents on the launch stream) */
  double min_ms;        /* fastest of the timed launches */
  uint64_t count;       /* match count (x|Aut| when unlabeled, test.cu:198) */
  uint64_t jobs;        /* jobs that passed the filter (src/job_queue.h:44-61) */
  uint64_t alg_bytes;   /* algorithmic bytes of one launch: list entries
                           read and written x 4, job and table entries */
  uint64_t set_ops;     /* set operations executed by one launch */
  uint64_t link_bytes;  /* host->HBM bytes paged during the run (out-of-core) */
  uint64_t tasks;       /* range tasks created by heavy-root splitting */
  int32_t waves;        /* wavefronts actually launched */
  int32_t chunk;
  /* per kernel phase of the last timed launch (hipEvents on the launch
     stream): 0 pendant-tree folding, 1 edge phase, 2 range-task phase
     (with re-split rounds, if any); all 0 in graph mode */
  double phase_ms[3];
  uint64_t phase_alg_bytes[3];
  uint64_t misses;      /* out-of-core: package misses over all rounds */
  uint64_t deferred;    /* out-of-core: branches deferred and resumed */
  int32_t rounds;       /* out-of-core matching rounds, or resident re-split
                           rounds + 1 (1 when neither ran) */
  int32_t packages;     /* out-of-core: packages in the graph */
  double link_ms;       /* out-of-core: host->HBM copy time (preload included) */
  double round_kernel_ms; /* out-of-core: kernel time of rounds 2.. (the rest of
                           kernel_ms minus link_ms is host-side round work) */
} pm_result;

typedef struct {
  uint64_t n, m;            /* vertices, directed edges */
  uint32_t max_degree;
  int32_t has_labels;
  int32_t on_device;
  int32_t reserved;
  uint64_t device_bytes;    /* HBM held by the graph */
} pm_graph_info;

const char* pm_last_error(void);
const char* pm_version(void);

/* -- graphs --------------------------------------------------------------- */

/* Load <prefix>.meta.txt/.vertex.bin/.edge.bin[/.label.bin] and copy the CSR
 * to HBM.  Replaces GraphPreprocessor(prefix) + to_gpu()
 * (src/graph.h:34-41, :169-184, :228-254; test.cu:19-20,47). */
int pm_graph_open(const char* prefix, uint32_t flags, pm_graph** out);

/* Build a graph from host CSR arrays (copied).  labels may be NULL.  Same
 * layout as read_bin_file (src/graph.h:228-254). */
int pm_graph_from_csr(uint64_t n, uint64_t m, const uint64_t* rowptr, const uint32_t* colidx,
                      const uint32_t* labels, uint32_t flags, pm_graph** out);

/* Seeded synthetic generators; the result is symmetric, deduplicated,
 * self-loop free and sorted, the layout the reference's converter produces
 * (graph_converter/compactify.cpp:10-51).  No reference counterpart. */
int pm_graph_generate_rmat(int32_t scale, double edge_factor, double a, double b, double c,
                           uint64_t seed, uint32_t flags, pm_graph** out);
int pm_graph_generate_chunglu(uint64_t n, double gamma, double dmax, double avg, uint64_t seed,
                              uint32_t flags, pm_graph** out);
/* uniform labels 0..nlabels-1 (nlabels <= 32); uploads them when on device */
int pm_graph_set_uniform_labels(pm_graph* g, uint32_t nlabels, uint64_t seed);

/* Write the graph in the reference's on-disk format (the files
 * read_bin_file reads, src/graph.h:228-254). */
int pm_graph_write(const pm_graph* g, const char* prefix);

int pm_graph_info_get(const pm_graph* g, pm_graph_info* info);
/* Borrow the host CSR arrays (valid until pm_graph_destroy). */
int pm_graph_host_csr(const pm_graph* g, const uint64_t** rowptr, const uint32_t** colidx,
                      const uint32_t** labels);
/* Copy the CSR to HBM (no-op when already there). */
int pm_graph_to_device(pm_graph* g);
void pm_graph_destroy(pm_graph* g);

/* -- patterns ------------------------------------------------------------- */

/* Pattern modes (the `labeled` argument below is a set of these bits). */
#define PM_LABELED 0x1        /* LABELED=true, src/config.h:14 */
#define PM_VERTEX_INDUCED 0x2 /* EDGE_INDUCED=false, src/config.h:15: pattern
                                 non-edges map to non-adjacent vertices
                                 (src/gpu_match.cu:466-470,570-587,734-738);
                                 rows must be resident in HBM */

/* Parse a gSpan-like .g file and plan it.  PM_LABELED (1) uses the vertex
 * labels of the file (LABELED=true, src/config.h:14); without it every
 * label is 1 and embeddings are counted with symmetry breaking.  Adding
 * PM_VERTEX_INDUCED counts vertex-induced matches instead of edge-induced
 * ones.  Replaces PatternPreprocessor(path) (src/pattern.h:48-58). */
int pm_pattern_load(const char* path, int32_t labeled, pm_pattern** out);
/* Same from the text of a .g file. */
int pm_pattern_parse(const char* text, int32_t labeled, pm_pattern** out);

typedef struct {
  int32_t k;
  int32_t labeled;
  int64_t automorphisms;  /* PatternMultiplicity, src/pattern.h:236 */
  int32_t rc_order;       /* unlabeled job filter r > c, src/job_queue.h:44 */
  int32_t nslots;
  int32_t degree[PM_MAX_PATTERN];
  int32_t label[PM_MAX_PATTERN];
  uint32_t sym_mask[PM_MAX_PATTERN];  /* phi(u) < phi(j) for bit j */
  uint32_t nbr_mask[PM_MAX_PATTERN];  /* earlier neighbours of u */
  int32_t vertex_induced;             /* PM_VERTEX_INDUCED was given */
} pm_pattern_info;

int pm_pattern_info_get(const pm_pattern* p, pm_pattern_info* info);
void pm_pattern_destroy(pm_pattern* p);

/* -- matching ------------------------------------------------------------- */

void pm_opts_default(pm_opts* o);

/* Count the matches of p in g on the GPU.  Replaces the timed region of
 * test.cu:171-200 (job generation src/job_queue.h:39-125 runs on the device
 * inside the launch). */
int pm_match(pm_graph* g, const pm_pattern* p, const pm_opts* opts, pm_result* res);

/* Count n patterns of 2..PM_MAX_PATTERN vertices, labeled and unlabeled
 * mixed, in ONE persistent launch: the job space is (pattern, edge chunk) pairs on
 * one cursor, range tasks and deferred pieces carry their pattern, and equal
 * pendant subtrees share one table.  results[i].count/jobs are pattern i's
 * and equal a separate pm_match of p[i]; the timing and traffic fields of
 * every result describe the whole launch.  Replaces running test.cu once per
 * pattern file (plot/plot_fig11.sh:112-123).  opts->per_edge needs n == 1. */
int pm_match_batch(pm_graph* g, const pm_pattern* const* p, int32_t n, const pm_opts* opts,
                   pm_result* results);

/* -- learned scheduling and paging ----------------------------------------
 * A 16 -> 64 -> 64 -> 2 MLP over per-vertex features.  Output 0 predicts
 * log2(1 + work of the jobs rooted at the vertex), output 1 log2(1 + rounds
 * in which the vertex's package was used); their training targets are the
 * traces pm_opts.root_work / package_use record.  With pm_opts.model set,
 * PM_LEARN_ORDER takes phase-1 jobs heaviest predicted root first and
 * PM_LEARN_PAGING ranks packages for eviction (PM_POLICY_LEARNED) and
 * prefetch.  No counterpart in the reference (CSR-order job queue,
 * src/job_queue.h:39-89; LRU buffer, src/buffer.cu:581-796); counts never
 * depend on the model. */
#define PM_FEATURES 16
#define PM_MODEL_PARAMS 5378 /* W1[16][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2] */

typedef struct pm_model pm_model;

typedef struct {
  int32_t epochs;  /* passes over the samples (0 = 20) */
  int32_t batch;   /* samples per Adam step (0 = 4096) */
  float lr;        /* Adam step size (0 = 3e-3) */
  int32_t pad;
  uint64_t seed;   /* sample order (0 = 1) */
} pm_train_opts;

/* He-uniform initial weights from `seed` (FP32 master copy on the host). */
int pm_model_create(uint64_t seed, pm_model** out);
void pm_model_destroy(pm_model* m);
/* copy the PM_MODEL_PARAMS FP32 parameters out / in (layout above,
 * row-major [in][out]) */
int pm_model_get(const pm_model* m, float* params);
int pm_model_set(pm_model* m, const float* params);
/* features of every vertex of g on the GPU: out = host float[n][16] */
int pm_graph_features(pm_graph* g, float* out);
/* inference with BF16 weights and inputs on MFMA, FP32 accumulation:
 * X float[n][16] -> Y float[n][2] */
int pm_model_predict(const pm_model* m, const float* X, uint64_t n, float* Y);
/* FP32 loss sum_{s,o} W[s][o] (yhat - y)^2 / sum W and its gradient with
 * respect to the parameters (W NULL: all ones) */
int pm_model_grad(const pm_model* m, const float* X, const float* Y, const float* W, uint64_t n,
                  float* grad, double* loss);
/* Adam on the GPU (FP32 forward/backward kernels) */
int pm_model_train(pm_model* m, const float* X, const float* Y, const float* W, uint64_t n,
                   const pm_train_opts* o, double* final_loss);

/* -- autotuning ------------------------------------------------------------ */

typedef struct {
  int32_t sample;      /* job sample for the timed trials (0 = about 1M edges) */
  int32_t max_orders;  /* labeled patterns: core orders timed per pattern (0 = 4) */
  int32_t repeats;     /* launches per trial, fastest counts (0 = 2) */
} pm_tune_opts;

typedef struct {
  double base_ms;   /* sampled batch time with the caller's options */
  double best_ms;   /* sampled batch time with the chosen options */
  int32_t trials;   /* timed launches */
  int32_t reordered; /* patterns whose measured-fastest order is not the file order */
  int32_t sample;   /* sample used */
  int32_t pad;
} pm_tune_report;

/* Pick, by timing sampled launches on g, (1) for each labeled pattern the
 * order of its core vertices 2.. (file order and the planner's cheapest
 * candidates; remembered in the pattern handle for this graph) and (2) the
 * launch granularities of the whole batch: chunk (jobs per claim, the steal
 * granularity), split (range-task size) and hub_factor (fused-tail strategy
 * switch).  *best = *base with those fields set.  Counts never depend on any
 * of them; every trial's counts are checked against the first and a
 * mismatch fails with PM_ERR_DEVICE.  No counterpart in the reference, whose
 * granularities are compile-time constants (src/config.h:9-25). */
int pm_autotune(pm_graph* g, pm_pattern* const* p, int32_t n, const pm_opts* base,
                const pm_tune_opts* t, pm_opts* best, pm_tune_report* report);

#ifdef __cplusplus
}
#endif

#endif /* PUMATCH_H */
