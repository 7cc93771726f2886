/*
 * graphops_amd.h -- C ABI of the MI355X (gfx950) PageRank family.
 *
 * Drop-in boundary for graphops v0.4.1's PageRank operators.  The reference
 * has no FFI of its own; each entry point below is what a thin Rust
 * `extern "C"` crate binds in place of the generic Rust function cited beside
 * it (file:line in the reference crate).  Plain pointers and sizes only; no
 * C++ exception or abort crosses this boundary.
 *
 * Conventions (mirroring the reference, pagerank.rs / ppr.rs / lib.rs):
 *   - Unchecked `*_run` calls never reject inputs; they compute on what they
 *     are given and reproduce the reference's edge cases (n == 0 returns
 *     {iterations 0, diff 0.0, converged}; max_iterations == 0 returns the
 *     initial vector with {0, +inf, not converged}).
 *   - `*_checked_run` calls validate first and return GOPS_INVALID_PARAMETER
 *     with the reference's message text, readable via gops_last_error().
 *   - Scores are written to a caller-owned double[n] (the reference returns an
 *     owned Vec<f64> of length n), indexed by the caller's node ids.  A NULL
 *     scores pointer leaves the result on the device (for timing).
 *   - A context owns one HIP stream and all device memory it allocates; it is
 *     not thread-safe.  Separate contexts are independent; nothing is global.
 */
#ifndef GRAPHOPS_AMD_H
#define GRAPHOPS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The first four mirror graphops::Error (lib.rs:108-123):
 * InvalidParameter(String) lib.rs:117-118, IndexOutOfBounds(usize)
 * lib.rs:113-114, DimensionMismatch(usize, usize) lib.rs:121-122. */
typedef enum gops_status {
    GOPS_OK = 0,
    GOPS_INVALID_PARAMETER = 1,
    GOPS_INDEX_OUT_OF_BOUNDS = 2,
    GOPS_DIMENSION_MISMATCH = 3,
    GOPS_DEVICE_ERROR = 4,
    GOPS_OUT_OF_MEMORY = 5,
    GOPS_NOT_SUPPORTED = 6
} gops_status;

typedef struct gops_context gops_context;
typedef struct gops_graph gops_graph;
typedef struct gops_edges gops_edges;
typedef struct gops_server gops_server;

/* PageRankConfig, pagerank.rs:37-49, field for field (usize -> uint64_t).
 * Default (pagerank.rs:51-59): {0.85, 100, 1e-6}. */
typedef struct gops_pagerank_config {
    double damping;
    uint64_t max_iterations;
    double tolerance;
} gops_pagerank_config;

/* PageRankRun, pagerank.rs:14-29, without `scores` (written to the caller's
 * buffer).  iterations == max_iterations when converged == 0. */
typedef struct gops_run_info {
    uint64_t iterations;
    double diff_l1;
    int32_t converged;
    int32_t reserved;
} gops_run_info;

/* Device storage of edge weights (the caller always passes f64 weights, the
 * type WeightedGraph::edge_weight returns, graph.rs:67-72). */
enum {
    GOPS_WEIGHT_NONE = 0, /* unweighted graph (edge weights ignored / absent) */
    GOPS_WEIGHT_F64 = 1,
    GOPS_WEIGHT_F32 = 2,  /* rounded to nearest f32 on upload */
    GOPS_WEIGHT_BF16 = 3  /* rounded to f32, then to nearest-even bf16 */
};

/* Context options (gops_context_set_option).  No reference counterpart:
 * the reference is f64 throughout. */
enum {
    GOPS_OPT_PRECISION = 1,     /* GOPS_PRECISION_MIXED (default) or _FP64 */
    GOPS_OPT_ANCHOR_PERIOD = 2, /* mixed: at most this many sweeps between f64 anchor
                                   sweeps (default 32) */
    GOPS_OPT_NORMALIZE = 3,     /* 1: rescale the result to sum to 1 in f64 (default 0,
                                   the reference never renormalises, lib.rs:11-12) */
    GOPS_OPT_TIMING = 4,        /* 1: time every sweep kernel with HIP events */
    GOPS_OPT_PPR_TILE = 5,      /* batched PPR: queries per tile, 8/16/32/64 (0 = the
                                   autotuned width, see gops_ppr_tile_info; -1 = one
                                   query per device run) */
    GOPS_OPT_ANCHOR_DROP = 6,   /* mixed, global PageRank: re-anchor in f64 once the
                                   residual has dropped by this factor since the last
                                   anchor (default 1e-2) */
    GOPS_OPT_F32_BELOW = 7,     /* mixed, global PageRank: f64 sweeps while the residual
                                   is above this times the tolerance (default 0: none) */
    GOPS_OPT_PPR_F32_BELOW = 8, /* mixed, personalized, one query per device run: same
                                   (default 1e4) */
    GOPS_OPT_PPR_ANCHOR_PERIOD = 9, /* mixed, personalized: f32 sweeps between f64
                                   anchors while the residual is above 1e3 times
                                   the tolerance (default 64) */
    GOPS_OPT_GRAPH_BUILD = 10,  /* where gops_graph_create builds the CSR of in-edges:
                                   GOPS_BUILD_DEVICE (default) or GOPS_BUILD_HOST; the
                                   arrays, and so every result, are the same */
    GOPS_OPT_TILE_F32_BELOW = 11, /* mixed, batched and served personalized queries (tiles):
                                   f64 sweeps while the residual is above this times the
                                   tolerance (default 0: f32 delta sweeps from the first,
                                   each row carrying its f32 rounding into its next
                                   delta, which keeps L1 near 1e-14 of the f64 reference;
                                   the f64 phase only adds time) */
    GOPS_OPT_LAUNCH = 13,       /* how a run's sweeps are queued: GOPS_LAUNCH_AUTO (default:
                                   queued; measured no slower than the replays on every
                                   benchmark graph), _QUEUED (the sweeps' kernels
                                   launched one by one, a
                                   few ahead of the last finished sweep) or _GRAPH (a
                                   HIP graph of a few sweeps, captured once per graph
                                   and kind of run, replayed until the device says
                                   stop); the results are the same bits */
    GOPS_OPT_DELTA_SWEEP = 12   /* mixed, one vector per run: how f32 delta sweeps read
                                   the graph: GOPS_DELTA_SWEEP_AUTO (default: blocked
                                   from 2^20 nodes, where the f32 source vector outgrows
                                   the L2; pull below), _PULL (gathers over the CSR of
                                   in-edges) or _BLOCKED (a scatter pass per chunk of
                                   sources that sums each row's run of edges, and an
                                   accumulate pass per bin of rows, both streaming,
                                   integer row sums; its layout is built on the graph's
                                   first such run) */,
    GOPS_OPT_ANCHOR_SWEEP = 14, /* mixed, one vector per run, with blocked delta sweeps:
                                   how the f64 anchor sweeps read the graph:
                                   GOPS_ANCHOR_SWEEP_AUTO (default: blocked as well,
                                   f64 run sums added as 128-bit fixed point) or _PULL
                                   (gathers of the f64 iterate over the CSR) */
    GOPS_OPT_BLOCKED_BIN_RUNS = 15 /* blocked layout, read when a graph's layout is
                                   built (its first blocked run): the accumulate pass's
                                   bins of at most 4096 consecutive rows also end before
                                   a row whose bound on its runs would take the bin
                                   past this many (default 131072; 0: bins of 4096 rows
                                   each).  Bins of more than 65536 runs, one
                                   workgroup's share, are split over several
                                   workgroups and combined by a further kernel, so at
                                   65536 none is split.  The results differ only by
                                   f32 rounding in the run sums */,
    GOPS_OPT_CENTRALITY_SWEEP = 16 /* how Katz, eigenvector and HITS sweeps read the
                                   graph: GOPS_CENTRALITY_SWEEP_AUTO (default: blocked
                                   from 2^21 nodes, pull below), _PULL (gathers over the
                                   CSR of in-edges or of out-edges) or _BLOCKED (blocked
                                   layouts' scatter and accumulate passes, f64 run sums
                                   added as 128-bit fixed point; the in-edge layout is
                                   shared with PageRank's blocked sweeps, the out-edge
                                   one is built for eigenvector and HITS on first use) */
};
enum { GOPS_CENTRALITY_SWEEP_AUTO = 0, GOPS_CENTRALITY_SWEEP_PULL = 1, GOPS_CENTRALITY_SWEEP_BLOCKED = 2 };
enum { GOPS_DELTA_SWEEP_AUTO = 0, GOPS_DELTA_SWEEP_PULL = 1, GOPS_DELTA_SWEEP_BLOCKED = 2 };
enum { GOPS_ANCHOR_SWEEP_AUTO = 0, GOPS_ANCHOR_SWEEP_PULL = 1 };
enum { GOPS_LAUNCH_AUTO = 0, GOPS_LAUNCH_QUEUED = 1, GOPS_LAUNCH_GRAPH = 2 };
enum { GOPS_PRECISION_MIXED = 0, GOPS_PRECISION_FP64 = 1 };
enum { GOPS_BUILD_DEVICE = 0, GOPS_BUILD_HOST = 1 };

/* Statistics of the most recent run on a context (no reference counterpart). */
typedef struct gops_run_stats {
    uint64_t n;             /* nodes */
    uint64_t nnz;           /* edges (adapter neighbour entries) */
    uint64_t sweeps_fp32;   /* f32 delta sweeps executed */
    uint64_t sweeps_fp64;   /* f64 anchor sweeps executed */
    double device_ms;       /* first sweep launch .. result ready, HIP events */
    double fp32_sweep_ms;   /* summed duration of f32 sweep kernels (GOPS_OPT_TIMING) */
    double fp64_sweep_ms;   /* summed duration of f64 sweep kernels (GOPS_OPT_TIMING) */
    double other_ms;        /* reserved */
    uint64_t sweeps_blocked; /* sweeps that ran in the blocked layout (f32 deltas, and
                                f64 anchors unless GOPS_ANCHOR_SWEEP_PULL) */
    int32_t blocked_fallback; /* 1: GOPS_DELTA_SWEEP_AUTO chose the blocked layout but it
                                 did not fit in free device memory, so the f32 delta
                                 sweeps ran as pulls (retried on later runs) */
    int32_t graph_replay;   /* 1: the sweeps ran as replays of a captured HIP graph */
    int32_t blocked_bins;   /* blocked runs: the layout's bins of rows (accumulate pass) */
    int32_t blocked_split_bins; /* of those, bins split over several workgroups */
} gops_run_stats;

/* ---- context ---------------------------------------------------------- */
gops_status gops_context_create(int device, gops_context **out);
/* Releases the caller's handle.  Graphs made on the context keep it (its
 * stream and memory) until the last of them is destroyed, so handles may be
 * destroyed in any order; the context handle must not be used afterwards. */
void gops_context_destroy(gops_context *ctx);
/* Message of the last failing call on ctx ("" if none).  For
 * GOPS_INVALID_PARAMETER it is Error's Display text, e.g.
 * "invalid parameter: damping must be in [0,1]" (lib.rs:117). */
const char *gops_last_error(const gops_context *ctx);
gops_status gops_context_set_option(gops_context *ctx, int option, double value);
gops_status gops_last_run_stats(const gops_context *ctx, gops_run_stats *out);

/* ---- graphs ----------------------------------------------------------- */
/* Graph from adapter output: for u in 0..n, targets[offsets[u]..offsets[u+1]]
 * are Graph::neighbors(u) (graph.rs:26-39) in adapter order, duplicates kept
 * (they count toward out_degree, pagerank.rs:137); edge_weights (nullable) are
 * WeightedGraph::edge_weight(u, v) per entry (graph.rs:67-72).  Replaces the
 * reference's per-call neighbour materialisation (pagerank.rs:136, 226-234;
 * ppr.rs:112, 133).  The lists are uploaded once and the CSR transpose is
 * built on the device (or on the host, GOPS_OPT_GRAPH_BUILD); the caller's
 * arrays may be freed on return.  Ids must be < n (GOPS_INDEX_OUT_OF_BOUNDS
 * otherwise); n and offsets[n] must be < 2^31. */
gops_status gops_graph_create(gops_context *ctx, uint64_t n, const uint64_t *offsets,
                              const uint32_t *targets, const double *edge_weights,
                              int weight_storage, gops_graph **out);
/* AdjacencyMatrix adapter (graph.rs:122-142): row-major n*n matrix; edge
 * u->v iff matrix[u*n+v] > 0.0 (NaN and negatives are not edges, +inf and
 * self-loops are); edge weight = the raw entry. */
gops_status gops_graph_from_dense(gops_context *ctx, uint64_t n, const double *matrix,
                                  int weight_storage, gops_graph **out);
/* Releases the caller's handle; a server on the graph keeps it until the
 * server is destroyed. */
void gops_graph_destroy(gops_graph *g);
gops_status gops_graph_info(const gops_graph *g, uint64_t *n, uint64_t *nnz);
/* Host copies of the preprocessed graph (no reference counterpart: the
 * reference keeps no preprocessed form).  Row r is the node order[r]; nodes
 * are ordered by in-degree, descending, ties by id.  Every pointer may be
 * null.  row_ptr: n+1 entries; col: m entries, the new ids of the row's
 * sources, ascending, duplicate edges in adapter order; weights: m entries
 * of the graph's weight storage (f64, f32 or bf16 as uint16; null for an
 * unweighted graph); order: n; inv_deg / inv_ws: n f64 values per row,
 * 1/out-degree and 1/out_wsum of the row's node (0 when zero). */
gops_status gops_graph_export(const gops_graph *g, int32_t *row_ptr, int32_t *col,
                              void *weights, uint32_t *order, double *inv_deg,
                              double *inv_ws);

/* ---- PageRank family -------------------------------------------------- */
/* pagerank_run, pagerank.rs:122-187 (pagerank, pagerank.rs:114-116, is this
 * call with the info ignored). */
gops_status gops_pagerank_run(gops_context *ctx, gops_graph *g,
                              const gops_pagerank_config *cfg, double *scores,
                              gops_run_info *info);
/* pagerank_checked_run, pagerank.rs:321-324 (and pagerank_checked,
 * pagerank.rs:91-94): PageRankConfig::validate (pagerank.rs:63-85) first. */
gops_status gops_pagerank_checked_run(gops_context *ctx, gops_graph *g,
                                      const gops_pagerank_config *cfg, double *scores,
                                      gops_run_info *info);
/* pagerank_weighted_run, pagerank.rs:210-289 (pagerank_weighted,
 * pagerank.rs:201-203).  Needs a graph created with edge weights. */
gops_status gops_pagerank_weighted_run(gops_context *ctx, gops_graph *g,
                                       const gops_pagerank_config *cfg, double *scores,
                                       gops_run_info *info);
/* pagerank_weighted_checked_run, pagerank.rs:328-336 (and
 * pagerank_weighted_checked, pagerank.rs:294-316): validate the config and
 * that every neighbour entry's weight is finite and >= 0, then run once. */
gops_status gops_pagerank_weighted_checked_run(gops_context *ctx, gops_graph *g,
                                               const gops_pagerank_config *cfg,
                                               double *scores, gops_run_info *info);
/* personalized_pagerank_run, ppr.rs:90-156 (personalized_pagerank,
 * ppr.rs:66-72).  personalization has len entries (the reference indexes it
 * for i < n, ppr.rs:127; a shorter vector is GOPS_DIMENSION_MISMATCH here
 * instead of a panic). */
gops_status gops_personalized_pagerank_run(gops_context *ctx, gops_graph *g,
                                           const gops_pagerank_config *cfg,
                                           const double *personalization, uint64_t len,
                                           double *scores, gops_run_info *info);
/* personalized_pagerank_checked_run, ppr.rs:77-85 (and
 * personalized_pagerank_checked, ppr.rs:11-44). */
gops_status gops_personalized_pagerank_checked_run(gops_context *ctx, gops_graph *g,
                                                   const gops_pagerank_config *cfg,
                                                   const double *personalization,
                                                   uint64_t len, double *scores,
                                                   gops_run_info *info);

/* Calibration (no reference counterpart): the best of reps timings (HIP
 * events, ms) of a kernel that makes only the f32 sweep's gathers on g --
 * every edge's source value in stored order, the col stream, a sum -- i.e.
 * the bound that the source-gather pattern of this graph sets for a sweep. */
gops_status gops_gather_probe(gops_context *ctx, gops_graph *g, uint32_t reps, double *best_ms);

/* ---- top_k ----------------------------------------------------------- */
/* top_k, topk.rs:17-42: the (id, score) pairs of the k largest finite,
 * positive scores, descending, computed on the device; *count (<= k)
 * entries are written to ids / values (each nullable, k entries).  When
 * scores tie at the k-th value, the entries kept are the ones the
 * reference's heap keeps; equal scores come out by ascending id (the
 * reference's unstable sort leaves their order open). */
gops_status gops_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                       uint32_t *ids, double *values, uint32_t *count);
/* top_k of the last pagerank / weighted / personalized run on g, whose
 * scores stay on the device (call the run with scores = NULL): only the k
 * pairs cross PCIe.  No reference counterpart (it composes the reference's
 * run and top_k). */
gops_status gops_result_top_k(gops_context *ctx, gops_graph *g, uint32_t k, uint32_t *ids,
                              double *values, uint32_t *count);
/* gops_result_top_k without allocating.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results, bit for bit; k is at most 4096
 * (GOPS_INVALID_PARAMETER beyond).  The selection's scratch is one
 * file-scope device array of about 400 KB, which every GPU the library runs
 * on holds until the process ends; every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (from several contexts or threads)
 * race. */
gops_status gops_result_top_k_noalloc(gops_context *ctx, gops_graph *g, uint32_t k,
                                      uint32_t *ids, double *values, uint32_t *count);

/* ---- Katz, eigenvector and HITS centrality ----------------------------- */
/* KatzConfig, katz.rs:25-53 (defaults alpha 0.1, beta 1.0, 100, 1e-6). */
typedef struct gops_katz_config {
    double alpha;
    double beta;
    uint64_t max_iterations;
    double tolerance;
} gops_katz_config;
/* katz_centrality_run, katz.rs:136-201 (katz_centrality, katz.rs:131-133, is
 * this call with the info ignored): x <- alpha A^T x + beta from x = beta,
 * stop on sum |x - x'| < tolerance.  KatzRun's fields are gops_run_info's.
 * The graph's edge weights, if any, are not used. */
gops_status gops_katz_centrality_run(gops_context *ctx, gops_graph *g,
                                     const gops_katz_config *cfg, double *scores,
                                     gops_run_info *info);
/* katz_centrality_checked, katz.rs:207-210: KatzConfig::validate (katz.rs:57-84,
 * messages verbatim) first. */
gops_status gops_katz_centrality_checked_run(gops_context *ctx, gops_graph *g,
                                             const gops_katz_config *cfg, double *scores,
                                             gops_run_info *info);
/* eigenvector_centrality_run, eigenvector.rs:49-125 (eigenvector_centrality,
 * eigenvector.rs:26-28, is this call with 100, 1e-6): x <- (A + I) x,
 * L2-normalised with a non-negative sum, from x = 1/sqrt(n).  (A x)[v] sums
 * over v's own neighbour list; the out-edge CSR it needs is built on the
 * device at the first call and kept with the graph. */
gops_status gops_eigenvector_centrality_run(gops_context *ctx, gops_graph *g,
                                            uint64_t max_iterations, double tolerance,
                                            double *scores, gops_run_info *info);
/* hits, centrality.rs:114-166: hub and authority scores (each nullable),
 * L2-normalised after every half step, from 1/sqrt(n); stops when both L2
 * changes are below tol.  The reference returns only the two vectors; info
 * (nullable) adds the iteration count, the last hub change (diff_l1) and
 * whether both changes were below tol. */
gops_status gops_hits(gops_context *ctx, gops_graph *g, uint64_t max_iter, double tol,
                      double *hub, double *authority, gops_run_info *info);

/* Batched personalized PageRank: num_queries independent calls of
 * personalized_pagerank_run (ppr.rs:90-156) in one device run over the shared
 * graph.  Query q's personalization is sparse: seeds[seed_offsets[q] ..
 * seed_offsets[q+1]] with weights seed_weights[...] (NULL: all 1.0), i.e. the
 * dense vector the reference would take with zeros elsewhere.  cfgs[q] is
 * query q's config.  scores (nullable) receives num_queries rows of n
 * doubles; infos (nullable) receives num_queries infos. */
gops_status gops_ppr_batch_run(gops_context *ctx, gops_graph *g, uint32_t num_queries,
                               const gops_pagerank_config *cfgs,
                               const uint64_t *seed_offsets, const uint32_t *seeds,
                               const double *seed_weights, double *scores,
                               gops_run_info *infos);

/* The pull sweeps' row plan: rows of more than row_min in-edges take a
 * workgroup each, rows of more than hub_chunk are split into chunks of it,
 * and runs of shorter rows give each group of lanes at most row_iters rows.
 * Defaults 4096 / 65536 / 8. */
typedef struct gops_row_plan {
    int32_t row_min;
    int32_t hub_chunk;
    int32_t row_iters;
    int32_t candidates;  /* plans timed by gops_graph_tune_rows */
    double default_us;   /* f32 pull sweep with the default cut-offs (us) */
    double tuned_us;     /* f32 pull sweep with the chosen ones (us) */
} gops_row_plan;

/* Autotune g's row plan: time f32 pull sweeps (10-sweep runs, best of 3,
 * HIP events) under a coordinate search over row_min {4096, 2048, 1024, 512},
 * hub_chunk {65536, 32768, 16384} and row_iters {8, 4, 16}, and keep the
 * fastest (a change must gain 1 %) for every later pull sweep on g, f64
 * ones included.  Results change only by f64 rounding (the order in which
 * a row's products are added).  The context's options and last run
 * statistics are left as they were.  No reference counterpart (tuning of
 * the loop of pagerank.rs:154-162). */
gops_status gops_graph_tune_rows(gops_context *ctx, gops_graph *g, gops_row_plan *out);

/* Set g's row plan to the given cut-offs (row_min in [64, 4096], hub_chunk
 * in [row_min, 2^26], row_iters in [1, 64]; the other fields are ignored).
 * No reference counterpart. */
gops_status gops_graph_set_rows(gops_context *ctx, gops_graph *g, const gops_row_plan *plan);

/* Tile width a batch of num_queries queries runs with on g under ctx's
 * options (0: the one-query-per-run path), and the autotuner's measured cost
 * of widths 8, 16, 32 and 64 in ms per query-sweep (nullable, 4 doubles;
 * infinity where a width does not fit in device memory).  Runs the
 * autotuner's trial on g if it has not run yet.  No reference counterpart
 * (diagnostics of the batched path). */
gops_status gops_ppr_tile_info(gops_context *ctx, gops_graph *g, uint32_t num_queries,
                               uint32_t *tile, double *ms_per_query_sweep);

/* ---- serving loop (continuous batching of PPR queries) ---------------- */
typedef struct gops_server_config {
    uint32_t max_slots;      /* query slots resident on the device (0 = from HBM budget) */
    double hbm_budget_bytes; /* device bytes the slot pool may use (0 = 90% of free) */
    uint32_t tile;           /* queries per tile (0 = autotuned) */
    uint32_t topk;           /* >0: results are top-k (id, score) instead of n scores */
} gops_server_config;

typedef struct gops_server_result {
    uint64_t ticket;
    gops_run_info info;
    double submit_ms;  /* host clock, ms since server creation */
    double finish_ms;
} gops_server_result;

/* No reference counterpart (the reference is a library call per query). */
gops_status gops_server_create(gops_context *ctx, gops_graph *g,
                               const gops_server_config *cfg, gops_server **out);
void gops_server_destroy(gops_server *s);
/* Queue one personalized PageRank query; returns its ticket. */
gops_status gops_server_submit(gops_server *s, const gops_pagerank_config *cfg,
                               const uint32_t *seeds, uint64_t num_seeds,
                               const double *seed_weights, uint64_t *ticket);
/* Advance the device by up to `sweeps` sweeps, admitting queued queries into
 * free slots first; returns the number of active queries via *active. */
gops_status gops_server_step(gops_server *s, uint32_t sweeps, uint32_t *active);
/* Collect a finished query: *ready = 1 and out / scores (n doubles,
 * nullable) are filled, or *ready = 0 when nothing has finished.  A server
 * created with topk > 0 selects each result's top_k (topk.rs:17-42) on the
 * device when the query finishes and keeps only those pairs: scores must be
 * NULL, and topk_ids / topk_scores (topk entries each, nullable) receive
 * them, padded with UINT32_MAX / 0.0 when fewer scores are positive. */
gops_status gops_server_poll(gops_server *s, int *ready, gops_server_result *out,
                             double *scores, uint32_t *topk_ids, double *topk_scores);
gops_status gops_server_stats(const gops_server *s, uint64_t *queued, uint64_t *active,
                              uint64_t *finished, uint32_t *slots);

/* ---- synthetic inputs (no reference counterpart: the reference has no
 *      generator) -------------------------------------------------------- */
enum { GOPS_GEN_UNIT = 0, GOPS_GEN_UNIFORM = 1, GOPS_GEN_LOGNORMAL = 2 };
enum { GOPS_ROUND_F64 = 0, GOPS_ROUND_F32 = 1, GOPS_ROUND_BF16 = 2 };

/* R-MAT: 2^scale nodes, edge_factor * 2^scale drawn edges, quadrant
 * probabilities (a, b, c, 1-a-b-c); self-loops and duplicates removed;
 * neighbour lists sorted by target.  Weights per unique edge (u, v) are drawn
 * from a counter RNG keyed by (seed, u, v): GOPS_GEN_UNIFORM in [p0, p1),
 * GOPS_GEN_LOGNORMAL(mu = p0, sigma = p1), then rounded per `rounding`. */
gops_status gops_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                          uint64_t seed, int weight_kind, double p0, double p1,
                          int rounding, gops_edges **out);
/* Directed Erdos-Renyi G(n, p) over ordered pairs u != v, unit weights. */
gops_status gops_gen_erdos_renyi(uint64_t n, double p, uint64_t seed, gops_edges **out);
/* Seed sets for batched PPR: sizes uniform in [min_seeds, max_seeds]
 * (log-uniform if log_uniform != 0), distinct nodes drawn proportional to
 * out-degree (the source of a uniformly drawn edge). */
gops_status gops_gen_seed_sets(const gops_edges *e, uint32_t num_queries, uint32_t min_seeds,
                               uint32_t max_seeds, int log_uniform, uint64_t seed,
                               uint64_t *offsets /* num_queries+1 */, uint32_t *seeds,
                               uint64_t seeds_capacity);
gops_status gops_edges_view(const gops_edges *e, uint64_t *n, uint64_t *m,
                            const uint64_t **offsets, const uint32_t **targets,
                            const double **weights);
void gops_edges_destroy(gops_edges *e);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHOPS_AMD_H */
