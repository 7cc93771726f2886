// internal.h -- shared host/device declarations of libgraphops_amd.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/graphops_amd.h"

namespace gops {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64s)
constexpr int kStageNnz = 4096;      // edges of a run of rows staged in LDS (f32: 16 KB)
constexpr int kHubChunk = 65536;     // rows longer than this are split into chunks of it
constexpr int kRowIters = 8;         // row iterations per group in one run
constexpr int kUnroll = 16;          // independent edge loads in flight per lane

// Work descriptor of one workgroup in a sweep (rows sorted by in-degree):
//   ROWS : x = log2(lanes per row), y = first row, z = end row; the run has
//          at most kStageNnz edges and rows of at most kStageNnz edges
//   ROW  : x = kRowFlag, y = row; one row of (kStageNnz, kHubChunk] edges
//   HUB  : x = kHubFlag | chunk index, y = hub id, z = first edge, w = end
//          edge; one chunk of a row longer than kHubChunk
constexpr int32_t kHubFlag = 1 << 28;
constexpr int32_t kRowFlag = 1 << 27;
struct BlockDesc {
  int32_t x, y, z, w;
};

// Sweep modes, chosen on the device by the last workgroup of the previous
// sweep (see pagerank.hip).
enum : int32_t { kModeDelta = 0, kModeDeltaFold = 1, kModeAnchor = 2 };

// Scalars a sweep chain keeps in device memory.
struct RunState {
  double dsum_in;      // dangling sum of the delta the next f32 sweep gathers
  double mass_in;      // sum of that delta (0 in exact arithmetic; removed by the sweep)
  double dsum_x;       // dangling sum of the iterate the next anchor sweep gathers
  double diff;         // L1 residual of the last sweep
  double damping;
  double tolerance;
  double inv_n;        // 1.0 / n
  double n_f64;        // n as f64
  double diff_anchor;  // residual of the last anchor sweep
  double anchor_drop;  // re-anchor once diff <= anchor_drop * diff_anchor
  double x_scale;      // the next anchor gathers x = x_scale * xbar
  double f32_below;    // f32 sweeps only once diff <= f32_below * tolerance (0: at once)
  double anchor_min;   // periodic anchors only while diff > anchor_min * tolerance
  int32_t f64_phase;   // still in the leading all-f64 phase
  int32_t pad1;
  // Global runs in the blocked layout: the isolated rows (no in-edges, no
  // out-edges) all hold the same iterate, kept here instead of in xbar;
  // n_iso = 0 when their rows are swept like the others.
  double x_iso;
  double n_iso;
  uint64_t iter;
  uint64_t max_iter;
  uint64_t since_anchor;
  uint64_t anchor_period;  // re-anchor at the latest after this many sweeps
  int32_t done;
  int32_t converged;
  int32_t mode;        // mode of the next sweep
  int32_t all_f64;     // every sweep an f64 anchor sweep
  int32_t yd_cur;      // yd[yd_cur] holds the delta the next sweep gathers
  int32_t y64_cur;     // y64[y64_cur] holds the iterate the next anchor gathers
  uint32_t ticket;
  int32_t kind_miss;   // a sweep found only the other kind's kernels launched (run stopped)
  // host-mapped progress word written after every sweep:
  // (iterations << 4) | (mode of the next sweep << 2) | (converged << 1) | done
  uint64_t *mirror;
  uint8_t *sweep_log;  // mode of every sweep (capacity sweep_log_cap)
  uint64_t sweep_log_cap;
};

// Host copy of the preprocessed graph (new ids = rows sorted by in-degree,
// descending, ties by original id).
struct HostCSRT {
  uint64_t n = 0, m = 0;
  std::vector<int32_t> row_ptr;  // n+1
  std::vector<int32_t> col;      // m, new source ids, ascending per row
  std::vector<double> w;         // m, max(edge_weight, 0) (empty if unweighted)
  std::vector<uint32_t> order;   // new -> original id
  std::vector<uint32_t> rank;    // original -> new id
  std::vector<double> inv_deg;   // new ids: 1/out-degree, 0 for dangling
  std::vector<double> inv_ws;    // new ids: 1/out_wsum, 0 for ws == 0 (weighted only)
  bool has_weights = false;
  // first invalid neighbour weight in adapter order (pagerank.rs:300-314):
  // 0 none, 1 non-finite, 2 negative
  int first_bad_weight = 0;
};

// Cut-offs of the pull sweeps' row plan (plan_blocks): rows of more than
// row_min in-edges (<= kStageNnz, the stage) take a workgroup of their own,
// rows of more than hub_chunk are split into chunks of it, and a run of short
// rows gives each group of lanes at most row_iters rows.
struct PlanParams {
  int32_t row_min = kStageNnz;
  int32_t hub_chunk = kHubChunk;
  int32_t row_iters = kRowIters;
};

struct Plan {
  std::vector<BlockDesc> desc;
  std::vector<int32_t> hub_row, hub_nchunks, hub_first;
  int num_hubs = 0;
};

// Device-resident graph.
struct DevGraph {
  uint64_t n = 0, m = 0;
  int weight_storage = GOPS_WEIGHT_NONE;
  int32_t *row_ptr = nullptr;
  int32_t *col = nullptr;
  void *w = nullptr;             // f64 / f32 / bf16 (uint16) per weight_storage
  float *inv_deg32 = nullptr;
  double *inv_deg64 = nullptr;
  float *inv_ws32 = nullptr;
  double *inv_ws64 = nullptr;
  uint32_t *rank = nullptr;      // original -> new
  uint32_t *order = nullptr;     // new -> original
  BlockDesc *desc = nullptr;
  int num_desc = 0;
  int32_t *hub_row = nullptr, *hub_nchunks = nullptr, *hub_first = nullptr;
  int num_hubs = 0;
  int num_hub_chunks = 0;
};

// Per-graph scratch of a single-vector run (allocated on first use).
struct RunBuffers {
  uint64_t n = 0;
  double *xbar = nullptr;        // the iterate x (f64)
  float *yd[2] = {nullptr, nullptr};   // prescaled deltas d*delta/deg (f32)
  double *y64[2] = {nullptr, nullptr}; // prescaled iterate d*x/deg (f64)
  float *q32 = nullptr;          // personalization (new ids), f32
  double *q64 = nullptr;         // personalization (new ids), f64
  double *out = nullptr;         // result in original ids
  double *partials = nullptr;    // per-workgroup partial sums
  double *hub_partial = nullptr; // per hub chunk partial sums
  RunState *state = nullptr;
  uint8_t *sweep_log = nullptr;
  size_t partial_slots = 0;
  bool has_result = false;       // out holds the last run's scores
  // captured chains of kChainSweeps sweeps, one per kind of run
  // (weighted, personalized, blocked: the bits of the index)
  hipGraphExec_t chain[8] = {};
};

// State of a Katz / eigenvector / HITS power iteration in device memory.
struct PowerState {
  double alpha, beta;   // Katz
  double tolerance;
  double norm, sign;    // normalisation of the last sweep's vector (eigenvector, HITS)
  double diff;          // residual of the last iteration (HITS: the hub change)
  double diff2;         // HITS: the authority change of the last iteration
  uint64_t iter, max_iter;
  int32_t done, converged;
  int32_t par;          // Katz: which prescaled buffer the next sweep gathers
  int32_t fix;          // blocked Katz: the next sweep's row sums in units of 2^-fix
  // host-mapped progress word: (iterations << 2) | (converged << 1) | done
  uint64_t *mirror;
};

// Per-graph buffers of the power iterations (allocated on first use).
struct PowerBuffers {
  uint64_t n = 0;
  double *v[4] = {nullptr, nullptr, nullptr, nullptr};  // iterates, prescaled copies, row sums;
                                                        // v[3]: results in caller ids
  double *partials = nullptr;
  double *hub_partial = nullptr;
  size_t hub_cap = 0;
  PowerState *state = nullptr;
  bool katz_blocked = false;     // the current Katz run sweeps the blocked layout
  bool power_blocked = false;    // the current eigenvector / HITS run sweeps blocked layouts
  uint64_t outdeg_max = 0;       // out-degree of the out-edge CSR's row 0, the largest
  uint64_t n_in = 0;             // rows with in-edges (rows are sorted by in-degree)
  uint64_t indeg_max = 0;        // in-degree of row 0, the largest
  bool have_degrees = false;     // n_in and indeg_max found
  double katz_bound_scale = 0.0; // alpha * indeg_max of the current Katz run
};

// Layout of the blocked f32 delta sweep (blocked.hip, pagerank.hip).  The
// sweep runs as two streaming passes instead of a gather pull:
//   scatter     per chunk of 2^kPbChunkShift sources, whose f32 deltas sit in
//               LDS: the chunk's edges are walked in (row, source) order and
//               the products of each run (consecutive edges of one row) are
//               summed; one f32 per run goes to its place in rsum
//   accumulate  per bin of at most 2^kPbBinShift consecutive rows, whose sums
//               sit in LDS as fixed-point int64: the bin's runs are read in
//               order and added.  A bin ends at 2^kPbBinShift rows or where
//               the next row's bound on its runs would take the bin past a
//               cap (gops_context::pb_bin_runs), so rows of many runs get
//               bins of fewer rows; a bin of more than kPbItem2 runs is split
//               over several workgroups and combined after
// Entries are the edges of the CSR of in-edges in chunk-major order (chunk of
// the source, row, source), padded: every segment (the entries of one chunk
// into one bin) starts at a multiple of kPbLane entries, and every chunk at a
// multiple of kPbWin.  A padding entry repeats the source of the entry before
// it with weight 0 and continues its run.  A wave takes a window of kPbWin
// entries, kPbLane consecutive entries per lane.  Runs are maximal sequences of
// entries of one row within one window; they are numbered in chunk-major order
// and stored in bin-major order (bin of the row, chunk, row).  The runs of one
// segment keep their relative order, so a run's bin-major place is its
// chunk-major number plus its segment's offset.
constexpr int kPbChunkShift = 14;
constexpr int kPbBinShift = 12;
constexpr int kPbLane = 8;                  // consecutive entries per lane
constexpr int kPbWin = 64 * kPbLane;        // entries per window (one wave's)
constexpr int kPbItem1 = 64;               // windows per scatter workgroup, at most
constexpr int kPbItem2 = 1 << 16;           // runs per accumulate workgroup, at most
constexpr uint16_t kPbSrcMask = (1u << kPbChunkShift) - 1;
constexpr uint16_t kPbRunHead = 1u << 14;   // entry starts a run
constexpr uint16_t kPbSegHead = 1u << 15;   // entry starts a segment (and a run)
// GOPS_DELTA_SWEEP_AUTO: blocked from here (f32 sweep, bf16 weights, R-MAT:
// scale 20 pull 105 us / blocked 101 us, scale 21 194 / 119, 22 357 / 172;
// C2, scale 20 with f32 weights, whole runs: pull 2.73 ms, blocked 2.21 ms)
constexpr uint64_t kPbAutoNodes = 1ull << 20;
// GOPS_CENTRALITY_SWEEP_AUTO: blocked Katz, eigenvector and HITS from here
// (measured at scale 24 only; not below 2^21)
constexpr uint64_t kPowerAutoNodes = 1ull << 21;

struct PbItem {
  int32_t x, y, z, w;
};

struct BlockedLayout {
  uint64_t m = 0;              // entries, padding included (a multiple of kPbWin)
  uint64_t num_runs = 0;
  uint64_t num_segs = 0;
  uint16_t *src = nullptr;     // source id - chunk base | kPbRunHead | kPbSegHead
  void *w = nullptr;           // weights in the graph's storage type (weighted graphs)
  uint8_t *mask = nullptr;     // 1 per edge, 0 per padding entry (the unweighted runs'
                               // weights)
  int32_t *win = nullptr;      // per window: the number of the run its first entry
                               // starts, and that entry's segment
  int32_t *seg_off = nullptr;  // per segment: bin-major place - chunk-major run number
  uint16_t *rrow = nullptr;    // bin-major runs: row id - bin base
  float *rsum = nullptr;       // bin-major runs: one sweep's run sums (8 bytes a run)
  double *rsum64 = nullptr;    // the same memory: an anchor sweep's run sums
  PbItem *items1 = nullptr;    // scatter: {chunk, first window, end window, 0}
  PbItem *items2 = nullptr;    // accumulate: {bin, first run, end run, slot or -1}
  PbItem *splits = nullptr;    // bins of several items: {bin, first slot, slots, first row}
  int32_t *bin_row0 = nullptr; // [bins + 1] first row of every bin, then n
  long long *bin_partial = nullptr;  // [slot][1 << kPbBinShift] fixed-point row sums, or
                                     // [slot][1 << kPbBinShift][2] (128-bit, anchors)
  uint8_t *bin_flags = nullptr;      // [slot][1 << kPbBinShift] non-finite run sums seen
  double *partials = nullptr;  // [kReduce][items2 + splits] sums of the epilogues
  int num_items1 = 0, num_items2 = 0, num_splits = 0, num_bins = 0, num_split_bins = 0;
  uint64_t n_iso = 0;  // isolated rows (no in-edges, no out-edges): the tail [n - n_iso, n)
};

constexpr uint64_t kSweepLogCap = 1 << 16;
#ifndef GOPS_TUNE_CHAIN
#define GOPS_TUNE_CHAIN 4
#endif
#ifndef GOPS_TUNE_CHAIN_AHEAD
#define GOPS_TUNE_CHAIN_AHEAD 1
#endif
constexpr int kChainSweeps = GOPS_TUNE_CHAIN;       // sweeps per captured chain
constexpr int kChainAhead = GOPS_TUNE_CHAIN_AHEAD;  // chains queued beyond the last seen sweep
constexpr size_t kMaxReduce = 8;  // partial sums per workgroup, at most
constexpr int kMaxSweepBlocks = 2048;  // grid of the sweep kernel (grid-stride over descriptors)

}  // namespace gops

namespace gops {
struct PprGraphCache;
struct TopkScratch;
struct TopkBatchScratch;
}

struct gops_graph {
  gops_context *ctx = nullptr;
  // the caller's handle, plus one per live server on the graph; atomic because
  // finalizers may release handles from any thread
  std::atomic<int> refs{1};
  gops::PprGraphCache *ppr_cache = nullptr;  // batched PPR plans (server.cpp)
  gops::DevGraph dev;
  gops::DevGraph fwd;      // out-edge CSR in new ids (row_ptr, col, plan only), built on demand
  bool has_fwd = false;
  gops::BlockedLayout pb;  // blocked delta sweep, built on demand
  bool has_pb = false;
  gops::BlockedLayout pb_fwd;  // blocked layout of fwd (eigenvector, HITS), built on demand
  bool has_pb_fwd = false;
  gops::RunBuffers run;
  gops::PowerBuffers pow;
  bool weighted = false;   // created with edge weights
  int first_bad_weight = 0;
  std::vector<uint32_t> order_host;    // new -> original id
  std::vector<uint32_t> rank_host;     // original -> new id
  gops::PlanParams plan_params;        // cut-offs of dev's row plan
};

namespace gops {
constexpr int kPinChunks = 3;  // pinned chunks of the staged host <-> device copies
}

struct gops_context {
  int device = 0;
  std::atomic<int> refs{1};  // the caller's handle, plus one per live graph made on it
  hipStream_t stream = nullptr;
  std::string last_error;
  int precision = GOPS_PRECISION_MIXED;
  int anchor_period = 32;
  // re-anchor every 1e2 of residual: the f32 chain's rounding since the last
  // anchor stays below 1e-4 of the residual it stops at on every graph tested
  // (1e-3 saves an anchor at R-MAT scale 24 but reaches 1.2e-4 on small
  // graphs with slow contraction)
  double anchor_drop = 1e-2;
  double f32_below = 0.0;          // global PageRank: f32 sweeps from the start
  double ppr_f32_below = 1e4;      // personalized: f64 until diff <= 1e4 * tolerance
  double tile_f32_below = 0.0;     // batched / served personalized (tiles): f32 with carried rounding
  double ppr_anchor_min = 1e3;     // personalized: no anchor once diff <= 1e3 * tolerance
  int ppr_anchor_period = 64;
  int normalize = 0;
  int timing = 0;
  int ppr_tile = 0;
  int graph_build = GOPS_BUILD_DEVICE;
  int delta_sweep = GOPS_DELTA_SWEEP_AUTO;
  int anchor_sweep = GOPS_ANCHOR_SWEEP_AUTO;
  // run bound that ends a blocked layout's bin (0: 2^12 rows each); bins of up
  // to twice one accumulate workgroup's runs measured best at R-MAT scale 24
  // (f32 sweep 491 us, anchor 730 us; kPbItem2: 498 / 731 us; 0: 491 / 740 us)
  int pb_bin_runs = 2 * gops::kPbItem2;
  int centrality_sweep = GOPS_CENTRALITY_SWEEP_AUTO;
  int launch = GOPS_LAUNCH_AUTO;
  gops_run_stats stats{};
  gops::TopkScratch *topk = nullptr;     // device top_k scratch (topk.hip), on first use
  gops::TopkBatchScratch *topk_batch = nullptr;  // batched top_k scratch, on first use
  void *noalloc_scratch = nullptr;       // device address of topk.hip's file-scope scratch
  double *staging = nullptr;              // device copy of host scores for gops_top_k
  size_t staging_cap = 0;
  gops::RunState *host_state = nullptr;  // pinned, 2 slots
  uint8_t *host_log = nullptr;           // pinned, kSweepLogCap: a run's sweep modes
  uint64_t *mirror_host = nullptr;        // mapped pinned progress word
  uint64_t *mirror_dev = nullptr;         // its device address
  std::vector<hipEvent_t> events;         // timing pool
  int num_cus = 256;
  void *pin[gops::kPinChunks] = {};            // staged copies (staging.cpp), on first use
  hipEvent_t pin_done[gops::kPinChunks] = {};  // each chunk's last transfer
};

struct gops_edges {
  uint64_t n = 0, m = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> nb;
  std::vector<double> w;
};

namespace gops {

// host.cpp
gops_status build_csrt(uint64_t n, const uint64_t *off, const uint32_t *nb, const double *ew,
                       HostCSRT &out, std::string &err);
void plan_blocks(const HostCSRT &g, Plan &plan, const PlanParams &pp = PlanParams{});
uint16_t f32_to_bf16(float f);

// graph_build.hip: build_csrt's arrays made on the device and kept there
// (g->dev apart from the block plan, g->order_host, g->rank_host,
// g->first_bad_weight); plan receives the block plan
gops_status build_on_device(gops_context *ctx, uint64_t n, const uint64_t *off, const uint32_t *nb,
                            const double *ew, int weight_storage, gops_graph *g, Plan &plan,
                            std::string &err);

// graph_build.hip: the out-edge CSR (rows = sources, new ids, targets
// ascending) from the in-edge CSR, with its block plan, into g->fwd
gops_status build_forward(gops_context *ctx, gops_graph *g, std::string &err);
// the block plan's arrays into dg (desc, hubs), synchronised on stream
hipError_t upload_plan_to(hipStream_t stream, DevGraph &dg, const Plan &p);

// centrality.hip: launches of one Katz / eigenvector / HITS iteration (all
// kernels return at once once the state says done), initial vectors, and
// the permutation back to caller ids
enum PowerAlgo { kKatz = 0, kEigen = 1, kHits = 2 };
hipError_t launch_power_init(gops_context *ctx, gops_graph *g, PowerAlgo algo, double value,
                             double prescaled);
hipError_t launch_power_iteration(gops_context *ctx, gops_graph *g, PowerAlgo algo);
// units 2^-fix of a blocked Katz run's first sweep (x = beta everywhere)
int katz_initial_fix(double alpha, uint64_t indeg_max, uint64_t n, double beta);
hipError_t launch_permute(gops_context *ctx, gops_graph *g, const double *src, double *out);

// topk.hip: top_k (topk.rs:17-42) of n device scores into host ids / values
gops_status device_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                         uint32_t *ids, double *vals, uint32_t *count, std::string &err);
void free_topk(TopkScratch *t);
void free_topk_batch(TopkBatchScratch *t);
// topk.hip: the allocation-free top_k.  noalloc_scratch gives the device
// address of the library's file-scope scratch array on `device` (the current
// device), looked up once per device; top_k_noalloc selects top_k (k <=
// kNoAllocTopK, n < 2^32) of n device scores into host ids / vals with that
// scratch at `base`, allocating nothing.  Calls that overlap on one GPU share
// the array and race.
constexpr int kMaxDevices = 64;
constexpr uint32_t kNoAllocTopK = 4096;
hipError_t noalloc_scratch(int device, void **out);
gops_status top_k_noalloc(gops_context *ctx, unsigned char *base, const double *scores,
                          uint64_t n, uint32_t k, uint32_t *ids, double *vals, uint32_t *count,
                          std::string &err);
// the same, queued on ctx's stream with the results left on the device
// (ids / vals: k entries, *count); n < 2^32
hipError_t launch_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                        uint32_t *ids, double *vals, uint32_t *count);
hipError_t launch_top_k_batch(gops_context *ctx, const double *scores, uint64_t n, uint32_t nq,
                              uint32_t k, uint32_t *ids, double *vals, uint32_t *count);

// api.cpp
enum RunKind { kPlain = 0, kWeighted = 1, kPersonalized = 2 };
gops_status run_single(gops_context *ctx, gops_graph *g, const gops_pagerank_config *cfg,
                       RunKind kind, const double *pers, uint64_t pers_len, double *scores,
                       gops_run_info *info);
gops_status ppr_batch_looped(gops_context *ctx, gops_graph *g, uint32_t nq,
                             const gops_pagerank_config *cfgs, const uint64_t *seed_off,
                             const uint32_t *seeds, const double *seed_w, double *scores,
                             gops_run_info *infos);

// server.cpp
void free_ppr_cache(gops_graph *g);

// kernels (pagerank.hip)
struct SweepSetup {
  int weighted;           // use edge weights and inv_ws
  int personalized;       // q vector instead of uniform
  int blocked;            // f32 delta sweeps in the blocked layout (g->pb)
  int anchors_blocked;    // and the f64 anchor sweeps
};
hipError_t launch_init(gops_context *ctx, gops_graph *g, const SweepSetup &s);
// kinds: the blocked kernels to launch when every sweep is blocked (1 f32
// delta, 2 f64 anchor, 3 both: each returns at once for the other kind)
hipError_t launch_sweep(gops_context *ctx, gops_graph *g, const SweepSetup &s, int kinds = 3);
hipError_t launch_finalize(gops_context *ctx, gops_graph *g, int normalize);
hipError_t launch_gather_probe(gops_context *ctx, gops_graph *g);

// blocked.hip: the layout of the blocked delta sweep, built once per graph
gops_status build_blocked(gops_context *ctx, gops_graph *g, std::string &err);
// the out-edge CSR (build_forward) and its blocked layout, once per graph
gops_status build_blocked_fwd(gops_context *ctx, gops_graph *g, std::string &err);
void free_blocked(BlockedLayout &pb);

// staging.cpp: copies of pageable host memory through pinned chunks on the
// host threads.  staged_upload returns once the source has been read (the
// transfers complete in stream order on s); staged_download returns with the
// data in dst.
hipError_t staged_upload(gops_context *ctx, void *dst, const void *src, size_t bytes,
                         hipStream_t s);
hipError_t staged_download(gops_context *ctx, void *dst, const void *src, size_t bytes,
                           hipStream_t s);
void free_pinned_chunks(gops_context *ctx);

// api.cpp: drop one reference; the last one frees the object (a graph's
// release drops its reference on the context).
void release_graph(gops_graph *g);

}  // namespace gops
