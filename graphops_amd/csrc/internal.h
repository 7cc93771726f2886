// internal.h -- shared host/device declarations of libgraphops_amd.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/graphops_amd.h"

namespace gops {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64s)
constexpr int kHubThreshold = 4096;  // in-degree above which a row is split over workgroups
constexpr int kHubChunk = 4096;      // edges per hub chunk (16 per thread)
constexpr int kRowsNnzTarget = 8192; // edges per row block
constexpr int kRowIters = 8;         // row iterations per group in one block

// Work descriptor of one workgroup in a sweep.
//   ROWS: x = log2(lanes per row), y = first row, z = end row
//   HUB : x = kHubFlag | chunk index, y = hub id, z = first edge, w = end edge
constexpr int32_t kHubFlag = 1 << 28;
struct BlockDesc {
  int32_t x, y, z, w;
};

// Scalars a sweep chain keeps in device memory.
struct RunState {
  double dsum_in;   // dangling sum of the vector the next sweep gathers
  double dsum_x;    // dangling sum of the current iterate (for f64 sweeps)
  double diff;      // L1 residual of the last sweep
  double damping;
  double tolerance;
  double inv_n;     // 1.0 / n
  double n_f64;     // n as f64
  uint64_t iter;
  uint64_t max_iter;
  int32_t done;
  int32_t converged;
  uint32_t ticket;
  uint32_t pad;
  // host-mapped progress word written after every sweep:
  // (iterations << 2) | (converged << 1) | done
  uint64_t *mirror;
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
  double *xbar = nullptr;        // f64 anchor iterate
  float *e = nullptr;            // f32 accumulated deltas since the anchor
  float *yd[2] = {nullptr, nullptr};   // prescaled deltas d*delta/deg (f32)
  double *y64[2] = {nullptr, nullptr}; // prescaled iterate d*x/deg (f64)
  float *q32 = nullptr;          // personalization (new ids), f32
  double *q64 = nullptr;         // personalization (new ids), f64
  double *out = nullptr;         // result in original ids
  double *partials = nullptr;    // per-workgroup partial sums
  double *hub_partial = nullptr; // per hub chunk partial sums
  uint32_t *hub_count = nullptr; // per hub arrival counters
  RunState *state = nullptr;
  size_t partial_slots = 0;
};

}  // namespace gops

struct gops_graph {
  gops_context *ctx = nullptr;
  gops::DevGraph dev;
  gops::RunBuffers run;
  bool weighted = false;   // created with edge weights
  int first_bad_weight = 0;
  std::vector<uint32_t> order_host;    // new -> original id
  std::vector<uint32_t> rank_host;     // original -> new id
};

struct gops_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int precision = GOPS_PRECISION_MIXED;
  int anchor_period = 32;
  int normalize = 0;
  int timing = 0;
  int ppr_tile = 0;
  gops_run_stats stats{};
  gops::RunState *host_state = nullptr;  // pinned, 2 slots
  uint64_t *mirror_host = nullptr;        // mapped pinned progress word
  uint64_t *mirror_dev = nullptr;         // its device address
  std::vector<hipEvent_t> events;         // timing pool
  int num_cus = 256;
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
void plan_blocks(const HostCSRT &g, Plan &plan);
uint16_t f32_to_bf16(float f);

// api.cpp
enum RunKind { kPlain = 0, kWeighted = 1, kPersonalized = 2 };
gops_status run_single(gops_context *ctx, gops_graph *g, const gops_pagerank_config *cfg,
                       RunKind kind, const double *pers, uint64_t pers_len, double *scores,
                       gops_run_info *info);
gops_status ppr_batch_looped(gops_context *ctx, gops_graph *g, uint32_t nq,
                             const gops_pagerank_config *cfgs, const uint64_t *seed_off,
                             const uint32_t *seeds, const double *seed_w, double *scores,
                             gops_run_info *infos);

// kernels (pagerank.hip)
struct SweepSetup {
  int weighted;           // use edge weights and inv_ws
  int personalized;       // q vector instead of uniform
};
hipError_t launch_init(gops_context *ctx, gops_graph *g, const SweepSetup &s);
hipError_t launch_fold(gops_context *ctx, gops_graph *g, const SweepSetup &s);
hipError_t launch_sweep_f64(gops_context *ctx, gops_graph *g, const SweepSetup &s, int y64_in,
                            int yd_out, bool write_y64, bool write_yd);
hipError_t launch_sweep_f32(gops_context *ctx, gops_graph *g, const SweepSetup &s, int parity);
hipError_t launch_finalize(gops_context *ctx, gops_graph *g, int normalize);

}  // namespace gops
