// ppr_batch.h -- device state and launchers of the batched personalized
// PageRank tiles (ppr_batch.hip) shared with the host side (server.cpp).
#pragma once

#include "internal.h"

namespace gops {

constexpr int kMaxTile = 64;  // queries per tile, at most
constexpr int kLaneQ = 4;     // queries per refill group (and per lane with f32 deltas)

// Queries per lane in the tile sweep.  The host's work descriptors depend
// on it (Q = T / this lanes per edge).  With f16 deltas 8 per lane (one
// 16-byte gather, 8 lanes per 64-query edge) measured slower than 4 (0.0624
// against about 0.058 ms per query-sweep at T = 64, scale 22): it spills
// registers at five waves per SIMD and halves the rows a pass can spread
// over.
constexpr int tile_lane_queries(bool delta16) { return delta16 ? 4 : 4; }
constexpr int kTileReduce = 64;  // workgroups of the first stage of the per-query sums

// Device state of one tile (T <= kMaxTile queries).  The host writes it while
// no kernel of the tile is queued; the finish kernel updates it every sweep.
struct TileState {
  int32_t yd_cur;        // yd buffer the next sweep gathers
  int32_t all_done;      // no live query left
  int32_t y64_par;       // y64 buffer (0: y64a, 1: y64b) the next sweep's anchors gather;
                         // flips every sweep, so each sweep writes the other one
  int32_t pad0;
  uint64_t sweep;        // sweeps executed by the tile
  uint64_t anchor_period;
  double anchor_drop;    // 0: no residual-triggered anchors
  double f32_below;      // f32 sweeps only once diff <= f32_below * tol
  double anchor_min;     // anchors only while diff > anchor_min * tol
  // per query slot
  double damping[kMaxTile];
  double tol[kMaxTile];
  double dsum_in[kMaxTile];
  double mass_in[kMaxTile];
  double dsum_x[kMaxTile];
  double init_dsum[kMaxTile];  // dangling part of the personalization (admission)
  double init_mass[kMaxTile];  // sum of the personalization (admission)
  double diff[kMaxTile];
  double diff_anchor[kMaxTile];
  double x_scale[kMaxTile];
  // f16 delta storage: the power-of-two scale of the stored deltas the next
  // sweep gathers (ysc_in) and of the ones it writes (ysc_out); 1 for f32
  double ysc_in[kMaxTile];
  double ysc_out[kMaxTile];
  uint64_t iter[kMaxTile];
  uint64_t max_iter[kMaxTile];
  uint64_t since_anchor[kMaxTile];
  int32_t mode[kMaxTile];  // mode of the query's next sweep (kModeDelta / DeltaFold / Anchor)
  int32_t f64_phase[kMaxTile];  // still in the leading all-f64 phase
  int32_t live[kMaxTile];
  int32_t converged[kMaxTile];
};

struct PprArgs {
  TileState *ts;
  const BlockDesc *desc;
  int num_desc;
  const int32_t *row_ptr;
  const int32_t *col;
  const float *inv32;   // 1/out-degree (0 = dangling), new ids
  const double *inv64;
  double *x;            // [n][T] iterate
  void *yd0, *yd1;      // [n][T] prescaled deltas, ping-pong: f32, or f16 times the
                        // query's scale (delta16)
  float *carry;         // [n][T] what the row's last stored delta rounded off, added
                        // to its next one (so the sums the sweeps propagate lag
                        // the f64 deltas by one rounding, not by all of them)
  const void *yin;      // picked per sweep
  void *yout;
  int delta16;          // storage of yd0 / yd1: 0 f32, 1 f16 (GOPS_OPT_TILE_DELTA)
  double *y64a, *y64b;  // [n][T] prescaled iterate for anchor sweeps, ping-pong
  double *partials;     // [grid][T][5] per sweep workgroup, then [kTileReduce][T][5]
                        // after the first reduction stage (at kMaxSweepBlocks * T * 5)
  double *hub_partial;  // [chunks][T]
  const int32_t *hub_row, *hub_nchunks, *hub_first;
  int num_hubs;
  const int32_t *seed_idx;  // [n]: first seed entry of the row, -1 if none
  const int32_t *seed_row;  // seed entries sorted by (row, query)
  const int32_t *seed_q;
  const double *seed_p;     // normalised personalization value (ppr.rs:104-109)
  int num_seed_entries;
  uint64_t *mirror;         // mapped host word: (sweeps << 8) | (free groups << 1) | all_done,
                            // a group being kLaneQ consecutive slots with no live query
};

hipError_t ppr_launch_sweep(hipStream_t s, int T, const PprArgs &a, int grid);
hipError_t ppr_launch_admit(hipStream_t s, int T, const PprArgs &a, int n, uint64_t clear_mask,
                            uint64_t admit_mask);
hipError_t ppr_launch_gather(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank, int n,
                             int q, double *out);
hipError_t ppr_launch_gather_mask(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank,
                                  int n, uint64_t mask, double *out);
hipError_t ppr_launch_seed_mark(hipStream_t s, int32_t *seed_idx, const int32_t *rows, int count,
                                int set);

}  // namespace gops
