// ppr_batch.h -- device state and launchers of the batched personalized
// PageRank tiles (ppr_batch.hip) shared with the host side (server.cpp).
#pragma once

#include "internal.h"

namespace gops {

constexpr int kMaxTile = 64;  // queries per tile, at most
constexpr int kLaneQ = 4;     // queries per lane in the tile sweep (one 16-byte gather),
                              // and per refill group
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
  float *yd0, *yd1;     // [n][T] prescaled deltas (f32), ping-pong
  float *carry;         // [n][T] what the row's last f32 delta rounded off, added
                        // to its next one (so the sums the sweeps propagate lag
                        // the f64 deltas by one rounding, not by all of them)
  const float *yin32;   // picked per sweep
  float *yout32;
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
  int zseed_first;          // first seed entry whose row has no in-edges (row >= n_in)
  int n_in;                 // rows with in-edges: ids [0, n_in) (rows sorted by in-degree)
  uint64_t *mirror;         // mapped host word: (sweeps << 8) | (free groups << 1) | all_done,
                            // a group being kLaneQ consecutive slots with no live query
};

hipError_t ppr_launch_sweep(hipStream_t s, int T, const PprArgs &a, int grid);
// zrows: rows without in-edges where the admitted columns may be nonzero (the
// departing queries' seed rows there), nz of them
hipError_t ppr_launch_admit(hipStream_t s, int T, const PprArgs &a, uint64_t clear_mask,
                            uint64_t admit_mask, const int32_t *zrows, int nz);
hipError_t ppr_launch_gather(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank, int n,
                             int q, double *out);
hipError_t ppr_launch_gather_mask(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank,
                                  int n, uint64_t mask, double *out);
hipError_t ppr_launch_seed_mark(hipStream_t s, int32_t *seed_idx, const int32_t *rows, int count,
                                int set);

}  // namespace gops
