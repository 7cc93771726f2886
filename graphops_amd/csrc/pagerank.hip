// pagerank.hip -- gfx950 sweep kernels of the PageRank family.
//
// One sweep = one iteration of pagerank.rs:142-180 / 239-281 or ppr.rs:117-149,
// written as a pull over the CSR of in-edges (rows renumbered by in-degree).
// Each sweep is one pass over HBM that fuses: the gather-multiply-reduce of
// the row, the teleport and dangling-redistribution terms, the L1 residual
// (f64) and the dangling sums the next sweep needs (f64).
//
// No workgroup of a sweep waits for, or synchronises with, another: on gfx950
// a device-scope fence writes back and invalidates the XCD's L2, which would
// evict the gathered vector.  Each workgroup writes its partial sums with plain
// stores; a one-workgroup finish kernel then combines the chunks of split hub
// rows and reduces the partials in a fixed order (so results do not depend on
// scheduling) and decides what the next sweep does.  The kernel boundary is
// the only synchronisation.
//
// Mixed precision uses a "delta" scheme (sweeps over x_{k+1} - x_k that gather
// it as f32, with periodic f64 anchor sweeps); see the sweep-chain section.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "sweep_common.h"
#include "blocked_common.h"

namespace gops {

struct SweepArgs {
  const BlockDesc *desc;
  int num_desc;
  const int32_t *row_ptr;
  const int32_t *col;
  const void *w;
  const float *inv32;
  const double *inv64;
  const float *yin32;
  float *yout32;
  const double *yin64;
  double *yout64;
  double *xbar;
  const float *q32;
  const double *q64;
  double *partials;     // kReduce per workgroup of the sweep kernel
  double *hub_partial;  // one per hub chunk
  const int32_t *hub_row;
  const int32_t *hub_nchunks;
  const int32_t *hub_first;
  int num_hubs;
  int n_hot;  // min(n, kStageNnz): sources whose f32 deltas a workgroup keeps in LDS
  RunState *st;
  int pb_active;     // f32 delta sweeps run in the blocked layout (pb)
  int pb_anchor;     // so do the f64 anchor sweeps
  int n;             // rows
  BlockedLayout pb;
  int n_live;  // rows the blocked epilogues sweep: n, or n - n_iso when the isolated
               // tail (no in-edges, no out-edges) is kept in RunState::x_iso
  int kinds;   // blocked kernels launched for this sweep: 1 the f32 delta kernels,
               // 2 the f64 anchor kernels, 3 both (each returns at once for the other)
};

constexpr int kReduce = 5;  // |delta|, dangling delta, dangling x, delta, x

// The epilogues of the isolated rows in one go: each would gather 0, add the
// same uniform term and read and write the same iterate, so their common
// iterate is updated once and each sum gets n_iso times its term, before
// finish_sweep reads the sums.  The arithmetic is that of DeltaEpilogue and
// AnchorEpilogue for a row with acc = 0, inv = 0 and no personalization.
__device__ void isolated_rows(RunState *st, int mode, double (&s)[kReduce]) {
  const double n_iso = st->n_iso;
  if (!(n_iso > 0.0)) return;
  const double d = st->damping;
  double delta, x;
  if (mode == kModeAnchor) {
    const double sc = st->x_scale;
    const double dsum = st->dsum_x * sc;
    const double fill = (1.0 - d) / st->n_f64 + d * dsum / st->n_f64;  // pagerank.rs:150-152
    const double t = sc * 0.0 + fill;
    delta = t - sc * st->x_iso;
    x = t;
    s[2] += n_iso * t;
  } else {
    const double ddsum = d * (st->dsum_in - st->mass_in);
    delta = 0.0 + ddsum * st->inv_n;
    x = st->x_iso + delta;
    if (mode == kModeDeltaFold) {
      s[2] += n_iso * x;
      s[4] += n_iso * x;
    }
  }
  st->x_iso = x;
  s[0] += n_iso * fabs(delta);
  s[1] += n_iso * delta;
  s[3] += n_iso * delta;
}

// ---------------------------------------------------------------------------
// The sweep chain.  Every launch runs the same two kernels; the mode of the
// sweep is chosen on the device by the finish kernel of the previous sweep
// (RunState::mode):
//   ANCHOR      f64: t = fill + sum w * y64[u] (y64 = d*x*inv of the current
//               iterate), delta = t - x, x = t, yd = d*delta*inv stored in f32
//   DELTA       delta' = sum w * yd[u] + d*dsum(delta)*q, f32 operands summed
//               in f64; x += delta' (f64), yd' = d*delta'*inv stored in f32
//   DELTA_FOLD  DELTA, and also y64 = d*x*inv and the sums of x for the
//               ANCHOR sweep that follows.
// delta is x_{k+1} - x_k, so sum |delta| is the reference's L1 residual
// (pagerank.rs:164-173), and the only f32 rounding is the storage of the
// gathered delta (4 bytes per node instead of 8).  A DELTA_FOLD + ANCHOR pair
// is scheduled once the residual has dropped by anchor_drop since the last
// anchor (or after anchor_period sweeps): the anchor recomputes delta from the
// iterate in f64, so the f32 storage rounding never propagates over more than
// one such segment.

// Bookkeeping of pagerank.rs:163-179: iterations counts the sweep, diff_l1 is
// its residual, stop on diff < tolerance (strict) or at the iteration cap;
// then the choice of the next sweep's mode and buffers.
__device__ void finish_sweep(RunState *st, int mode, const double (&s)[kReduce]) {
  st->mass_in = s[3];
  const double diff = s[0];
  const uint64_t it = st->iter + 1;
  int next;
  if (mode == kModeAnchor) {
    st->dsum_in = s[1];
    st->dsum_x = s[2];
    st->diff_anchor = diff;
    st->since_anchor = 0;
    st->x_scale = 1.0;
    if (st->f64_phase && (st->all_f64 || diff > st->f32_below * st->tolerance)) {
      st->y64_cur ^= 1;  // this anchor wrote the next anchor's gathered iterate
      next = kModeAnchor;
    } else {
      st->f64_phase = 0;
      next = kModeDelta;
    }
  } else if (mode == kModeDeltaFold) {
    st->dsum_in = s[1];
    st->dsum_x = s[2];
    // the folded iterate sums to 1 in exact arithmetic (pagerank.rs:133 starts
    // at 1/n per node and every sweep conserves the total); the anchor rescales
    // it by 1/sum so that rounding leaves no error along the stationary
    // direction, which the iteration itself would remove only at rate d
    st->x_scale = 1.0 / s[4];
    next = kModeAnchor;
  } else {
    st->dsum_in = s[1];
    const uint64_t since = st->since_anchor + 1;
    st->since_anchor = since;
    const bool anchor = diff <= st->anchor_drop * st->diff_anchor ||
                        (since + 1 >= st->anchor_period && diff > st->anchor_min * st->tolerance);
    next = anchor ? kModeDeltaFold : kModeDelta;
  }
  st->yd_cur ^= 1;
  st->mode = next;
  int conv = 0, done = 0;
  if (diff < st->tolerance) {
    conv = 1;
    done = 1;
  } else if (it >= st->max_iter) {
    done = 1;
  }
  st->iter = it;
  st->diff = diff;
  st->converged = conv;
  st->done = done;
  if (it - 1 < st->sweep_log_cap) st->sweep_log[it - 1] = (uint8_t)mode;
  if (st->mirror) {
    __hip_atomic_store(st->mirror,
                       (it << 4) | ((uint64_t)next << 2) | (uint64_t)(conv << 1) | (uint64_t)done,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Per-row epilogue of an f32 delta sweep.
struct DeltaEpilogue {
  const SweepArgs &a;
  double d, ddsum, uni;
  bool fold;
  __device__ DeltaEpilogue(const SweepArgs &a_, const RunState *st, bool fold_) : a(a_), fold(fold_) {
    d = st->damping;
    // d * (dangling sum - total sum) of the input delta: the total is 0 in
    // exact arithmetic; removing it keeps the f32 storage rounding from
    // building up along the stationary direction (decaying only at rate d)
    ddsum = d * (st->dsum_in - st->mass_in);
    uni = ddsum * st->inv_n;
  }
  // the row's operands, loaded apart from the update so that callers can
  // have several rows' loads in flight
  struct Ops {
    double q, x, inv64;
    float inv;
  };
  __device__ __forceinline__ Ops load(int r) const {
    Ops o;
    o.q = a.q32 ? (double)a.q32[r] : 0.0;
    o.inv = __builtin_nontemporal_load(a.inv32 + r);
    o.x = __builtin_nontemporal_load(a.xbar + r);
    o.inv64 = fold ? a.inv64[r] : 0.0;
    return o;
  }
  __device__ __forceinline__ void apply(int r, double acc, const Ops &o,
                                        double (&red)[kReduce]) const {
    const double delta = acc + (a.q32 ? ddsum * o.q : uni);
    const double x = o.x + delta;
    __builtin_nontemporal_store(x, a.xbar + r);
    a.yout32[r] = (float)(d * delta * (double)o.inv);
    red[0] += fabs(delta);
    red[3] += delta;
    if (o.inv == 0.0f) red[1] += delta;
    if (fold) {
      a.yout64[r] = (d * x) * o.inv64;
      if (o.inv64 == 0.0) red[2] += x;
      red[4] += x;
    }
  }
  __device__ __forceinline__ void operator()(int r, double acc, double (&red)[kReduce]) const {
    apply(r, acc, load(r), red);
  }
};

// Per-row epilogue of an f64 anchor sweep: x = sc * xbar is the iterate
// (sc = 1 except after a mixed-precision fold).
struct AnchorEpilogue {
  const SweepArgs &a;
  double d, sc, fill_u, ddsum;
  bool write_y64;
  __device__ AnchorEpilogue(const SweepArgs &a_, const RunState *st) : a(a_) {
    d = st->damping;
    sc = st->x_scale;
    const double nf = st->n_f64;
    const double dsum = st->dsum_x * sc;
    fill_u = (1.0 - d) / nf + d * dsum / nf;  // pagerank.rs:150-152
    ddsum = d * dsum;
    write_y64 = st->f64_phase != 0;
  }
  __device__ __forceinline__ void operator()(int r, double acc, double (&red)[kReduce]) const {
    double fill = fill_u;
    if (a.q64) {  // ppr.rs:126-127
      const double q = a.q64[r];
      fill = (1.0 - d) * q + ddsum * q;
    }
    const double t = sc * acc + fill;
    const double delta = t - sc * a.xbar[r];
    a.xbar[r] = t;
    const double inv = a.inv64[r];
    a.yout32[r] = (float)((d * delta) * inv);
    if (write_y64) a.yout64[r] = (d * t) * inv;
    red[0] += fabs(delta);
    red[3] += delta;
    if (inv == 0.0) {
      red[1] += delta;
      red[2] += t;
    }
  }
};

// f32 sweep: f32 operands, summed in f64 (long rows: the product in f64;
// runs of short rows stage it in LDS as f32)
template <typename WT>
struct EdgeF32 {
  using P = float;                               // staged product
  static constexpr int kStageUnroll = kUnroll;   // loads in flight per lane
  const SweepArgs &a;
  const float *hot;  // LDS copy of yin32[0, nhot): the most gathered sources
  int nhot;
  __device__ __forceinline__ P stage_weight(int k) const { return wload32<WT>(a.w, k); }
  __device__ __forceinline__ P stage_gather(int c) const {
    return c < nhot ? hot[c] : a.yin32[c];
  }
  __device__ __forceinline__ double operator()(int k) const {
    const int c = __builtin_nontemporal_load(a.col + k);
    return (double)wload32<WT>(a.w, k) * (double)a.yin32[c];
  }
};

// f64 sweep: f64 products and sums (staged in LDS as f64)
template <typename WT>
struct EdgeF64 {
  using P = double;
  static constexpr int kStageUnroll = kUnroll / 2;
  const SweepArgs &a;
  __device__ __forceinline__ P stage_weight(int k) const { return wload64<WT>(a.w, k); }
  __device__ __forceinline__ P stage_gather(int c) const { return a.yin64[c]; }
  __device__ __forceinline__ double operator()(int k) const {
    const int c = __builtin_nontemporal_load(a.col + k);
    return wload64<WT>(a.w, k) * a.yin64[c];
  }
};

__device__ __forceinline__ void pick_buffers(SweepArgs &a, const RunState *st, float *yd0,
                                             float *yd1, double *y640, double *y641) {
  const int ycur = st->yd_cur;
  a.yin32 = ycur ? yd1 : yd0;
  a.yout32 = ycur ? yd0 : yd1;
  const int c64 = st->y64_cur;
  if (st->mode == kModeAnchor) {
    a.yin64 = c64 ? y641 : y640;
    a.yout64 = c64 ? y640 : y641;
  } else {
    a.yout64 = c64 ? y641 : y640;  // the fold writes the iterate the anchor will gather
  }
}

// The sweep: a fixed grid strides over the work descriptors.  Buffers are
// picked from the device state, so every launch of a chain has identical
// arguments.
template <typename WT>
__global__ __launch_bounds__(kBlock) void sweep_kernel(SweepArgs a, float *yd0, float *yd1,
                                                       double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done) return;
  if (a.pb_active && (st->mode != kModeAnchor || a.pb_anchor)) return;  // the blocked kernels do it
  pick_buffers(a, st, yd0, yd1, y640, y641);
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  __shared__ double stage[kStageNnz];  // a run's products (f32 or f64)
  if (st->mode == kModeAnchor) {
    const AnchorEpilogue epi(a, st);
    const EdgeF64<WT> load{a};
    for (int i = blockIdx.x; i < a.num_desc; i += gridDim.x)
      do_desc(a, a.desc[i], load, epi, red, stage);
  } else {
    // f32 products use the lower half of the stage; the upper half keeps the
    // deltas of the first kStageNnz sources (ids sorted by in-degree: about
    // 15% of all gathers at R-MAT scale 24), read from LDS instead of memory
    float *hot = reinterpret_cast<float *>(stage) + kStageNnz;
    const int nhot = a.n_hot;
    for (int i = threadIdx.x; i < nhot; i += kBlock) hot[i] = a.yin32[i];
    __syncthreads();
    const DeltaEpilogue epi(a, st, st->mode == kModeDeltaFold);
    const EdgeF32<WT> load{a, hot, nhot};
    for (int i = blockIdx.x; i < a.num_desc; i += gridDim.x)
      do_desc(a, a.desc[i], load, epi, red, stage);
  }
  block_sum<kReduce>(red, stage);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kReduce; i++) a.partials[(size_t)blockIdx.x * kReduce + i] = red[i];
  }
}

// The finish kernel's workgroup: enough threads that the partials of a
// sweep's workgroups (2048 pulled, about 9500 blocked) load in one or two
// rounds.
constexpr int kFinishBlock = 1024;

// red += partials[b] for b = threadIdx.x, + kFinishBlock, ... < count, in
// that order, with eight rows' loads in flight at a time
__device__ __forceinline__ void add_partials(const double *partials, int count,
                                             double (&red)[kReduce]) {
  constexpr int U = 8;
  for (int b0 = threadIdx.x; b0 < count; b0 += U * kFinishBlock) {
    double v[U][kReduce];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b = b0 + u * kFinishBlock;
#pragma unroll
      for (int i = 0; i < kReduce; i++) v[u][i] = b < count ? partials[(size_t)b * kReduce + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (b0 + u * kFinishBlock < count) {
#pragma unroll
        for (int i = 0; i < kReduce; i++) red[i] += v[u][i];
      }
    }
  }
}

// The blocked kernels' partial sums are stored sum-major, [kReduce][parts]
// with one part per accumulate and combine workgroup, so that a wave of the
// finish kernel loads 64 consecutive values of one sum (512 bytes) per
// instruction.
__device__ __forceinline__ double *pb_partial(const SweepArgs &a, int i, int part) {
  return a.pb.partials + (size_t)i * (size_t)(a.pb.num_items2 + a.pb.num_splits) + part;
}

// red[i] += partials[i][b] for b = threadIdx.x, + kFinishBlock, ... < count,
// in that order, with ten parts' loads of every sum in flight at a time (one
// round for R-MAT scale 24)
__device__ __forceinline__ void add_partials_by_sum(const double *partials, int count,
                                                    double (&red)[kReduce]) {
  constexpr int U = 10;
  for (int b0 = threadIdx.x; b0 < count; b0 += U * kFinishBlock) {
    double v[kReduce][U];
#pragma unroll
    for (int i = 0; i < kReduce; i++) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int b = b0 + u * kFinishBlock;
        v[i][u] = b < count ? partials[(size_t)i * count + b] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (b0 + u * kFinishBlock < count) {
#pragma unroll
        for (int i = 0; i < kReduce; i++) red[i] += v[i][u];
      }
    }
  }
}

// One workgroup: hub-row epilogues (chunk partials summed in chunk order),
// the fixed-order reduction of the sweep's partials, the bookkeeping.
__global__ __launch_bounds__(kFinishBlock) void sweep_finish_kernel(SweepArgs a, float *yd0,
                                                                    float *yd1, double *y640,
                                                                    double *y641,
                                                                    int sweep_blocks) {
  RunState *st = a.st;
  if (st->done) return;
  const int mode = st->mode;
  if (a.pb_anchor && !(a.kinds & (mode == kModeAnchor ? 2 : 1))) {
    // the host launched the other kind's kernels, which returned at once: stop
    // the run with an error rather than finish a sweep that did not run
    if (threadIdx.x == 0) {
      st->kind_miss = 1;
      st->done = 1;
      if (st->mirror)
        __hip_atomic_store(st->mirror, (st->iter << 4) | 1ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  pick_buffers(a, st, yd0, yd1, y640, y641);
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  auto hubs = [&](const auto &epi) {
    for (int h = threadIdx.x; h < a.num_hubs; h += kFinishBlock) {
      const int first = a.hub_first[h], nch = a.hub_nchunks[h];
      double s = 0.0;
      for (int c = 0; c < nch; c++) s += a.hub_partial[first + c];
      epi(a.hub_row[h], s, red);
    }
  };
  if (a.pb_active && (mode != kModeAnchor || a.pb_anchor)) {
    // a blocked sweep: hub rows were summed in their bins
    add_partials_by_sum(a.pb.partials, a.pb.num_items2 + a.pb.num_splits, red);
  } else {
    if (mode == kModeAnchor) {
      hubs(AnchorEpilogue(a, st));
    } else {
      hubs(DeltaEpilogue(a, st, mode == kModeDeltaFold));
    }
    add_partials(a.partials, sweep_blocks, red);
  }
  __shared__ double scratch[kReduce * (kFinishBlock / 64)];
  block_sum_waves<kReduce, kFinishBlock / 64>(red, scratch);
  if (threadIdx.x == 0) {
    if (a.pb_active && (mode != kModeAnchor || a.pb_anchor)) isolated_rows(st, mode, red);
    finish_sweep(st, mode, red);
  }
}

// ---------------------------------------------------------------------------
// Blocked f32 delta sweeps (layout: blocked.hip).  sweep_kernel returns at
// once for them; the scatter and accumulate kernels (and the combine kernel
// for bins split over several workgroups) do the sweep, and the finish kernel
// reduces their partial sums.  Anchor sweeps are pulls, or blocked too (the
// f64 kernels below).
//
// Row sums are fixed-point int64, so that the order in which workgroups and
// lanes add does not change a bit of the result.  The products of a delta
// sweep add up, over all rows, to at most d * sum|delta| <= the previous
// sweep's residual (a source's out-weights sum to its normaliser), so with
// the scale 2^(61 - e), residual < 2^e, no row sum or partial sum reaches
// 2^62.  One unit is 2^-62 of that residual: far below the f32 rounding of
// the products (2^-24 of each).
__device__ __forceinline__ int pb_exponent(const RunState *st) {
  int e = 0;
  const double diff = st->diff;
  if (diff > 0.0 && diff <= 1.7e308) (void)frexp(diff, &e);
  const int k = 61 - e;
  return k < -1000 ? -1000 : (k > 1000 ? 1000 : k);
}

// Scatter: one chunk of sources (their f32 deltas in LDS); the item's windows
// of kPbWin chunk-major entries, one per wave at a time, kPbLane consecutive
// entries per lane.  A lane sums its entries' products run by run in order;
// the sum still open at its last entry is carried into the next lanes by a
// segmented scan over the wave.  Every run's last entry writes the run's sum
// to its bin-major place: the runs a lane closes are consecutive run numbers,
// from the run of its first entry (the window's first run plus the run heads
// before that entry), all in the lane's segment (segments start at a lane's
// first entry), whose offset one load brings.
//
// 1024 threads per scatter workgroup: the 64 KB of LDS allows two workgroups
// per CU, and a larger workgroup puts more waves beside each other's loads
// (measured at R-MAT scale 24: 256 threads 579 us, 512 553 us, 1024 543 us
// per f32 sweep); the register budget of that occupancy lets all of a wave's
// windows' loads be in flight at once.
template <typename WT>
__global__ __launch_bounds__(kScatterBlock) __attribute__((amdgpu_waves_per_eu(2 * kScatterBlock / 256, 2 * kScatterBlock / 256)))
void pb_scatter_kernel(SweepArgs a, float *yd0, float *yd1) {
  const RunState *st = a.st;
  if (st->done || st->mode == kModeAnchor) return;
  const float *yin = st->yd_cur ? yd1 : yd0;
  constexpr int kChunk = 1 << kPbChunkShift;
  __shared__ float z[kChunk];
  const PbItem it = a.pb.items1[blockIdx.x];  // {chunk, first window, end window}
  const int base = it.x << kPbChunkShift;
  const int len = min(kChunk, a.n - base);
  if (len == kChunk && (base & 3) == 0) {
    // all of the chunk's loads in flight at once (16-byte loads)
    constexpr int V = kChunk / 4 / kScatterBlock;
    float4 v[V];
#pragma unroll
    for (int j = 0; j < V; j++)
      v[j] = *reinterpret_cast<const float4 *>(yin + base + 4 * ((int)threadIdx.x + j * kScatterBlock));
#pragma unroll
    for (int j = 0; j < V; j++)
      *reinterpret_cast<float4 *>(z + 4 * ((int)threadIdx.x + j * kScatterBlock)) = v[j];
  } else {
    for (int i = threadIdx.x; i < len; i += kScatterBlock) z[i] = yin[base + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint32_t nruns = (uint32_t)a.pb.num_runs, nseg = (uint32_t)a.pb.num_segs;
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  constexpr int U = 4;  // windows' loads in flight per wave
  constexpr int kWaves = kScatterBlock / 64;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int w0 = it.y + wave; w0 < it.z; w0 += U * kWaves) {
    pb_u32x4 sq[U];
    float wv[U][kPbLane];
    int run0[U], seg0[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int wi = min(w0 + u * kWaves, it.z - 1);  // clamped: every load in bounds
      const size_t e = (size_t)wi * kPbWin + (size_t)lane * kPbLane;
      sq[u] = __builtin_nontemporal_load(reinterpret_cast<const pb_u32x4 *>(a.pb.src + e));
      pb_load_w8<WT>(std::is_same<WT, NoW>::value ? (const void *)a.pb.mask : a.pb.w, e, wv[u]);
      run0[u] = a.pb.win[2 * wi];
      seg0[u] = a.pb.win[2 * wi + 1];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the loads together, ahead of their uses
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (w0 + u * kWaves >= it.z) break;  // the same in every lane
      uint32_t sw[kPbLane];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        sw[2 * i] = sq[u][i] & 0xffffu;
        sw[2 * i + 1] = sq[u][i] >> 16;
      }
      // products, and the in-lane sums run by run (q[j]: the sum of the
      // entries of j's run up to j within the lane)
      float q[kPbLane];
      uint32_t hm = 0;
#pragma unroll
      for (int j = 0; j < kPbLane; j++) {
        const float pj = wv[u][j] * z[sw[j] & kPbSrcMask];
        const bool h = (sw[j] & kPbRunHead) != 0;
        hm |= (h ? 1u : 0u) << j;
        q[j] = (j == 0 || h) ? pj : q[j - 1] + pj;
      }
      // the open run carried across lanes
      const unsigned long long lheads = __ballot(hm != 0);  // bit 0: the window's first entry
      const int start = 63 - __clzll(lheads & upto);
      const float S = pb_seg_scan(q[kPbLane - 1], start);
      const float X = __shfl_up(S, 1);  // the open run's sum up to the previous lane
      // run heads before and segment heads up to each lane (packed counts)
      const uint32_t sh = (sw[0] & kPbSegHead) ? 1u : 0u;
      const uint32_t cnt = (uint32_t)__popc(hm) | (sh << 16);
      const uint32_t before = pb_int_scan(cnt) - cnt;
      const uint32_t sh0 = __builtin_amdgcn_readfirstlane(sh);
      const uint32_t run = (uint32_t)run0[u] + (before & 0xffffu) + (hm & 1u) - 1u;
      const uint32_t seg = (uint32_t)seg0[u] + (before >> 16) + sh - sh0;
      const int32_t off = a.pb.seg_off[seg < nseg ? seg : 0];
      // the ballot runs in every lane, ahead of the test that may skip it
      const unsigned long long first_heads = __ballot((hm & 1u) != 0);
      const bool next_head = lane == 63 || ((first_heads >> (lane + 1)) & 1ull);
      uint32_t place = run + (uint32_t)off;
#pragma unroll
      for (int j = 0; j < kPbLane; j++) {
        const bool tail = j + 1 < kPbLane ? ((hm >> (j + 1)) & 1u) != 0 : next_head;
        const float t = (hm & ((2u << j) - 1u)) ? q[j] : q[j] + X;
        if (tail) {
          // the layout keeps both in range; the checks keep a damaged layout
          // from writing outside rsum
          if (seg < nseg && place < nruns) a.pb.rsum[place] = t;
          place++;
        }
      }
    }
  }
}

// c * 2^k as a fixed-point integer, truncated toward zero: the f32's
// mantissa shifted by its exponent (integer operations only).  |c| * 2^k
// stays below 2^61 (the scale's bound); truncation costs at most one unit,
// 2^-62 of the previous residual, per run.  Finite c only.
__device__ __forceinline__ long long pb_fixed(float c, int k) {
  const uint32_t b = __float_as_uint(c);
  const int ex = (int)((b >> 23) & 0xffu);
  if (ex == 0) return 0;  // zero or subnormal: below one unit at any scale used
  const uint64_t mant = (b & 0x7fffffu) | 0x800000u;
  const int sh = ex - 150 + k;  // c = mant * 2^(ex - 150)
  uint64_t v;
  if (sh >= 0) v = sh < 40 ? mant << sh : 0;  // sh >= 40 only for values beyond the bound
  else v = sh > -24 ? mant >> -sh : 0;
  return (b >> 31) ? -(long long)v : (long long)v;
}

// Non-finite run sums (an infinite or NaN weight or delta) carry through as
// they would in a pull's f64 sum: per row, bit 0 NaN, bit 1 +inf, bit 2 -inf.
__device__ __forceinline__ uint32_t pb_nonfinite_bits(float c) {
  const uint32_t b = __float_as_uint(c);
  return (b & 0x7fffffu) ? 1u : ((b >> 31) ? 4u : 2u);
}

__device__ __forceinline__ double pb_row_value(long long s, uint32_t flags, double inv_scale) {
  if (flags) {
    if ((flags & 1u) || (flags & 6u) == 6u) return __builtin_nan("");
    return (flags & 2u) ? __builtin_inf() : -__builtin_inf();
  }
  return (double)s * inv_scale;
}

// Accumulate: one bin of rows (or a part of its runs), fixed-point sums in
// LDS, one atomic per run; then the rows' epilogues, or the part's sums for
// the combine kernel.
// Threads of an f32 accumulate workgroup: with 36 KB of LDS four workgroups
// fit on a CU, so 512 threads keep 32 waves per CU against 16 at 256
// (measured at R-MAT scale 24 on one box: f32 sweep 500.8 us at 256 threads,
// 491.2 us at 512, 548.4 us at 1024, two workgroups per CU).
#ifndef GOPS_TUNE_ACC_BLOCK
#define GOPS_TUNE_ACC_BLOCK 512
#endif
constexpr int kAccBlock = GOPS_TUNE_ACC_BLOCK;
__global__ __launch_bounds__(kAccBlock) void pb_accumulate_kernel(SweepArgs a, float *yd0, float *yd1,
                                                               double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || st->mode == kModeAnchor) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  constexpr int kRows = 1 << kPbBinShift;
  __shared__ unsigned long long acc[kRows];
  __shared__ uint32_t nonfin[kRows / 4];  // a byte of pb_nonfinite_bits per row
  const PbItem it = a.pb.items2[blockIdx.x];
  const int row0 = a.pb.bin_row0[it.x];
  if (row0 >= a.n_live) {  // a bin of the isolated tail: no runs, no rows
    if (threadIdx.x < kReduce) *pb_partial(a, threadIdx.x, blockIdx.x) = 0.0;
    return;
  }
  const int rows = a.pb.bin_row0[it.x + 1] - row0;  // at most kRows
  for (int i = threadIdx.x; i < kRows; i += kAccBlock) acc[i] = 0ull;
  for (int i = threadIdx.x; i < kRows / 4; i += kAccBlock) nonfin[i] = 0u;
  __syncthreads();
  const int k = pb_exponent(st);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int U = 16;
  // waves start on a multiple of 64 runs: whole lines per wave
  if (it.y < it.z) {
    for (int f0 = it.y & ~63; f0 < it.z; f0 += U * kAccBlock) {
      float cv[U];
      int key[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int f = f0 + u * kAccBlock + wave * 64 + lane;  // 64 consecutive runs per wave
        const bool ok = f >= it.y && f < it.z;
        const int ff = ok ? f : it.y;
        cv[u] = __builtin_nontemporal_load(a.pb.rsum + ff);
        const int d = __builtin_nontemporal_load(a.pb.rrow + ff);
        key[u] = ok ? d : -1;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (key[u] < 0) continue;
        const float c = cv[u];
        if (((__float_as_uint(c) >> 23) & 0xffu) == 0xffu)
          atomicOr(&nonfin[key[u] >> 2], pb_nonfinite_bits(c) << (8 * (key[u] & 3)));
        else
          atomicAdd(&acc[key[u]], (unsigned long long)pb_fixed(c, k));
      }
    }
  }
  __syncthreads();
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (it.w < 0) {
    const DeltaEpilogue epi(a, st, st->mode == kModeDeltaFold);
    const double inv_scale = ldexp(1.0, -k);
    constexpr int R = 4;  // rows per thread with their loads in flight together
    const int live = min(rows, a.n_live - row0);  // the isolated tail is the finish kernel's
    for (int i0 = threadIdx.x; i0 < live; i0 += R * kAccBlock) {
      DeltaEpilogue::Ops o[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        const int i = i0 + u * kAccBlock;
        o[u] = epi.load(row0 + (i < live ? i : i0));
      }
#pragma unroll
      for (int u = 0; u < R; u++) {
        const int i = i0 + u * kAccBlock;
        if (i < live) {
          const uint32_t fl = (nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu;
          epi.apply(row0 + i, pb_row_value((long long)acc[i], fl, inv_scale), o[u], red);
        }
      }
    }
  } else {
    long long *out = a.pb.bin_partial + ((size_t)it.w << kPbBinShift);
    uint8_t *fout = a.pb.bin_flags + ((size_t)it.w << kPbBinShift);
    for (int i = threadIdx.x; i < rows; i += kAccBlock) {
      out[i] = (long long)acc[i];
      fout[i] = (uint8_t)((nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
  }
  __syncthreads();
  if constexpr (kAccBlock == kBlock)
    block_sum<kReduce>(red, reinterpret_cast<double *>(acc));
  else
    block_sum_waves<kReduce, kAccBlock / 64>(red, reinterpret_cast<double *>(acc));
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kReduce; i++) *pb_partial(a, i, blockIdx.x) = red[i];
  }
}

// the parts of a bin split over several workgroups in a blocked anchor (below)
__device__ void pb_combine64(const SweepArgs &a, const RunState *st);

// Combine: kBlock rows of a bin whose runs were split over several
// accumulate workgroups: the parts' sums in slot order, then the epilogues.
__global__ __launch_bounds__(kBlock) void pb_combine_kernel(SweepArgs a, float *yd0, float *yd1,
                                                            double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || (st->mode == kModeAnchor && !a.pb_anchor)) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  if (st->mode == kModeAnchor) {  // a blocked anchor's parts
    pb_combine64(a, st);
    return;
  }
  const PbItem sp = a.pb.splits[blockIdx.x];  // {bin, first slot, slots, first row of the bin}
  const double inv_scale = ldexp(1.0, -pb_exponent(st));
  const int row0 = a.pb.bin_row0[sp.x] + sp.w;
  const int i = sp.w + (int)threadIdx.x;  // row within the bin
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (row0 + (int)threadIdx.x < a.pb.bin_row0[sp.x + 1]) {
    // the hottest bins are split into dozens of parts: eight parts' loads in
    // flight at a time (integer sums: the grouping leaves the bits alone)
    const size_t at = ((size_t)sp.y << kPbBinShift) + i;
    const long long *bp = a.pb.bin_partial + at;
    const uint8_t *fp = a.pb.bin_flags + at;
    long long s = 0;
    uint32_t fl = 0;
    int j = 0;
    for (; j + 8 <= sp.z; j += 8) {
      long long v[8];
      uint32_t f[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        v[u] = bp[(size_t)(j + u) << kPbBinShift];
        f[u] = fp[(size_t)(j + u) << kPbBinShift];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        s += v[u];
        fl |= f[u];
      }
    }
    for (; j < sp.z; j++) {
      s += bp[(size_t)j << kPbBinShift];
      fl |= fp[(size_t)j << kPbBinShift];
    }
    const DeltaEpilogue epi(a, st, st->mode == kModeDeltaFold);
    if (row0 + (int)threadIdx.x < a.n_live)
      epi(row0 + (int)threadIdx.x, pb_row_value(s, fl, inv_scale), red);
  }
  block_sum<kReduce>(red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kReduce; q++)
      *pb_partial(a, q, a.pb.num_items2 + blockIdx.x) = red[q];
  }
}

// ---------------------------------------------------------------------------
// Blocked f64 anchor sweeps: the same layout, the iterate's prescaled f64
// values y64 = d * x * inv in place of the f32 deltas.  The scatter holds a
// chunk's 16 384 values in 128 KB of LDS (one workgroup per CU) and sums each
// run in f64, in the same lane order as the f32 scatter; the accumulate adds
// the runs of a row as 128-bit fixed-point integers with 2^-100 units, so the
// row sum is one rounding of the exact sum of the run sums, whatever the
// order of the atomics.  An anchor's products are w * d * x_u / wsum_u <= d *
// x_u, so every row sum is at most d * sum(x) ~ 1, far below the 2^26 that
// the format holds; a run sum beyond it (a layout or input fault) makes the
// row NaN rather than wrap.
// Scatter of an anchor sweep: pb_scatter_kernel's windows and run logic on
// f64 products (f64 weights as stored, the others converted exactly) with
// f64 run sums.
template <typename WT>
__global__ __launch_bounds__(kScatterBlock) __attribute__((amdgpu_waves_per_eu(kScatterBlock / 256, kScatterBlock / 256)))
void pb_scatter64_kernel(SweepArgs a, double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || st->mode != kModeAnchor) return;
  const double *yin = st->y64_cur ? y641 : y640;
  __shared__ double z[1 << kPbChunkShift];
  const PbItem it = a.pb.items1[blockIdx.x];  // {chunk, first window, end window}
  pb_load_chunk64(yin, a.n, it, z);
  pb_scatter64_windows<WT>(a.pb, it, z);
}

// Accumulate of an anchor sweep: one bin of rows (or a part of its runs),
// 128-bit fixed-point sums in LDS (an atomic on the low word, its carry into
// the high word), then the rows' anchor epilogues or the part's sums.
constexpr int kAcc64Block = 1024;  // 68 KB of LDS: two workgroups, 32 waves per CU
__global__ __launch_bounds__(kAcc64Block) void pb_accumulate64_kernel(SweepArgs a, float *yd0, float *yd1,
                                                                 double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || st->mode != kModeAnchor) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  constexpr int kRows = 1 << kPbBinShift;
  __shared__ unsigned long long acc_lo[kRows], acc_hi[kRows];
  __shared__ uint32_t nonfin[kRows / 4];
  const PbItem it = a.pb.items2[blockIdx.x];
  const int row0 = a.pb.bin_row0[it.x];
  if (row0 >= a.n_live) {  // a bin of the isolated tail: no runs, no rows
    if (threadIdx.x < kReduce) *pb_partial(a, threadIdx.x, blockIdx.x) = 0.0;
    return;
  }
  const int rows = a.pb.bin_row0[it.x + 1] - row0;  // at most kRows
  for (int i = threadIdx.x; i < kRows; i += kAcc64Block) {
    acc_lo[i] = 0ull;
    acc_hi[i] = 0ull;
  }
  for (int i = threadIdx.x; i < kRows / 4; i += kAcc64Block) nonfin[i] = 0u;
  __syncthreads();
  pb_accumulate128_runs<kAcc64Block>(a.pb, it, kAnchorFix, acc_lo, acc_hi, nonfin);
  __syncthreads();
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (it.w < 0) {
    const AnchorEpilogue epi(a, st);
    const int live = min(rows, a.n_live - row0);  // the isolated tail is the finish kernel's
    for (int i = threadIdx.x; i < live; i += kAcc64Block) {
      const uint32_t fl = (nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu;
      epi(row0 + i, pb_row_value128(acc_lo[i], acc_hi[i], fl, kAnchorFix), red);
    }
  } else {
    long long *out = a.pb.bin_partial + ((size_t)it.w << (kPbBinShift + 1));
    uint8_t *fout = a.pb.bin_flags + ((size_t)it.w << kPbBinShift);
    for (int i = threadIdx.x; i < rows; i += kAcc64Block) {
      out[2 * i] = (long long)acc_lo[i];
      out[2 * i + 1] = (long long)acc_hi[i];
      fout[i] = (uint8_t)((nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
  }
  __syncthreads();
  block_sum_waves<kReduce, kAcc64Block / 64>(red, reinterpret_cast<double *>(acc_lo));
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kReduce; i++) *pb_partial(a, i, blockIdx.x) = red[i];
  }
}

// Combine of an anchor sweep: the parts' 128-bit sums in slot order, then the
// anchor epilogues.
__device__ void pb_combine64(const SweepArgs &a, const RunState *st) {
  const PbItem sp = a.pb.splits[blockIdx.x];  // {bin, first slot, slots, first row of the bin}
  const int row0 = a.pb.bin_row0[sp.x] + sp.w;
  const int i = sp.w + (int)threadIdx.x;
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (row0 + (int)threadIdx.x < a.pb.bin_row0[sp.x + 1]) {
    unsigned long long lo, hi;
    uint32_t fl;
    pb_combine128_row(a.pb, sp, i, lo, hi, fl);
    const AnchorEpilogue epi(a, st);
    if (row0 + (int)threadIdx.x < a.n_live)
      epi(row0 + (int)threadIdx.x, pb_row_value128(lo, hi, fl, kAnchorFix), red);
  }
  block_sum<kReduce>(red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kReduce; q++)
      *pb_partial(a, q, a.pb.num_items2 + blockIdx.x) = red[q];
  }
}

// ---------------------------------------------------------------------------
// init: x0 = 1/n (pagerank.rs:133) or the normalised personalization
// (ppr.rs:104-110); x = x0, y64[0] = d*x0*inv; partial dangling sums of x0.
__global__ __launch_bounds__(kBlock) void init_kernel(SweepArgs a, int n) {
  const RunState *st = a.st;
  const double d = st->damping;
  double red[1] = {0.0};
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
    const double x = a.q64 ? a.q64[v] : st->inv_n;
    a.xbar[v] = x;
    const double inv = a.inv64[v];
    a.yout64[v] = (d * x) * inv;
    if (inv == 0.0) red[0] += x;
  }
  block_sum<1>(red);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = red[0];
}

// out[o] = x[rank[o]] in the caller's ids; partial sums of x.
__global__ __launch_bounds__(kBlock) void finalize_kernel(SweepArgs a, const uint32_t *rank,
                                                          double *out, int n) {
  double red[1] = {0.0};
  const double x_iso = a.st->x_iso;
  // the run kept the isolated tail's iterate apart if its state says so
  const int n_live = a.st->n_iso > 0.0 ? a.n_live : n;
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) {
    const int r = (int)rank[o];
    const double x = r >= n_live ? x_iso : a.xbar[r];
    out[o] = x;
    red[0] += x;
  }
  block_sum<1>(red);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = red[0];
}

// One workgroup: *dst = sum of partials[0..count) in a fixed order.
__global__ __launch_bounds__(kBlock) void sum_partials_kernel(const double *partials, int count,
                                                              double *dst) {
  double red[1] = {0.0};
  for (int b = threadIdx.x; b < count; b += kBlock) red[0] += partials[b];
  block_sum<1>(red);
  if (threadIdx.x == 0) *dst = red[0];
}

// reference normalize() semantics (topk.rs:54-61): divide by the sum if > 0
__global__ __launch_bounds__(kBlock) void scale_kernel(const RunState *st, double *out, int n) {
  const double s = st->dsum_x;
  if (!(s > 0.0)) return;
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) out[o] /= s;
}

// The f32 sweep's gathers alone (calibration): every edge's source value is
// gathered in stored order, kUnroll loads in flight per lane, and summed, so
// that nothing but the col stream and the gathers touches memory.
__global__ __launch_bounds__(kBlock) void gather_probe_kernel(const int32_t *col, const float *y,
                                                              int m, double *partials) {
  double acc = 0.0;
  const int stride = gridDim.x * kBlock;
  for (int k0 = blockIdx.x * kBlock + threadIdx.x; k0 < m; k0 += kUnroll * stride) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int k = k0 + u * stride;
      v[u] = k < m ? y[__builtin_nontemporal_load(col + k)] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc += (double)v[u];
  }
  double red[1] = {acc};
  block_sum<1>(red);
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

hipError_t launch_gather_probe(gops_context *ctx, gops_graph *g) {
  const int grid = kMaxSweepBlocks;
  gather_probe_kernel<<<grid, kBlock, 0, ctx->stream>>>(g->dev.col, g->run.yd[0], (int)g->dev.m,
                                                        g->run.partials);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
static SweepArgs make_args(gops_graph *g, const SweepSetup &s) {
  SweepArgs a{};
  DevGraph &dg = g->dev;
  RunBuffers &rb = g->run;
  a.desc = dg.desc;
  a.num_desc = dg.num_desc;
  a.row_ptr = dg.row_ptr;
  a.col = dg.col;
  a.w = s.weighted ? dg.w : nullptr;
  a.inv32 = s.weighted ? dg.inv_ws32 : dg.inv_deg32;
  a.inv64 = s.weighted ? dg.inv_ws64 : dg.inv_deg64;
  a.xbar = rb.xbar;
  a.q32 = s.personalized ? rb.q32 : nullptr;
  a.q64 = s.personalized ? rb.q64 : nullptr;
  a.partials = rb.partials;
  a.hub_partial = rb.hub_partial;
  a.hub_row = dg.hub_row;
  a.hub_nchunks = dg.hub_nchunks;
  a.hub_first = dg.hub_first;
  a.num_hubs = dg.num_hubs;
  a.n_hot = (int)std::min<uint64_t>(dg.n, kStageNnz);
  a.st = rb.state;
  a.pb_active = s.blocked && g->has_pb;
  a.pb_anchor = a.pb_active && s.anchors_blocked;
  a.n = (int)dg.n;
  if (a.pb_active) a.pb = g->pb;
  // isolated rows become one shared iterate when every sweep is blocked and
  // the teleport is uniform
  a.n_live = (a.pb_anchor && !s.personalized) ? (int)(dg.n - g->pb.n_iso) : (int)dg.n;
  a.kinds = 3;
  return a;
}

static int elementwise_grid(uint64_t n) {
  uint64_t blocks = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  return (int)blocks;
}

int sweep_grid(const gops_graph *g) { return std::min(g->dev.num_desc, kMaxSweepBlocks); }

hipError_t launch_init(gops_context *ctx, gops_graph *g, const SweepSetup &s) {
  SweepArgs a = make_args(g, s);
  a.yout64 = g->run.y64[0];
  const int n = (int)g->dev.n;
  const int grid = elementwise_grid(n);
  init_kernel<<<grid, kBlock, 0, ctx->stream>>>(a, n);
  sum_partials_kernel<<<1, kBlock, 0, ctx->stream>>>(a.partials, grid, &a.st->dsum_x);
  return hipGetLastError();
}

template <typename WT>
static void launch_wt(gops_context *ctx, gops_graph *g, const SweepArgs &a) {
  RunBuffers &rb = g->run;
  const int grid = sweep_grid(g);
  if (!a.pb_anchor)  // pull anchors (and every sweep when not blocked)
    sweep_kernel<WT><<<grid, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1], rb.y64[0], rb.y64[1]);
  if (a.pb_active) {
    if (a.kinds & 1) {
      if (a.pb.num_items1)
        pb_scatter_kernel<WT><<<a.pb.num_items1, kScatterBlock, 0, ctx->stream>>>(a, rb.yd[0],
                                                                                 rb.yd[1]);
      pb_accumulate_kernel<<<a.pb.num_items2, kAccBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1],
                                                                        rb.y64[0], rb.y64[1]);
    }
    if (a.pb_anchor && (a.kinds & 2)) {
      if (a.pb.num_items1)
        pb_scatter64_kernel<WT><<<a.pb.num_items1, kScatterBlock, 0, ctx->stream>>>(a, rb.y64[0],
                                                                                   rb.y64[1]);
      pb_accumulate64_kernel<<<a.pb.num_items2, kAcc64Block, 0, ctx->stream>>>(
          a, rb.yd[0], rb.yd[1], rb.y64[0], rb.y64[1]);
    }
    // the parts of split bins, of either kind: after both accumulates
    if (a.pb.num_splits)
      pb_combine_kernel<<<a.pb.num_splits, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1],
                                                                     rb.y64[0], rb.y64[1]);
  }
  sweep_finish_kernel<<<1, kFinishBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1], rb.y64[0], rb.y64[1],
                                                     grid);
}

hipError_t launch_sweep(gops_context *ctx, gops_graph *g, const SweepSetup &s, int kinds) {
  SweepArgs a = make_args(g, s);
  // one kind's kernels only where every sweep is blocked (else the pull takes
  // the anchors and the blocked kernels run for the deltas)
  a.kinds = a.pb_anchor ? kinds : 3;
  if (!s.weighted) {
    launch_wt<NoW>(ctx, g, a);
  } else {
    switch (g->dev.weight_storage) {
      case GOPS_WEIGHT_F32: launch_wt<float>(ctx, g, a); break;
      case GOPS_WEIGHT_BF16: launch_wt<Bf16W>(ctx, g, a); break;
      default: launch_wt<double>(ctx, g, a); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_finalize(gops_context *ctx, gops_graph *g, int normalize) {
  SweepSetup s{0, 0, 0, 0};
  SweepArgs a = make_args(g, s);
  const int n = (int)g->dev.n;
  a.n_live = g->has_pb ? (int)(g->dev.n - g->pb.n_iso) : n;  // if the run's state says so
  const int grid = elementwise_grid(n);
  finalize_kernel<<<grid, kBlock, 0, ctx->stream>>>(a, g->dev.rank, g->run.out, n);
  sum_partials_kernel<<<1, kBlock, 0, ctx->stream>>>(a.partials, grid, &a.st->dsum_x);
  if (normalize) scale_kernel<<<grid, kBlock, 0, ctx->stream>>>(g->run.state, g->run.out, n);
  return hipGetLastError();
}

}  // namespace gops
