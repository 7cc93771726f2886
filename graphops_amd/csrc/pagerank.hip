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
  int n;             // rows
  BlockedLayout pb;
};

constexpr int kReduce = 5;  // |delta|, dangling delta, dangling x, delta, x

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
    __hip_atomic_store(st->mirror, (it << 2) | (uint64_t)(conv << 1) | (uint64_t)done,
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
  if (a.pb_active && st->mode != kModeAnchor) return;  // the blocked kernels do this sweep
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

// red += partials[b] for b = threadIdx.x, + kBlock, ... < count, in that
// order, with four rows' loads in flight at a time
__device__ __forceinline__ void add_partials(const double *partials, int count,
                                             double (&red)[kReduce]) {
  for (int b0 = threadIdx.x; b0 < count; b0 += 4 * kBlock) {
    double v[4][kReduce];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int b = b0 + u * kBlock;
#pragma unroll
      for (int i = 0; i < kReduce; i++) v[u][i] = b < count ? partials[(size_t)b * kReduce + i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (b0 + u * kBlock < count) {
#pragma unroll
        for (int i = 0; i < kReduce; i++) red[i] += v[u][i];
      }
    }
  }
}

// One workgroup: hub-row epilogues (chunk partials summed in chunk order),
// the fixed-order reduction of the sweep's partials, the bookkeeping.
__global__ __launch_bounds__(kBlock) void sweep_finish_kernel(SweepArgs a, float *yd0, float *yd1,
                                                              double *y640, double *y641,
                                                              int sweep_blocks) {
  RunState *st = a.st;
  if (st->done) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  const int mode = st->mode;
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  auto hubs = [&](const auto &epi) {
    for (int h = threadIdx.x; h < a.num_hubs; h += kBlock) {
      const int first = a.hub_first[h], nch = a.hub_nchunks[h];
      double s = 0.0;
      for (int c = 0; c < nch; c++) s += a.hub_partial[first + c];
      epi(a.hub_row[h], s, red);
    }
  };
  if (a.pb_active && mode != kModeAnchor) {
    // a blocked sweep: hub rows were summed in their bins
    const int parts = a.pb.num_items2 + a.pb.num_splits;
    add_partials(a.pb.partials, parts, red);
  } else {
    if (mode == kModeAnchor) {
      hubs(AnchorEpilogue(a, st));
    } else {
      hubs(DeltaEpilogue(a, st, mode == kModeDeltaFold));
    }
    add_partials(a.partials, sweep_blocks, red);
  }
  block_sum<kReduce>(red);
  if (threadIdx.x == 0) finish_sweep(st, mode, red);
}

// ---------------------------------------------------------------------------
// Blocked f32 delta sweeps (layout: blocked.hip).  sweep_kernel returns at
// once for them; the scatter and accumulate kernels (and the combine kernel
// for bins split over several workgroups) do the sweep, and the finish kernel
// reduces their partial sums.  Anchor sweeps stay pulls.
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

// Scatter: one chunk of sources (their f32 deltas in LDS); every edge of the
// chunk, in chunk-major order, writes its product to its bin-major place.
// Consecutive lanes take consecutive entries, so that the writes of a run
// (entries of one chunk into one bin) are contiguous.
template <typename WT>
__global__ __launch_bounds__(kBlock) void pb_scatter_kernel(SweepArgs a, float *yd0, float *yd1) {
  const RunState *st = a.st;
  if (st->done || st->mode == kModeAnchor) return;
  const float *yin = st->yd_cur ? yd1 : yd0;
  constexpr int kChunk = 1 << kPbChunkShift;
  __shared__ float z[kChunk];
  const PbItem it = a.pb.items1[blockIdx.x];
  const int base = it.x << kPbChunkShift;
  const int len = min(kChunk, a.n - base);
  if (len == kChunk && (base & 3) == 0) {
    // all of the chunk's loads in flight at once (16-byte loads)
    constexpr int V = kChunk / 4 / kBlock;
    float4 v[V];
#pragma unroll
    for (int j = 0; j < V; j++)
      v[j] = *reinterpret_cast<const float4 *>(yin + base + 4 * ((int)threadIdx.x + j * kBlock));
#pragma unroll
    for (int j = 0; j < V; j++)
      *reinterpret_cast<float4 *>(z + 4 * ((int)threadIdx.x + j * kBlock)) = v[j];
  } else {
    for (int i = threadIdx.x; i < len; i += kBlock) z[i] = yin[base + i];
  }
  __syncthreads();
  constexpr int U = 24;
  // Waves start on a multiple of 64 entries (lanes before the item's first
  // entry idle), so that a wave's 2-byte loads fill one 128 B line: from an
  // unaligned start every wave touches two lines per load (at R-MAT scale 24,
  // 0.1 GB more fetched per sweep and 3.5 % more time).
  for (int e = (it.y & ~63) + (int)threadIdx.x; e < it.z; e += U * kBlock) {
    int sl[U];
    uint32_t p[U];
    float w[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = e + u * kBlock;
      const int kk = k >= it.y && k < it.z ? k : it.y;
      sl[u] = __builtin_nontemporal_load(a.pb.src + kk);
      const int sg = __builtin_nontemporal_load(a.pb.seg + kk);
      p[u] = (uint32_t)(kk + a.pb.seg_off[it.w + sg]);
      w[u] = wload32<WT>(a.pb.w, kk);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = e + u * kBlock;
      if (k >= it.y && k < it.z) a.pb.contrib[p[u]] = w[u] * z[sl[u]];
    }
  }
}

// c * 2^k as a fixed-point integer, truncated toward zero: the f32's
// mantissa shifted by its exponent (integer operations only).  |c| * 2^k
// stays below 2^61 (the scale's bound); truncation costs at most one unit,
// 2^-62 of the previous residual, per product.
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

// A 64-bit DPP move (row_shr / row_bcast), 0 where the control selects no lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ long long pb_dpp(long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)((unsigned long long)v >> 32), CTRL, ROWS, 0xF, true);
  return (long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Adds one value per lane of a wave to acc[key] (key < 0: no value).  A
// bin's entries come in runs of one row, so lanes with equal keys side by
// side are summed first (a segmented scan: each lane adds the lanes from its
// run's head), and only each run's last lane makes an LDS atomic.
__device__ __forceinline__ void pb_add(long long v, int key, unsigned long long *acc) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(heads & upto);
  // inclusive scan within rows of 16 lanes (row_shr 1, 2, 4, 8), then across
  // rows (row_bcast 15: rows 1 and 3 add the last lane of the row before;
  // row_bcast 31: rows 2 and 3 add lane 31); a lane adds only lanes of its run
  const int r = lane & 15;
  long long o;
  o = pb_dpp<0x111, 0xF>(v); if (r >= 1 && lane - 1 >= start) v += o;
  o = pb_dpp<0x112, 0xF>(v); if (r >= 2 && lane - 2 >= start) v += o;
  o = pb_dpp<0x114, 0xF>(v); if (r >= 4 && lane - 4 >= start) v += o;
  o = pb_dpp<0x118, 0xF>(v); if (r >= 8 && lane - 8 >= start) v += o;
  o = pb_dpp<0x142, 0xA>(v); if (((lane >> 4) & 1) && (lane | 15) - 16 >= start) v += o;
  o = pb_dpp<0x143, 0xC>(v); if (lane >= 32 && 31 >= start) v += o;
  const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
  if (tail && key >= 0) atomicAdd(&acc[key], (unsigned long long)v);
}

// Accumulate: one bin of rows (or a part of its entries), fixed-point sums in
// LDS; then the rows' epilogues, or the part's sums for the combine kernel.
__global__ __launch_bounds__(kBlock) void pb_accumulate_kernel(SweepArgs a, float *yd0, float *yd1,
                                                               double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || st->mode == kModeAnchor) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  constexpr int kRows = 1 << kPbBinShift;
  __shared__ unsigned long long acc[kRows];
  const PbItem it = a.pb.items2[blockIdx.x];
  for (int i = threadIdx.x; i < kRows; i += kBlock) acc[i] = 0ull;
  __syncthreads();
  const int k = pb_exponent(st);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int U = 16;
  // waves start on a multiple of 64 entries, as in the scatter: whole lines
  // per wave (unaligned, the two waves sharing each line both fetched it:
  // 0.65 GB more per sweep at R-MAT scale 24)
  for (int f0 = it.y & ~63; f0 < it.z; f0 += U * kBlock) {
    float cv[U];
    int key[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int f = f0 + u * kBlock + wave * 64 + lane;  // 64 consecutive entries per wave
      const bool ok = f >= it.y && f < it.z;
      const int ff = ok ? f : it.y;
      cv[u] = __builtin_nontemporal_load(a.pb.contrib + ff);
      const int d = __builtin_nontemporal_load(a.pb.dst + ff);
      key[u] = ok ? d : -1 - lane;
    }
#pragma unroll
    for (int u = 0; u < U; u++) pb_add(pb_fixed(cv[u], k), key[u], acc);
  }
  __syncthreads();
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int row0 = it.x << kPbBinShift;
  const int rows = min(kRows, a.n - row0);
  if (it.w < 0) {
    const DeltaEpilogue epi(a, st, st->mode == kModeDeltaFold);
    const double inv_scale = ldexp(1.0, -k);
    constexpr int R = 4;  // rows per thread with their loads in flight together
    for (int i0 = threadIdx.x; i0 < rows; i0 += R * kBlock) {
      DeltaEpilogue::Ops o[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        const int i = i0 + u * kBlock;
        o[u] = epi.load(row0 + (i < rows ? i : i0));
      }
#pragma unroll
      for (int u = 0; u < R; u++) {
        const int i = i0 + u * kBlock;
        if (i < rows) epi.apply(row0 + i, (double)(long long)acc[i] * inv_scale, o[u], red);
      }
    }
  } else {
    long long *out = a.pb.bin_partial + ((size_t)it.w << kPbBinShift);
    for (int i = threadIdx.x; i < rows; i += kBlock) out[i] = (long long)acc[i];
  }
  __syncthreads();
  block_sum<kReduce>(red, reinterpret_cast<double *>(acc));
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kReduce; i++) a.pb.partials[(size_t)blockIdx.x * kReduce + i] = red[i];
  }
}

// Combine: kBlock rows of a bin whose entries were split over several
// accumulate workgroups: the parts' sums in slot order, then the epilogues.
__global__ __launch_bounds__(kBlock) void pb_combine_kernel(SweepArgs a, float *yd0, float *yd1,
                                                            double *y640, double *y641) {
  const RunState *st = a.st;
  if (st->done || st->mode == kModeAnchor) return;
  pick_buffers(a, st, yd0, yd1, y640, y641);
  const PbItem sp = a.pb.splits[blockIdx.x];  // {bin, first slot, slots, first row of the bin}
  const double inv_scale = ldexp(1.0, -pb_exponent(st));
  const int row0 = (sp.x << kPbBinShift) + sp.w;
  const int i = sp.w + (int)threadIdx.x;  // row within the bin
  double red[kReduce] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < (1 << kPbBinShift) && row0 + (int)threadIdx.x < a.n) {
    // the hottest bins are split into dozens of parts: eight parts' loads in
    // flight at a time (integer sums: the grouping leaves the bits alone)
    const long long *bp = a.pb.bin_partial + ((size_t)sp.y << kPbBinShift) + i;
    long long s = 0;
    int j = 0;
    for (; j + 8 <= sp.z; j += 8) {
      long long v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = bp[(size_t)(j + u) << kPbBinShift];
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    for (; j < sp.z; j++) s += bp[(size_t)j << kPbBinShift];
    const DeltaEpilogue epi(a, st, st->mode == kModeDeltaFold);
    epi(row0 + (int)threadIdx.x, (double)s * inv_scale, red);
  }
  block_sum<kReduce>(red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kReduce; q++)
      a.pb.partials[(size_t)(a.pb.num_items2 + blockIdx.x) * kReduce + q] = red[q];
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
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) {
    const double x = a.xbar[rank[o]];
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
  a.n = (int)dg.n;
  if (a.pb_active) a.pb = g->pb;
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
  sweep_kernel<WT><<<grid, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1], rb.y64[0], rb.y64[1]);
  if (a.pb_active) {
    if (a.pb.num_items1)
      pb_scatter_kernel<WT><<<a.pb.num_items1, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1]);
    pb_accumulate_kernel<<<a.pb.num_items2, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1],
                                                                      rb.y64[0], rb.y64[1]);
    if (a.pb.num_splits)
      pb_combine_kernel<<<a.pb.num_splits, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1],
                                                                     rb.y64[0], rb.y64[1]);
  }
  sweep_finish_kernel<<<1, kBlock, 0, ctx->stream>>>(a, rb.yd[0], rb.yd[1], rb.y64[0], rb.y64[1],
                                                     grid);
}

hipError_t launch_sweep(gops_context *ctx, gops_graph *g, const SweepSetup &s) {
  SweepArgs a = make_args(g, s);
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
  SweepSetup s{0, 0, 0};
  SweepArgs a = make_args(g, s);
  const int n = (int)g->dev.n;
  const int grid = elementwise_grid(n);
  finalize_kernel<<<grid, kBlock, 0, ctx->stream>>>(a, g->dev.rank, g->run.out, n);
  sum_partials_kernel<<<1, kBlock, 0, ctx->stream>>>(a.partials, grid, &a.st->dsum_x);
  if (normalize) scale_kernel<<<grid, kBlock, 0, ctx->stream>>>(g->run.state, g->run.out, n);
  return hipGetLastError();
}

}  // namespace gops
