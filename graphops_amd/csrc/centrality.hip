// centrality.hip -- gfx950 kernels of Katz, eigenvector and HITS centrality
// (katz.rs:136-201, eigenvector.rs:49-125, centrality.rs:114-166): power
// iterations in f64 over the same pull machinery as PageRank (work
// descriptors, LDS-staged runs, split hub rows, no inter-workgroup waits).
//
//   Katz         x <- beta + A^T (alpha x): one pull over the in-edge CSR
//                gathering the prescaled alpha * x, residual sum |x - x'|.
//   eigenvector  t = x + A x: a pull over the out-edge CSR gathering x
//                (its rows ordered by out-degree, mapped to nodes);
//                then x' = t * sign / |t|_2 (sign of sum t), residual
//                sum |x - x'|.
//   HITS         a = A^T h (in-edge CSR), a' = a / |a|_2; h = A a' (out-edge
//                CSR), h' = h / |h|_2; residuals |a - a'|_2 and |h - h'|_2.
//
// Katz also runs blocked (from 2^21 nodes by default): PageRank's blocked
// layout of the in-edges (blocked.hip) with unit weights, a scatter pass of
// f64 run sums of alpha * x and an accumulate pass of 128-bit fixed-point
// row sums whose units follow a bound on the row sums (below).
//
// Every kernel of an iteration first reads PowerState::done and returns at
// once when the chain has finished, so the host can queue iterations ahead.
// Reductions are per workgroup and then summed in a fixed order by one
// workgroup, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "sweep_common.h"
#include "blocked_common.h"

namespace gops {

constexpr int kPowR = 2;  // reductions per workgroup

struct PowerArgs {
  const BlockDesc *desc;
  int num_desc;
  const int32_t *row_ptr;
  const int32_t *col;
  double *hub_partial;
  const int32_t *hub_row;
  const int32_t *hub_nchunks;
  const int32_t *hub_first;
  int num_hubs;
  const uint32_t *rows;  // out-edge CSR: the node of each row (null: row = node)
  const double *yin;  // the gathered vector (Katz: picked by PowerState::par)
  double *x;          // Katz: the iterate; eigenvector: the iterate (its own term)
  double *t;          // eigenvector, HITS: the row sums
  double *partials;
  PowerState *st;
};

// f64 gathers of an unweighted graph (the staged product is 1.0 * y, exact)
struct GatherF64 {
  using P = double;
  static constexpr int kStageUnroll = kUnroll / 2;
  const double *y;
  const int32_t *col;
  __device__ __forceinline__ P stage_weight(int) const { return 1.0; }
  __device__ __forceinline__ P stage_gather(int c) const { return y[c]; }
  __device__ __forceinline__ double operator()(int k) const {
    return y[__builtin_nontemporal_load(col + k)];
  }
};

// katz.rs:166-174: x'[v] = beta + sum over in-edges of alpha * x[u]
struct KatzEpi {
  const PowerArgs &a;
  double alpha, beta;
  double *yout;
  __device__ __forceinline__ void operator()(int r, double acc, double (&red)[kPowR]) const {
    const double t = beta + acc;
    red[0] += fabs(a.x[r] - t);
    a.x[r] = t;
    yout[r] = alpha * t;
  }
};

// eigenvector.rs:76-84: t[v] = x[v] + sum over v's neighbours of x[u]
struct EigenEpi {
  const PowerArgs &a;
  __device__ __forceinline__ void operator()(int r, double acc, double (&red)[kPowR]) const {
    const int v = a.rows ? (int)a.rows[r] : r;
    const double t = a.x[v] + acc;
    a.t[v] = t;
    red[0] += t * t;
    red[1] += t;
  }
};

// centrality.rs:136-152: t[v] = sum of the gathered vector over the row
struct SumEpi {
  const PowerArgs &a;
  __device__ __forceinline__ void operator()(int r, double acc, double (&red)[kPowR]) const {
    a.t[a.rows ? (int)a.rows[r] : r] = acc;
    red[0] += acc * acc;
  }
};

__device__ __forceinline__ void publish(PowerState *st, double diff, bool conv) {
  const uint64_t it = st->iter + 1;
  const bool done = conv || it >= st->max_iter;
  st->iter = it;
  st->diff = diff;
  st->converged = conv;
  st->done = done;
  if (st->mirror) {
    __hip_atomic_store(st->mirror, (it << 2) | (uint64_t)(conv ? 2 : 0) | (uint64_t)done,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// body(epilogue, gathered vector) with the algorithm's epilogue
template <int ALGO, typename F>
__device__ __forceinline__ void with_epilogue(const PowerArgs &a, double *y0, double *y1,
                                              const F &body) {
  const PowerState *st = a.st;
  if constexpr (ALGO == kKatz) {
    const KatzEpi epi{a, st->alpha, st->beta, st->par ? y0 : y1};
    body(epi, st->par ? y1 : y0);
  } else if constexpr (ALGO == kEigen) {
    body(EigenEpi{a}, a.yin);
  } else {
    body(SumEpi{a}, a.yin);
  }
}

// One pull over a CSR (in-edges or out-edges) with the algorithm's row epilogue.
template <int ALGO>
__global__ __launch_bounds__(kBlock) void power_sweep_kernel(PowerArgs a, double *y0, double *y1) {
  if (a.st->done) return;
  __shared__ double stage[kStageNnz];
  double red[kPowR] = {0.0, 0.0};
  with_epilogue<ALGO>(a, y0, y1, [&](const auto &epi, const double *yin) {
    const GatherF64 load{yin, a.col};
    for (int i = blockIdx.x; i < a.num_desc; i += gridDim.x)
      do_desc(a, a.desc[i], load, epi, red, stage);
  });
  block_sum<kPowR>(red, stage);
  if (threadIdx.x == 0) {
    a.partials[(size_t)blockIdx.x * kPowR] = red[0];
    a.partials[(size_t)blockIdx.x * kPowR + 1] = red[1];
  }
}

// One workgroup: hub-row epilogues (chunks summed in chunk order), the
// fixed-order sum of the sweep's partials, and the step's bookkeeping.
template <int ALGO>
__global__ __launch_bounds__(kBlock) void power_finish_kernel(PowerArgs a, double *y0, double *y1,
                                                              int sweep_blocks) {
  PowerState *st = a.st;
  if (st->done) return;
  double red[kPowR] = {0.0, 0.0};
  with_epilogue<ALGO>(a, y0, y1, [&](const auto &epi, const double *) {
    for (int h = threadIdx.x; h < a.num_hubs; h += kBlock) {
      const int first = a.hub_first[h], nch = a.hub_nchunks[h];
      double s = 0.0;
      for (int c = 0; c < nch; c++) s += a.hub_partial[first + c];
      epi(a.hub_row[h], s, red);
    }
  });
  for (int b = threadIdx.x; b < sweep_blocks; b += kBlock) {
    red[0] += a.partials[(size_t)b * kPowR];
    red[1] += a.partials[(size_t)b * kPowR + 1];
  }
  block_sum<kPowR>(red);
  if (threadIdx.x == 0) {
    if constexpr (ALGO == kKatz) {  // katz.rs:187-192
      st->par ^= 1;
      publish(st, red[0], red[0] < st->tolerance);
    } else {  // eigenvector.rs:87-94, centrality.rs:168-175
      st->norm = sqrt(red[0]);
      st->sign = (ALGO == kEigen && !(red[1] >= 0.0)) ? -1.0 : 1.0;
    }
  }
}

// x' = t * sign / norm (t itself when norm is not > 0); partial residuals:
// sum |x - x'| (L1) or sum (x - x')^2 (L2).  x is overwritten with x'.
template <bool L2>
__global__ __launch_bounds__(kBlock) void power_scale_kernel(const double *t, double *x, int n,
                                                             const PowerState *st,
                                                             double *partials) {
  if (st->done) return;
  const double norm = st->norm, sign = st->sign;
  const bool scale = norm > 0.0;
  double red[1] = {0.0};
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
    const double nv = scale ? t[v] * sign / norm : t[v];
    const double dv = x[v] - nv;
    red[0] += L2 ? dv * dv : fabs(dv);
    x[v] = nv;
  }
  block_sum<1>(red);
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// One workgroup: the residual of a scale step and the iteration's end.
//   ALGO kEigen: diff = sum |x - x'| (eigenvector.rs:101-116)
//   ALGO kHits, stage 0: the authority change, kept for stage 1
//   ALGO kHits, stage 1: the hub change; stop when both are below tol
//   (centrality.rs:156-162)
template <int ALGO>
__global__ __launch_bounds__(kBlock) void power_step_kernel(PowerState *st, const double *partials,
                                                            int count, int stage) {
  if (st->done) return;
  double red[1] = {0.0};
  for (int b = threadIdx.x; b < count; b += kBlock) red[0] += partials[b];
  block_sum<1>(red);
  if (threadIdx.x == 0) {
    if constexpr (ALGO == kEigen) {
      publish(st, red[0], red[0] < st->tolerance);
    } else if (stage == 0) {
      st->diff2 = sqrt(red[0]);
    } else {
      const double hd = sqrt(red[0]);
      publish(st, hd, st->diff2 < st->tolerance && hd < st->tolerance);
    }
  }
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double *v, int n, double value) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = value;
}

__global__ __launch_bounds__(kBlock) void permute_kernel(const double *src, const uint32_t *rank,
                                                         double *out, int n) {
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock)
    out[o] = src[rank[o]];
}

// ---------------------------------------------------------------------------
// Blocked Katz sweeps.  Rows [0, n_in) have in-edges; a row without in-edges
// keeps x = beta (its sum is 0 every sweep: katz.rs:166-174 gives
// beta + nothing), so the sweep leaves it and its prescaled values, set to
// alpha * beta in both buffers at the start, alone; it adds |beta - beta| = 0
// to the residual.  Row sums are added as 128-bit fixed point in units of
// 2^-fix: every run sum and row sum is at most alpha * (max in-degree) *
// max|x| <= alpha * (max in-degree) * sum|x| < 2^(125 - fix), so none wraps,
// and one unit is at least 2^-100 of that bound's largest row, far below the
// f64 rounding of the reference's sequential sum.
struct KatzPbArgs {
  BlockedLayout pb;
  int n;
  int n_in;
  double *x;          // the iterate
  double *y0, *y1;    // alpha * x, ping-pong: the sweep gathers y1 when st->par, else y0
  double *partials;   // [2][items2 + splits]: sum |x - x'| and sum |x'| per workgroup
  PowerState *st;
  double bound_scale; // alpha * the largest in-degree (duplicate entries counted)
};

template <typename WT>
__global__ __launch_bounds__(kScatterBlock) __attribute__((amdgpu_waves_per_eu(kScatterBlock / 256, kScatterBlock / 256)))
void katz_pb_scatter_kernel(KatzPbArgs a) {
  const PowerState *st = a.st;
  if (st->done) return;
  __shared__ double z[1 << kPbChunkShift];
  const PbItem it = a.pb.items1[blockIdx.x];
  pb_load_chunk64(st->par ? a.y1 : a.y0, a.n, it, z);
  pb_scatter64_windows<WT>(a.pb, it, z);
}

// katz.rs:166-174 and 180-185: x' = beta + sum, |x - x'| and |x'|
__device__ __forceinline__ void katz_pb_row(const KatzPbArgs &a, const PowerState *st, int r,
                                            double sum, double (&red)[kPowR]) {
  const double t = st->beta + sum;
  red[0] += fabs(a.x[r] - t);
  red[1] += fabs(t);
  a.x[r] = t;
  (st->par ? a.y0 : a.y1)[r] = st->alpha * t;
}

__device__ __forceinline__ double *katz_pb_partial(const KatzPbArgs &a, int i, int part) {
  return a.partials + (size_t)i * (size_t)(a.pb.num_items2 + a.pb.num_splits) + part;
}

constexpr int kKatzAccBlock = 1024;
__global__ __launch_bounds__(kKatzAccBlock) void katz_pb_accumulate_kernel(KatzPbArgs a) {
  const PowerState *st = a.st;
  if (st->done) return;
  constexpr int kRows = 1 << kPbBinShift;
  __shared__ unsigned long long acc_lo[kRows], acc_hi[kRows];
  __shared__ uint32_t nonfin[kRows / 4];
  const PbItem it = a.pb.items2[blockIdx.x];
  const int row0 = a.pb.bin_row0[it.x];
  if (row0 >= a.n_in) {  // rows without in-edges: nothing to add, nothing changes
    if (threadIdx.x < kPowR) *katz_pb_partial(a, threadIdx.x, blockIdx.x) = 0.0;
    return;
  }
  const int rows = a.pb.bin_row0[it.x + 1] - row0;
  for (int i = threadIdx.x; i < kRows; i += kKatzAccBlock) {
    acc_lo[i] = 0ull;
    acc_hi[i] = 0ull;
  }
  for (int i = threadIdx.x; i < kRows / 4; i += kKatzAccBlock) nonfin[i] = 0u;
  __syncthreads();
  const int fix = st->fix;
  pb_accumulate128_runs<kKatzAccBlock>(a.pb, it, fix, acc_lo, acc_hi, nonfin);
  __syncthreads();
  double red[kPowR] = {0.0, 0.0};
  if (it.w < 0) {
    const int live = min(rows, a.n_in - row0);
    for (int i = threadIdx.x; i < live; i += kKatzAccBlock) {
      const uint32_t fl = (nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu;
      katz_pb_row(a, st, row0 + i, pb_row_value128(acc_lo[i], acc_hi[i], fl, fix), red);
    }
  } else {  // a part of a split bin: its sums for the combine kernel
    long long *out = a.pb.bin_partial + ((size_t)it.w << (kPbBinShift + 1));
    uint8_t *fout = a.pb.bin_flags + ((size_t)it.w << kPbBinShift);
    for (int i = threadIdx.x; i < rows; i += kKatzAccBlock) {
      out[2 * i] = (long long)acc_lo[i];
      out[2 * i + 1] = (long long)acc_hi[i];
      fout[i] = (uint8_t)((nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
  }
  __syncthreads();
  block_sum_waves<kPowR, kKatzAccBlock / 64>(red, reinterpret_cast<double *>(acc_lo));
  if (threadIdx.x == 0) {
    *katz_pb_partial(a, 0, blockIdx.x) = red[0];
    *katz_pb_partial(a, 1, blockIdx.x) = red[1];
  }
}

// kBlock rows of a bin split over several accumulate workgroups
__global__ __launch_bounds__(kBlock) void katz_pb_combine_kernel(KatzPbArgs a) {
  const PowerState *st = a.st;
  if (st->done) return;
  const PbItem sp = a.pb.splits[blockIdx.x];  // {bin, first slot, slots, first row of the bin}
  const int r = a.pb.bin_row0[sp.x] + sp.w + (int)threadIdx.x;
  double red[kPowR] = {0.0, 0.0};
  if (r < a.pb.bin_row0[sp.x + 1] && r < a.n_in) {
    unsigned long long lo, hi;
    uint32_t fl;
    pb_combine128_row(a.pb, sp, sp.w + (int)threadIdx.x, lo, hi, fl);
    katz_pb_row(a, st, r, pb_row_value128(lo, hi, fl, st->fix), red);
  }
  block_sum<kPowR>(red);
  if (threadIdx.x == 0) {
    *katz_pb_partial(a, 0, a.pb.num_items2 + blockIdx.x) = red[0];
    *katz_pb_partial(a, 1, a.pb.num_items2 + blockIdx.x) = red[1];
  }
}

// Units of the row sums of a sweep whose input vector sums to xsum in
// absolute value: 2^-fix with alpha * (max in-degree) * xsum < 2^(125 - fix).
__host__ __device__ inline int katz_fix(double bound_scale, double xsum) {
  const double b = bound_scale * xsum;
  if (!(b > 0.0) || !(b <= 1.7e308)) return 100;
  int e = 0;
  (void)frexp(b, &e);  // b < 2^e
  const int fix = 125 - e;
  return fix < -2000 ? -2000 : (fix > 2000 ? 2000 : fix);
}

// One workgroup: the sweep's partials in a fixed order, the stop rule of
// katz.rs:180-192, and the next sweep's units (the rows without in-edges add
// (n - n_in) |beta| to sum |x|).
__global__ __launch_bounds__(kBlock) void katz_pb_finish_kernel(KatzPbArgs a) {
  PowerState *st = a.st;
  if (st->done) return;
  const int parts = a.pb.num_items2 + a.pb.num_splits;
  double red[kPowR] = {0.0, 0.0};
  for (int b = threadIdx.x; b < parts; b += kBlock) {
    red[0] += *katz_pb_partial(a, 0, b);
    red[1] += *katz_pb_partial(a, 1, b);
  }
  block_sum<kPowR>(red);
  if (threadIdx.x == 0) {
    st->fix = katz_fix(a.bound_scale, red[1] + (double)(a.n - a.n_in) * fabs(st->beta));
    st->par ^= 1;
    publish(st, red[0], red[0] < st->tolerance);
  }
}

// ---------------------------------------------------------------------------
// Blocked eigenvector and HITS passes: t = the sums of a unit-weight blocked
// layout (the in-edge layout: HITS authority; the out-edge CSR's: eigenvector
// and the HITS hub, rows mapped to nodes), the f64 scatter and the 128-bit
// accumulate of the anchors and of Katz.  The gathered vector has unit L2
// norm (or 1/sqrt(n) in every node at the start), so no entry exceeds 1 and
// no row sum exceeds the layout's largest row length: the units 2^-fix are
// fixed for the run with that bound below 2^(124 - fix).  Every row is
// written, rows without entries too (their sums are 0).
struct PowerPbArgs {
  BlockedLayout pb;
  int n;
  const uint32_t *rows;  // the node of each row (out-edge CSR), or null (row = node)
  const double *yin;     // the gathered vector
  const double *x;       // eigenvector: the iterate (its own term)
  double *t;             // the row sums, by node
  double *partials;      // [2][items2 + splits]
  PowerState *st;
  int fix;
};

template <typename WT>
__global__ __launch_bounds__(kScatterBlock) __attribute__((amdgpu_waves_per_eu(kScatterBlock / 256, kScatterBlock / 256)))
void power_pb_scatter_kernel(PowerPbArgs a) {
  if (a.st->done) return;
  __shared__ double z[1 << kPbChunkShift];
  const PbItem it = a.pb.items1[blockIdx.x];
  pb_load_chunk64(a.yin, a.n, it, z);
  pb_scatter64_windows<WT>(a.pb, it, z);
}

// eigenvector.rs:76-84 (t[v] = x[v] + sum; sum t^2, sum t) and
// centrality.rs:136-152 (t[v] = sum; sum t^2), as EigenEpi and SumEpi
template <int ALGO>
__device__ __forceinline__ void power_pb_row(const PowerPbArgs &a, int r, double sum,
                                             double (&red)[kPowR]) {
  const int v = a.rows ? (int)a.rows[r] : r;
  const double t = ALGO == kEigen ? a.x[v] + sum : sum;
  a.t[v] = t;
  red[0] += t * t;
  red[1] += t;
}

__device__ __forceinline__ double *power_pb_partial(const PowerPbArgs &a, int i, int part) {
  return a.partials + (size_t)i * (size_t)(a.pb.num_items2 + a.pb.num_splits) + part;
}

template <int ALGO>
__global__ __launch_bounds__(kKatzAccBlock) void power_pb_accumulate_kernel(PowerPbArgs a) {
  if (a.st->done) return;
  constexpr int kRows = 1 << kPbBinShift;
  __shared__ unsigned long long acc_lo[kRows], acc_hi[kRows];
  __shared__ uint32_t nonfin[kRows / 4];
  const PbItem it = a.pb.items2[blockIdx.x];
  const int row0 = a.pb.bin_row0[it.x];
  const int rows = a.pb.bin_row0[it.x + 1] - row0;
  for (int i = threadIdx.x; i < kRows; i += kKatzAccBlock) {
    acc_lo[i] = 0ull;
    acc_hi[i] = 0ull;
  }
  for (int i = threadIdx.x; i < kRows / 4; i += kKatzAccBlock) nonfin[i] = 0u;
  __syncthreads();
  pb_accumulate128_runs<kKatzAccBlock>(a.pb, it, a.fix, acc_lo, acc_hi, nonfin);
  __syncthreads();
  double red[kPowR] = {0.0, 0.0};
  if (it.w < 0) {
    for (int i = threadIdx.x; i < rows; i += kKatzAccBlock) {
      const uint32_t fl = (nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu;
      power_pb_row<ALGO>(a, row0 + i, pb_row_value128(acc_lo[i], acc_hi[i], fl, a.fix), red);
    }
  } else {
    long long *out = a.pb.bin_partial + ((size_t)it.w << (kPbBinShift + 1));
    uint8_t *fout = a.pb.bin_flags + ((size_t)it.w << kPbBinShift);
    for (int i = threadIdx.x; i < rows; i += kKatzAccBlock) {
      out[2 * i] = (long long)acc_lo[i];
      out[2 * i + 1] = (long long)acc_hi[i];
      fout[i] = (uint8_t)((nonfin[i >> 2] >> (8 * (i & 3))) & 0xffu);
    }
  }
  __syncthreads();
  block_sum_waves<kPowR, kKatzAccBlock / 64>(red, reinterpret_cast<double *>(acc_lo));
  if (threadIdx.x == 0) {
    *power_pb_partial(a, 0, blockIdx.x) = red[0];
    *power_pb_partial(a, 1, blockIdx.x) = red[1];
  }
}

template <int ALGO>
__global__ __launch_bounds__(kBlock) void power_pb_combine_kernel(PowerPbArgs a) {
  if (a.st->done) return;
  const PbItem sp = a.pb.splits[blockIdx.x];  // {bin, first slot, slots, first row of the bin}
  const int r = a.pb.bin_row0[sp.x] + sp.w + (int)threadIdx.x;
  double red[kPowR] = {0.0, 0.0};
  if (r < a.pb.bin_row0[sp.x + 1]) {
    unsigned long long lo, hi;
    uint32_t fl;
    pb_combine128_row(a.pb, sp, sp.w + (int)threadIdx.x, lo, hi, fl);
    power_pb_row<ALGO>(a, r, pb_row_value128(lo, hi, fl, a.fix), red);
  }
  block_sum<kPowR>(red);
  if (threadIdx.x == 0) {
    *power_pb_partial(a, 0, a.pb.num_items2 + blockIdx.x) = red[0];
    *power_pb_partial(a, 1, a.pb.num_items2 + blockIdx.x) = red[1];
  }
}

// One workgroup: the pass's partials in a fixed order; the norm and sign of t
// (eigenvector.rs:87-94, centrality.rs:154-155), as power_finish_kernel.
template <int ALGO>
__global__ __launch_bounds__(kBlock) void power_pb_finish_kernel(PowerPbArgs a) {
  PowerState *st = a.st;
  if (st->done) return;
  const int parts = a.pb.num_items2 + a.pb.num_splits;
  double red[kPowR] = {0.0, 0.0};
  for (int b = threadIdx.x; b < parts; b += kBlock) {
    red[0] += *power_pb_partial(a, 0, b);
    red[1] += *power_pb_partial(a, 1, b);
  }
  block_sum<kPowR>(red);
  if (threadIdx.x == 0) {
    st->norm = sqrt(red[0]);
    st->sign = (ALGO == kEigen && !(red[1] >= 0.0)) ? -1.0 : 1.0;
  }
}

static int elementwise_blocks(uint64_t n) {
  uint64_t blocks = (n + kBlock * 8 - 1) / (kBlock * 8);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 2048));
}

// rows: the node of each row of dg (the out-edge CSR's order), or null
static PowerArgs power_args(gops_graph *g, const DevGraph &dg, const uint32_t *rows) {
  PowerArgs a{};
  a.desc = dg.desc;
  a.num_desc = dg.num_desc;
  a.row_ptr = dg.row_ptr;
  a.col = dg.col;
  a.hub_partial = g->pow.hub_partial;
  a.hub_row = dg.hub_row;
  a.hub_nchunks = dg.hub_nchunks;
  a.hub_first = dg.hub_first;
  a.num_hubs = dg.num_hubs;
  a.rows = rows;
  a.partials = g->pow.partials;
  a.st = g->pow.state;
  return a;
}

template <int ALGO>
static void sweep(gops_context *ctx, const PowerArgs &a, double *y0, double *y1) {
  const int grid = std::min(a.num_desc, kMaxSweepBlocks);
  power_sweep_kernel<ALGO><<<grid, kBlock, 0, ctx->stream>>>(a, y0, y1);
  power_finish_kernel<ALGO><<<1, kBlock, 0, ctx->stream>>>(a, y0, y1, grid);
}

// units of a blocked pass's row sums: entries of at most 1, rows of at most
// maxdeg entries
static int unit_fix(uint64_t maxdeg) {
  int e = 0;
  (void)frexp((double)std::max<uint64_t>(maxdeg, 1), &e);  // maxdeg < 2^e
  return 124 - e;
}

template <int ALGO>
static void pb_pass(gops_context *ctx, gops_graph *g, const BlockedLayout &lay, uint64_t maxdeg,
                    const uint32_t *rows, const double *yin, const double *x, double *t) {
  PowerPbArgs a{};
  a.pb = lay;
  a.n = (int)g->dev.n;
  a.rows = rows;
  a.yin = yin;
  a.x = x;
  a.t = t;
  a.partials = lay.partials;
  a.st = g->pow.state;
  a.fix = unit_fix(maxdeg);
  if (lay.num_items1)
    power_pb_scatter_kernel<NoW><<<lay.num_items1, kScatterBlock, 0, ctx->stream>>>(a);
  power_pb_accumulate_kernel<ALGO><<<lay.num_items2, kKatzAccBlock, 0, ctx->stream>>>(a);
  if (lay.num_splits)
    power_pb_combine_kernel<ALGO><<<lay.num_splits, kBlock, 0, ctx->stream>>>(a);
  power_pb_finish_kernel<ALGO><<<1, kBlock, 0, ctx->stream>>>(a);
}

template <int ALGO, bool L2>
static void scale_step(gops_context *ctx, gops_graph *g, const double *t, double *x, int stage) {
  const int n = (int)g->dev.n;
  const int grid = elementwise_blocks(n);
  PowerBuffers &pb = g->pow;
  power_scale_kernel<L2><<<grid, kBlock, 0, ctx->stream>>>(t, x, n, pb.state, pb.partials);
  power_step_kernel<ALGO><<<1, kBlock, 0, ctx->stream>>>(pb.state, pb.partials, grid, stage);
}

// Buffer roles: Katz v0 = x, v1/v2 = alpha*x (ping-pong); eigenvector
// v0 = x, v1 = t; HITS v0 = hub, v2 = authority, v1 = t.  value: the start
// value of every node; prescaled: Katz's first gathered value, alpha * beta.
hipError_t launch_power_init(gops_context *ctx, gops_graph *g, PowerAlgo algo, double value,
                             double prescaled) {
  const int n = (int)g->dev.n;
  const int grid = elementwise_blocks(n);
  PowerBuffers &pb = g->pow;
  fill_kernel<<<grid, kBlock, 0, ctx->stream>>>(pb.v[0], n, value);
  if (algo == kKatz) {
    fill_kernel<<<grid, kBlock, 0, ctx->stream>>>(pb.v[1], n, prescaled);
    // blocked sweeps leave the rows without in-edges in both buffers
    if (pb.katz_blocked) fill_kernel<<<grid, kBlock, 0, ctx->stream>>>(pb.v[2], n, prescaled);
  }
  if (algo == kHits) fill_kernel<<<grid, kBlock, 0, ctx->stream>>>(pb.v[2], n, value);
  return hipGetLastError();
}

int katz_initial_fix(double alpha, uint64_t indeg_max, uint64_t n, double beta) {
  return katz_fix(alpha * (double)indeg_max, (double)n * fabs(beta));
}

hipError_t launch_power_iteration(gops_context *ctx, gops_graph *g, PowerAlgo algo) {
  PowerBuffers &pb = g->pow;
  if (algo == kKatz && pb.katz_blocked) {
    KatzPbArgs a{};
    a.pb = g->pb;
    a.n = (int)g->dev.n;
    a.n_in = (int)pb.n_in;
    a.x = pb.v[0];
    a.y0 = pb.v[1];
    a.y1 = pb.v[2];
    a.partials = g->pb.partials;
    a.st = pb.state;
    a.bound_scale = pb.katz_bound_scale;
    if (a.pb.num_items1)
      katz_pb_scatter_kernel<NoW><<<a.pb.num_items1, kScatterBlock, 0, ctx->stream>>>(a);
    katz_pb_accumulate_kernel<<<a.pb.num_items2, kKatzAccBlock, 0, ctx->stream>>>(a);
    if (a.pb.num_splits)
      katz_pb_combine_kernel<<<a.pb.num_splits, kBlock, 0, ctx->stream>>>(a);
    katz_pb_finish_kernel<<<1, kBlock, 0, ctx->stream>>>(a);
  } else if (algo == kKatz) {
    PowerArgs a = power_args(g, g->dev, nullptr);
    a.x = pb.v[0];
    sweep<kKatz>(ctx, a, pb.v[1], pb.v[2]);
  } else if (algo == kEigen && pb.power_blocked) {
    pb_pass<kEigen>(ctx, g, g->pb_fwd, pb.outdeg_max, g->fwd.order, pb.v[0], pb.v[0], pb.v[1]);
    scale_step<kEigen, false>(ctx, g, pb.v[1], pb.v[0], 0);
  } else if (algo == kHits && pb.power_blocked) {
    // authority: in-edges, gathers hub; hub: out-edges, gathers the new authority
    pb_pass<kHits>(ctx, g, g->pb, pb.indeg_max, nullptr, pb.v[0], nullptr, pb.v[1]);
    scale_step<kHits, true>(ctx, g, pb.v[1], pb.v[2], 0);
    pb_pass<kHits>(ctx, g, g->pb_fwd, pb.outdeg_max, g->fwd.order, pb.v[2], nullptr, pb.v[1]);
    scale_step<kHits, true>(ctx, g, pb.v[1], pb.v[0], 1);
  } else if (algo == kEigen) {
    PowerArgs a = power_args(g, g->fwd, g->fwd.order);
    a.yin = pb.v[0];
    a.x = pb.v[0];
    a.t = pb.v[1];
    sweep<kEigen>(ctx, a, nullptr, nullptr);
    scale_step<kEigen, false>(ctx, g, pb.v[1], pb.v[0], 0);
  } else {
    PowerArgs a = power_args(g, g->dev, nullptr);  // authority: in-edges, gathers hub
    a.yin = pb.v[0];
    a.t = pb.v[1];
    sweep<kHits>(ctx, a, nullptr, nullptr);
    scale_step<kHits, true>(ctx, g, pb.v[1], pb.v[2], 0);
    PowerArgs h = power_args(g, g->fwd, g->fwd.order);  // hub: out-edges, gathers the new authority
    h.yin = pb.v[2];
    h.t = pb.v[1];
    sweep<kHits>(ctx, h, nullptr, nullptr);
    scale_step<kHits, true>(ctx, g, pb.v[1], pb.v[0], 1);
  }
  return hipGetLastError();
}

hipError_t launch_permute(gops_context *ctx, gops_graph *g, const double *src, double *out) {
  const int n = (int)g->dev.n;
  permute_kernel<<<elementwise_blocks(n), kBlock, 0, ctx->stream>>>(src, g->dev.rank, out, n);
  return hipGetLastError();
}

}  // namespace gops
