// ppr_batch.hip -- batched personalized PageRank: many independent
// personalized_pagerank_run calls (ppr.rs:90-156) over one resident graph.
//
// A tile holds T queries.  Every per-node vector of the tile is stored
// row-major [n][T], so one gather of a source row fetches the T queries'
// values of that source contiguously (T = 16: one 64-byte line for 16 queries,
// where a single-vector gather uses 4 bytes of the line it fetches).  Each
// sweep streams the CSR of in-edges once for the whole tile.
//
// Arithmetic per query is the single-vector sweep chain of pagerank.hip
// (delta scheme: f32 gathers, f64 iterate, f64 anchor sweeps, mass
// renormalisation at anchors), with per-query damping, tolerance, iteration
// cap, residual, stop and sweep mode: lanes of different queries may run
// different modes in one launch, so a query's arithmetic does not depend on
// the other queries of its tile.  A query that has stopped keeps its result
// and its lanes idle.
// Personalization is sparse: a per-node index (seed_idx) marks the rows that
// are seeds of some query of the tile, and only those rows read the tile's
// seed entries for the teleport / dangling terms (ppr.rs:125-128), which are
// zero everywhere else.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "ppr_batch.h"

namespace gops {

namespace {

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
constexpr int kSwzXor16 = 0x1F | (0x10 << 10);

__device__ __forceinline__ double xor_lanes(double v, int mask) {
  // lane ^ mask for mask in {1, 2, 16, 32} (the masks the tile layouts use)
  if (mask == 16) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), kSwzXor16);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), kSwzXor16);
    return __hiloint2double(hi, lo);
  }
  return __shfl_xor(v, mask);
}

// Sum over the lanes whose index differs only in bits [lo_bit, hi_bit).
__device__ __forceinline__ double lane_bits_sum(double v, int lo_bit, int hi_bit) {
  for (int b = lo_bit; b < hi_bit; b++) v += xor_lanes(v, 1 << b);
  return v;
}

// Per-lane partial sums of a sweep, per query: L1 residual, dangling part of
// the delta, dangling part of the iterate, total delta, total iterate.  Named
// fields, each updated once per epilogue: conditional adds in different
// branches get merged into one add through a select between field addresses,
// which keeps the partial sums in scratch memory instead of registers.
struct Red {
  double l1 = 0.0, dang = 0.0, dang_x = 0.0, mass = 0.0, sum_x = 0.0;
};

}  // namespace

// The prescaled deltas a sweep gathers are stored in f32; what a value
// rounds off goes into the row's carry, so it reaches the next sweep
// instead of being lost.
__device__ __forceinline__ float delta_put(double v) { return (float)v; }

// What a sweep's epilogue needs of its query, read from the tile state once
// per thread (lane q of the epilogue always handles query q): reloading it
// for every row would put a chain of global loads on each row's path.
struct QState {
  bool live, first, f64_phase;
  int c64;  // this sweep's anchors gather y64[c64]; the next gathers the other
  double d, x_scale, dsum_x, dsum_in, mass_in;
};

__device__ __forceinline__ QState load_qstate(const TileState *ts, int q) {
  QState z;
  z.live = ts->live[q] != 0;
  z.first = ts->iter[q] == 0;
  z.f64_phase = ts->f64_phase[q] != 0;
  z.c64 = ts->y64_par;
  z.d = ts->damping[q];
  z.x_scale = ts->x_scale[q];
  z.dsum_x = ts->dsum_x[q];
  z.dsum_in = ts->dsum_in[q];
  z.mass_in = ts->mass_in[q];
  return z;
}

// The loads of one (row, query) epilogue, issued together for all of a
// thread's rows before any of them is finished (ppr_epilogue).
struct EpiLoad {
  int sidx = -1;
  double inv = 0.0, x = 0.0;
  float c = 0.0f;
};

template <int T>
__device__ __forceinline__ EpiLoad epi_load(const PprArgs &a, const QState &qs, int mode, int r,
                                            int q) {
  EpiLoad ld;
  const size_t at = (size_t)r * T + q;
  ld.sidx = a.seed_idx[r];
  ld.inv = a.inv64[r];
  // a query's first delta sweep reads neither x nor the carry (ppr_epilogue)
  if (mode == kModeAnchor || !qs.first) ld.x = __builtin_nontemporal_load(a.x + at);
  if (mode != kModeAnchor && !qs.first) ld.c = a.carry[at];
  return ld;
}

// Per (row, query) epilogue.  mode: kModeDelta / kModeDeltaFold / kModeAnchor.
// acc is the row's gathered sum; ld holds the row's loads (epi_load).
template <int T>
__device__ __forceinline__ void ppr_epilogue(const PprArgs &a, const QState &qs, int mode,
                                             int r, int q, double acc, const EpiLoad &ld,
                                             Red &red) {
  if (!qs.live) return;
  const double d = qs.d;
  double seed_p = 0.0;
  const int sidx = ld.sidx;
  if (sidx >= 0) {
    for (int j = sidx; j < a.num_seed_entries && a.seed_row[j] == r; j++)
      if (a.seed_q[j] == q) seed_p += a.seed_p[j];
  }
  const size_t at = (size_t)r * T + q;
  const int c64 = qs.c64;
  float *yout = a.yout32;
  // every path sets all five contributions; one update of red at the end
  double delta, dang_x = 0.0, sum_x = 0.0;
  bool dangling;
  if (mode == kModeAnchor) {
    const double sc = qs.x_scale;
    const double dsum = qs.dsum_x * sc;
    const double t = sc * acc + ((1.0 - d) * seed_p + (d * dsum) * seed_p);  // ppr.rs:126-127
    delta = t - sc * ld.x;
    a.x[at] = t;
    const double inv = ld.inv;
    const double v = (d * delta) * inv;  // the delta chain restarts here: no carry
    const float yo = delta_put(v);
    yout[at] = yo;
    a.carry[at] = (float)(v - (double)yo);
    if (qs.f64_phase) (c64 ? a.y64a : a.y64b)[at] = (d * t) * inv;  // next anchor's input
    dangling = inv == 0.0;
    dang_x = dangling ? t : 0.0;
  } else {
    // delta' = A delta: the dangling and (removed) total sums go to the seeds
    delta = acc + d * (qs.dsum_in - qs.mass_in) * seed_p;
    const double inv = ld.inv;
    // A query's first sweep (admitted with p-hat as its input delta) starts
    // from x = p-hat and the carry of p-hat's rounding on its seeds, 0
    // elsewhere: neither column is read (nor cleared at admission) for it.
    const bool first = qs.first;
    double x_old, c_old;
    if (first) {
      x_old = seed_p;
      const double v0 = (d * seed_p) * inv;
      c_old = v0 - (double)delta_put(v0);
    } else {
      x_old = ld.x;
      c_old = (double)ld.c;
    }
    const double x = x_old + delta;
    __builtin_nontemporal_store(x, a.x + at);
    // the stored delta carries the rounding of the row's previous one:
    // summed over the sweeps, the roundings cancel but for the last
    const double v = (d * delta) * inv + c_old;
    const float yo = delta_put(v);
    yout[at] = yo;
    a.carry[at] = (float)(v - (double)yo);
    dangling = inv == 0.0;
    if (mode == kModeDeltaFold) {
      (c64 ? a.y64a : a.y64b)[at] = (d * x) * inv;  // the anchor that follows
      dang_x = dangling ? x : 0.0;
      sum_x = x;
    }
  }
  red.l1 += fabs(delta);
  red.dang += dangling ? delta : 0.0;
  red.dang_x += dang_x;
  red.mass += delta;
  red.sum_x += sum_x;
}

// Per-query partial sums of a workgroup (lane q = tid % T holds query q)
// -> out[q * 5 + i] in thread q < T.
template <int T>
__device__ __forceinline__ void tile_block_sum(Red &red, double *out) {
  __shared__ double sh[kBlock / 64][T][5];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int lb = 0;
  while ((1 << lb) < T) lb++;
  const double s0 = lane_bits_sum(red.l1, lb, 6), s1 = lane_bits_sum(red.dang, lb, 6),
               s2 = lane_bits_sum(red.dang_x, lb, 6), s3 = lane_bits_sum(red.mass, lb, 6),
               s4 = lane_bits_sum(red.sum_x, lb, 6);
  if (lane < T) {
    double *o = sh[wave][lane];
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
    o[3] = s3;
    o[4] = s4;
  }
  __syncthreads();
  if (threadIdx.x < T) {
#pragma unroll
    for (int i = 0; i < 5; i++)
      out[threadIdx.x * 5 + i] = (sh[0][threadIdx.x][i] + sh[1][threadIdx.x][i]) +
                                 (sh[2][threadIdx.x][i] + sh[3][threadIdx.x][i]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// The tile sweep.  Lane layout: LQ (kLaneQ) consecutive queries per lane,
// so Q = T / LQ lanes cover one edge's T queries with one 16-byte gather
// each; a wave gathers 64 / Q edges per load instruction.  A row gets 2^s slices of Q lanes (slice i sums edges
// i, i + 2^s, ...), and the slices are summed by lane shuffles.

// The LQ queries of a lane: sums, modes.
template <int LQ>
struct QV {
  double v[LQ];
};

template <int LQ>
struct LaneModes {
  bool a[LQ];           // query j of the lane is in an f64 anchor sweep
  bool need32, need64;  // some query of the lane gathers deltas / the f64 iterate
  bool live;            // some query of the lane is live (else the lane gathers nothing)
};

// LQ consecutive stored deltas, as loaded (one load per lane)
template <int LQ>
using DeltaVec = float __attribute__((ext_vector_type(LQ)));

template <int LQ>
__device__ __forceinline__ DeltaVec<LQ> load_deltas(const float *p) {
  return *reinterpret_cast<const DeltaVec<LQ> *>(p);
}

// acc += f in edge order, or nothing (f belongs to a missing edge)
template <int LQ, typename V>
__device__ __forceinline__ void add_deltas(QV<LQ> &acc, const V &f, bool have) {
#pragma unroll
  for (int j = 0; j < LQ; j++) acc.v[j] = have ? acc.v[j] + (double)(float)f[j] : acc.v[j];
}

template <int T, int LQ>
__device__ __forceinline__ void add_f64(const double *y64, int c, int q0, QV<LQ> &acc) {
  const double *p = y64 + (size_t)c * T + q0;
#pragma unroll
  for (int j = 0; j < LQ; j += 2) {
    const double2 v = *reinterpret_cast<const double2 *>(p + j);
    acc.v[j] += v.x;
    acc.v[j + 1] += v.y;
  }
}

// queries of the lane in different modes: each takes its own vector
template <int T, int LQ>
__device__ __forceinline__ void add_mixed(const float *yin, const double *y64, const LaneModes<LQ> &m,
                                          int c, int q0, QV<LQ> &acc) {
  const size_t off = (size_t)c * T + q0;
  const DeltaVec<LQ> f = load_deltas<LQ>(yin + off);
#pragma unroll
  for (int j = 0; j < LQ; j += 2) {
    const double2 v = *reinterpret_cast<const double2 *>(y64 + off + j);
    acc.v[j] += m.a[j] ? v.x : (double)(float)f[j];
    acc.v[j + 1] += m.a[j + 1] ? v.y : (double)(float)f[j + 1];
  }
}

// Sums of one lane's slice of a row: edges b, b + step, ... < e, added in
// edge order.  Delta lanes keep four edges' loads in flight, f64 and mixed
// lanes two.
template <int T, int LQ>
__device__ __forceinline__ QV<LQ> row_slice_sum(const PprArgs &a, const float *yin, const double *y64,
                                                const LaneModes<LQ> &m, int b, int e, int step,
                                                int q0) {
  QV<LQ> acc = {};
  if (!m.live) return acc;  // stopped queries and free slots: no loads
  int k = b;
  if (!m.need64) {
    // batches of four edges, the last one partial: every load of a batch is
    // issued before the first add (a slice of a short row would otherwise
    // wait out two dependent loads per edge); a missing edge repeats the
    // batch's first one, and its value is not added
    for (; k < e; k += 4 * step) {
      const int k1 = k + step, k2 = k + 2 * step, k3 = k + 3 * step;
      const bool h1 = k1 < e, h2 = k2 < e, h3 = k3 < e;
      const int c0 = a.col[k];
      const int c1 = a.col[h1 ? k1 : k], c2 = a.col[h2 ? k2 : k], c3 = a.col[h3 ? k3 : k];
      const DeltaVec<LQ> f0 = load_deltas<LQ>(yin + (size_t)c0 * T + q0);
      const DeltaVec<LQ> f1 = load_deltas<LQ>(yin + (size_t)c1 * T + q0);
      const DeltaVec<LQ> f2 = load_deltas<LQ>(yin + (size_t)c2 * T + q0);
      const DeltaVec<LQ> f3 = load_deltas<LQ>(yin + (size_t)c3 * T + q0);
      add_deltas<LQ>(acc, f0, true);
      add_deltas<LQ>(acc, f1, h1);
      add_deltas<LQ>(acc, f2, h2);
      add_deltas<LQ>(acc, f3, h3);
    }
  } else if (!m.need32) {
    if (LQ == 4) {  // 8 lanes' worth of f64 per edge at LQ = 8: one edge at a time
      for (; k + step < e; k += 2 * step) {
        const int c0 = a.col[k], c1 = a.col[k + step];
        add_f64<T>(y64, c0, q0, acc);
        add_f64<T>(y64, c1, q0, acc);
      }
    }
    for (; k < e; k += step) add_f64<T>(y64, a.col[k], q0, acc);
  } else {
    if (LQ == 4) {
      for (; k + step < e; k += 2 * step) {
        const int c0 = a.col[k], c1 = a.col[k + step];
        add_mixed<T>(yin, y64, m, c0, q0, acc);
        add_mixed<T>(yin, y64, m, c1, q0, acc);
      }
    }
    for (; k < e; k += step) add_mixed<T>(yin, y64, m, a.col[k], q0, acc);
  }
  return acc;
}

template <int LQ>
__device__ __forceinline__ void qv_lane_sum(QV<LQ> &v, int lo_bit, int hi_bit) {
#pragma unroll
  for (int j = 0; j < LQ; j++) v.v[j] = lane_bits_sum(v.v[j], lo_bit, hi_bit);
}

// Sum over the block's slices of Q lanes (threads with the same tid % Q);
// valid in threads < Q.
template <int T, int LQ>
__device__ __forceinline__ QV<LQ> block_slices_sum(QV<LQ> v) {
  constexpr int Q = T / LQ;
  __shared__ double sh[kBlock / 64][Q][LQ];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int lq = 0;
  while ((1 << lq) < Q) lq++;
  qv_lane_sum(v, lq, 6);
  if (lane < Q) {
#pragma unroll
    for (int j = 0; j < LQ; j++) sh[wave][lane][j] = v.v[j];
  }
  __syncthreads();
  QV<LQ> r = {};
  if (threadIdx.x < Q) {
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < LQ; j++) r.v[j] = (sh[0][t][j] + sh[1][t][j]) + (sh[2][t][j] + sh[3][t][j]);
  }
  __syncthreads();
  return r;
}

// Row sums leave the gather phase in the slice-0 lanes, LQ queries each;
// they go through LDS so that the epilogue runs one query per lane (lane
// q = tid % T always handles query q), with one query's state in registers.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_sweep_kernel(PprArgs a) {
  constexpr int LQ = kLaneQ;
  constexpr int Q = T / LQ;  // lanes per edge
  static_assert(Q >= 1 && kBlock % T == 0, "lane layout");
  __shared__ double rs[kBlock * LQ];  // row sums of a pass: rows (<= kBlock / Q) x T
  const TileState *ts = a.ts;
  if (ts->all_done) return;
  const int ycur = ts->yd_cur;
  a.yin32 = ycur ? a.yd1 : a.yd0;
  a.yout32 = ycur ? a.yd0 : a.yd1;
  const float *yin = a.yin32;
  const double *y64 = ts->y64_par ? a.y64b : a.y64a;
  // gather phase: queries q0 .. q0 + LQ - 1 of this lane
  const int qg = threadIdx.x % Q;
  const int q0 = qg * LQ;
  LaneModes<LQ> m;
  m.need64 = false;
  m.need32 = false;
  m.live = false;
#pragma unroll
  for (int j = 0; j < LQ; j++) {
    m.a[j] = ts->mode[q0 + j] == kModeAnchor;
    m.need64 = m.need64 || m.a[j];
    m.need32 = m.need32 || !m.a[j];
    m.live = m.live || ts->live[q0 + j] != 0;
  }
  // epilogue phase: query q of this lane
  const int q = threadIdx.x % T;
  const int mode = ts->mode[q];
  const QState qs = load_qstate(ts, q);
  int lq = 0;
  while ((1 << lq) < Q) lq++;
  Red red;
  for (int di = blockIdx.x; di < a.num_desc; di += gridDim.x) {
    const BlockDesc bd = a.desc[di];
    if ((bd.x & (kHubFlag | kRowFlag)) == 0) {
      // a run of rows, L = Q * 2^s lanes per row; ngrp rows per pass
      const int ls = bd.x;
      const int L = Q << ls;
      const int g = threadIdx.x / L;
      const int slice = (threadIdx.x % L) / Q;
      const int ngrp = kBlock / L;
      // each pass's row bounds are loaded one pass ahead
      int rb = 0, re = 0;
      if (bd.y + g < bd.z) {
        rb = a.row_ptr[bd.y + g];
        re = a.row_ptr[bd.y + g + 1];
      }
      for (int base = bd.y; base < bd.z; base += ngrp) {
        const int r = base + g;
        int nb = 0, ne = 0;
        if (r + ngrp < bd.z) {
          nb = a.row_ptr[r + ngrp];
          ne = a.row_ptr[r + ngrp + 1];
        }
        QV<LQ> acc = {};
        if (r < bd.z) acc = row_slice_sum<T>(a, yin, y64, m, rb + slice, re, 1 << ls, q0);
        rb = nb;
        re = ne;
        qv_lane_sum(acc, lq, lq + ls < 6 ? lq + ls : 6);  // Q << ls <= 64
        if (slice == 0) {
          double *o = rs + g * T + q0;
#pragma unroll
          for (int j = 0; j < LQ; j++) o[j] = acc.v[j];
        }
        __syncthreads();
        // at most LQ (row, query) pairs per thread, all with query q: the
        // loads of four of them first, then their epilogues
        const int pairs = (bd.z - base < ngrp ? bd.z - base : ngrp) * T;
#pragma unroll
        for (int j0 = 0; j0 < LQ; j0 += 4) {
          if ((int)threadIdx.x + j0 * kBlock >= pairs) break;
          EpiLoad ld[4];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int p = (int)threadIdx.x + (j0 + j) * kBlock;
            if (p < pairs && qs.live) ld[j] = epi_load<T>(a, qs, mode, base + p / T, q);
          }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int p = (int)threadIdx.x + (j0 + j) * kBlock;
            if (p < pairs) ppr_epilogue<T>(a, qs, mode, base + p / T, q, rs[p], ld[j], red);
          }
        }
        __syncthreads();
      }
    } else {
      // one long row (or one chunk of a hub row) over the whole block
      const int slice = threadIdx.x / Q;
      int b, e, r = -1;
      if (bd.x & kRowFlag) {
        r = bd.y;
        b = a.row_ptr[r];
        e = a.row_ptr[r + 1];
      } else {
        b = bd.z;
        e = bd.w;
      }
      const QV<LQ> acc =
          block_slices_sum<T>(row_slice_sum<T>(a, yin, y64, m, b + slice, e, kBlock / Q, q0));
      if (threadIdx.x < Q) {
        double *o = rs + q0;
#pragma unroll
        for (int j = 0; j < LQ; j++) o[j] = acc.v[j];
      }
      __syncthreads();
      if (threadIdx.x < T) {
        if (r >= 0) {
          const EpiLoad ld = qs.live ? epi_load<T>(a, qs, mode, r, q) : EpiLoad{};
          ppr_epilogue<T>(a, qs, mode, r, q, rs[q], ld, red);
        } else {
          const int h = bd.y, chunk = bd.x & (kHubFlag - 1);
          a.hub_partial[(size_t)(a.hub_first[h] + chunk) * T + q] = rs[q];
        }
      }
      __syncthreads();
    }
  }
  // Rows without in-edges (ids from n_in: the rows are sorted by in-degree)
  // get delta = teleport only, which is 0 for every query they do not seed:
  // only the ones that seed some query of the tile are swept, from their
  // seed entries (sorted by row; the first entry of each row takes it).
  {
    constexpr int RPB = kBlock / T;  // rows per workgroup pass
    const int j0 = a.zseed_first;
    for (int j = j0 + (int)blockIdx.x * RPB + (int)threadIdx.x / T; j < a.num_seed_entries;
         j += (int)gridDim.x * RPB) {
      const int r = a.seed_row[j];
      if (j > j0 && a.seed_row[j - 1] == r) continue;  // the row's first entry sweeps it
      const EpiLoad ld = qs.live ? epi_load<T>(a, qs, mode, r, q) : EpiLoad{};
      ppr_epilogue<T>(a, qs, mode, r, q, 0.0, ld, red);
    }
  }
  tile_block_sum<T>(red, a.partials + (size_t)blockIdx.x * T * 5);
}

// First stage of the per-query sums: workgroup b adds the partials of sweep
// workgroups [b * per, (b + 1) * per) in order.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_reduce_kernel(PprArgs a, int sweep_blocks) {
  if (a.ts->all_done) return;
  const int per = (sweep_blocks + kTileReduce - 1) / kTileReduce;
  const int b0 = blockIdx.x * per;
  const int b1 = b0 + per < sweep_blocks ? b0 + per : sweep_blocks;
  double *stage = a.partials + (size_t)kMaxSweepBlocks * T * 5;
  for (int v = threadIdx.x; v < T * 5; v += kBlock) {
    double s = 0.0;
    for (int b = b0; b < b1; b++) s += a.partials[(size_t)b * T * 5 + v];
    stage[(size_t)blockIdx.x * T * 5 + v] = s;
  }
}

// Per-query bookkeeping (pagerank.rs:163-179 for every query) and each
// query's next mode (the single-vector schedule of pagerank.hip).  One
// workgroup.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_finish_kernel(PprArgs a, int sweep_blocks) {
  TileState *ts = a.ts;
  if (ts->all_done) return;
  const int ycur = ts->yd_cur;
  a.yin32 = ycur ? a.yd1 : a.yd0;
  a.yout32 = ycur ? a.yd0 : a.yd1;
  const int q = threadIdx.x % T;
  const int mode = ts->mode[q];
  const QState qs = load_qstate(ts, q);
  Red red;
  // hub rows: (hub, query) pairs; thread t always gets query t % T
  for (int p = threadIdx.x; p < a.num_hubs * T; p += kBlock) {
    const int h = p / T;
    const int first = a.hub_first[h], nch = a.hub_nchunks[h];
    double s = 0.0;
    for (int c = 0; c < nch; c++) s += a.hub_partial[(size_t)(first + c) * T + q];
    const EpiLoad ld = qs.live ? epi_load<T>(a, qs, mode, a.hub_row[h], q) : EpiLoad{};
    ppr_epilogue<T>(a, qs, mode, a.hub_row[h], q, s, ld, red);
  }
  const double *stage = a.partials + (size_t)kMaxSweepBlocks * T * 5;
  for (int p = threadIdx.x; p < kTileReduce * T; p += kBlock) {
    const int b = p / T;
    {
      const double *pp = stage + ((size_t)b * T + q) * 5;
      red.l1 += pp[0];
      red.dang += pp[1];
      red.dang_x += pp[2];
      red.mass += pp[3];
      red.sum_x += pp[4];
    }
  }
  __shared__ double tot[T * 5];
  tile_block_sum<T>(red, tot);
  if (threadIdx.x < T && ts->live[threadIdx.x]) {
    const int qq = threadIdx.x;
    const double *s = tot + qq * 5;
    const double diff = s[0];
    ts->mass_in[qq] = s[3];
    ts->dsum_in[qq] = s[1];
    int next;
    if (mode == kModeAnchor) {
      ts->dsum_x[qq] = s[2];
      ts->diff_anchor[qq] = diff;
      ts->x_scale[qq] = 1.0;
      ts->since_anchor[qq] = 0;
      if (ts->f64_phase[qq] && diff > ts->f32_below * ts->tol[qq]) {
        next = kModeAnchor;  // this anchor wrote the next one's input
      } else {
        ts->f64_phase[qq] = 0;
        next = kModeDelta;
      }
    } else if (mode == kModeDeltaFold) {
      ts->dsum_x[qq] = s[2];
      ts->x_scale[qq] = 1.0 / s[4];
      next = kModeAnchor;
    } else {
      const uint64_t since = ts->since_anchor[qq] + 1;
      ts->since_anchor[qq] = since;
      // no anchors once the residual is within anchor_min of the tolerance:
      // the f32 rounding of the sweeps left is far below the tolerance then
      const bool anchor = diff > ts->anchor_min * ts->tol[qq] &&
                          (diff <= ts->anchor_drop * ts->diff_anchor[qq] ||
                           since + 1 >= ts->anchor_period);
      next = anchor ? kModeDeltaFold : kModeDelta;
    }
    ts->mode[qq] = next;
    const uint64_t it = ts->iter[qq] + 1;
    ts->iter[qq] = it;
    ts->diff[qq] = diff;
    if (diff < ts->tol[qq]) {
      ts->converged[qq] = 1;
      ts->live[qq] = 0;
    } else if (it >= ts->max_iter[qq]) {
      ts->live[qq] = 0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int live = 0, free_groups = 0;  // groups of kLaneQ slots (one lane's queries) with no live query
    for (int g = 0; g < T; g += kLaneQ) {
      int gl = 0;
      for (int i = g; i < g + kLaneQ; i++) gl += ts->live[i];
      live += gl;
      free_groups += gl == 0;
    }
    ts->yd_cur = ycur ^ 1;
    ts->y64_par ^= 1;
    ts->sweep += 1;
    ts->all_done = live == 0;
    if (a.mirror)
      __hip_atomic_store(a.mirror,
                         (ts->sweep << 8) | ((uint64_t)free_groups << 1) | (uint64_t)(live == 0),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Admission, step 1: clear the admitted queries' columns of the f32 deltas
// the next sweep gathers, and, for an anchor admission, of x and of the y64
// buffer the next sweep's anchors gather.  One thread per row and group of
// four queries; the tile's other queries are not touched.
//
// An anchor admission's first sweep reads x and y64.  A delta admission's
// first sweep reads only the gathered deltas: it takes x and the carry as
// p-hat and its rounding (ppr_epilogue), and a fold sweep writes the whole
// y64 column before any anchor reads it.
//
// Rows without in-edges (r >= n_in) are swept only while they are seeds of
// some query of the tile (ppr_sweep_kernel), and such a sweep leaves a
// column that is not seeded at the row at 0 (nothing is gathered, and the
// teleport term is 0 off the query's seeds).  So a column is nonzero there
// only at the seed rows of the query that held it: the host passes the
// zero-in seed rows of the departing queries (zrows), and all six vectors of
// the admitted columns are cleared at those rows alone (the tile starts at 0).
template <int T>
__device__ __forceinline__ void clear_row(const PprArgs &a, int r, unsigned m4, int qg,
                                          bool zero_in, bool anchored, double *y64, float *yin,
                                          float *yout) {
  const size_t at = (size_t)r * T + qg * kLaneQ;
  if (zero_in) {
    if (m4 == 0xFu) {
      const double2 z = {0.0, 0.0};
      const float4 zf = {0.0f, 0.0f, 0.0f, 0.0f};
      for (double *v : {a.x, a.y64a, a.y64b}) {
        *reinterpret_cast<double2 *>(v + at) = z;
        *reinterpret_cast<double2 *>(v + at + 2) = z;
      }
      for (float *v : {a.carry, yin, yout}) *reinterpret_cast<float4 *>(v + at) = zf;
    } else {
#pragma unroll
      for (int j = 0; j < kLaneQ; j++) {
        if ((m4 >> j) & 1u) {
          a.x[at + j] = 0.0;
          a.carry[at + j] = 0.0f;
          yin[at + j] = 0.0f;
          yout[at + j] = 0.0f;
          a.y64a[at + j] = 0.0;
          a.y64b[at + j] = 0.0;
        }
      }
    }
    return;
  }
  if (m4 == 0xFu) {
    const double2 z = {0.0, 0.0};
    if (anchored) {
      *reinterpret_cast<double2 *>(a.x + at) = z;
      *reinterpret_cast<double2 *>(a.x + at + 2) = z;
      *reinterpret_cast<double2 *>(y64 + at) = z;
      *reinterpret_cast<double2 *>(y64 + at + 2) = z;
    }
    *reinterpret_cast<float4 *>(yin + at) = float4{0.0f, 0.0f, 0.0f, 0.0f};
  } else {
#pragma unroll
    for (int j = 0; j < kLaneQ; j++) {
      if ((m4 >> j) & 1u) {
        if (anchored) {
          a.x[at + j] = 0.0;
          y64[at + j] = 0.0;
        }
        yin[at + j] = 0.0f;
      }
    }
  }
}

template <int T>
__global__ __launch_bounds__(kBlock) void ppr_clear_kernel(PprArgs a, uint64_t clear_mask,
                                                           const int32_t *zrows, int nz) {
  constexpr int Q = T / kLaneQ;
  const int qg = threadIdx.x % Q;
  const unsigned m4 = (unsigned)(clear_mask >> (qg * kLaneQ)) & 0xFu;
  if (!m4) return;
  const bool anchored = a.ts->f32_below > 0.0;
  double *y64 = anchored ? (a.ts->y64_par ? a.y64b : a.y64a) : nullptr;
  float *yin = a.ts->yd_cur ? a.yd1 : a.yd0;
  float *yout = a.ts->yd_cur ? a.yd0 : a.yd1;
  const int rows_per_block = kBlock / Q;
  const int first = blockIdx.x * rows_per_block + threadIdx.x / Q;
  const int stride = gridDim.x * rows_per_block;
  for (int r = first; r < a.n_in; r += stride)
    clear_row<T>(a, r, m4, qg, false, anchored, y64, yin, yout);
  for (int i = first; i < nz; i += stride)
    clear_row<T>(a, zrows[i], m4, qg, true, anchored, y64, yin, yout);
}

// Admission, step 2: the admitted queries start from their personalization
// (x = p-hat on the seeds, ppr.rs:110).  Their first sweep computes
// T(p-hat) - p-hat.  With a leading f64 phase (f32_below > 0) it is an f64
// anchor sweep over y64 = d p-hat / deg.  Without, it is a delta sweep whose
// input delta is p-hat itself: d A p-hat + d (dangling(p-hat) - sum(p-hat))
// p-hat is T(p-hat) - p-hat, as sum(p-hat) = 1 (ppr.rs:125-137).
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_admit_kernel(PprArgs a, uint64_t admit_mask) {
  TileState *ts = a.ts;
  const bool anchored = ts->f32_below > 0.0;
  double *y64 = ts->y64_par ? a.y64b : a.y64a;
  float *yin = ts->yd_cur ? a.yd1 : a.yd0;
  for (int j = threadIdx.x; j < a.num_seed_entries; j += kBlock) {
    const int qq = a.seed_q[j];
    if (!((admit_mask >> qq) & 1ull)) continue;
    const int r = a.seed_row[j];
    const size_t at = (size_t)r * T + qq;
    const double p = a.seed_p[j];
    const double v = (ts->damping[qq] * p) * a.inv64[r];
    if (anchored) {
      a.x[at] = p;  // one entry per (row, query): the host merges duplicate seeds
      y64[at] = v;
    } else {
      yin[at] = delta_put(v);  // x and the carry: ppr_epilogue's first sweep
    }
  }
  __syncthreads();
  if (threadIdx.x < T) {
    const int qq = threadIdx.x;
    if ((admit_mask >> qq) & 1ull) {
      ts->dsum_x[qq] = ts->init_dsum[qq];
      ts->x_scale[qq] = 1.0;
      ts->live[qq] = 1;
      ts->iter[qq] = 0;
      ts->converged[qq] = 0;
      ts->diff[qq] = 0.0;
      ts->mode[qq] = anchored ? kModeAnchor : kModeDelta;
      ts->since_anchor[qq] = 0;
      ts->f64_phase[qq] = anchored;
      ts->dsum_in[qq] = ts->init_dsum[qq];
      ts->mass_in[qq] = ts->init_mass[qq];
      ts->diff_anchor[qq] = ts->init_mass[qq];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int live = 0;
    for (int i = 0; i < T; i++) live += ts->live[i];
    ts->all_done = live == 0;
  }
}

// out[o] = x[rank[o]][q]: query q's scores in the caller's ids.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_gather_kernel(PprArgs a, const uint32_t *rank, int n,
                                                            int q, double *out) {
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock)
    out[o] = a.x[(size_t)rank[o] * T + q];
}

// out[j * n + o] = x[rank[o]][q] for the queries q of mask (j: q's rank
// among them), in the caller's ids: one pass over the rows for all of them,
// each row's values read from the lines of one 8 T-byte row.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_gather_mask_kernel(PprArgs a, const uint32_t *rank,
                                                                 int n, uint64_t mask,
                                                                 double *out) {
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) {
    const double *row = a.x + (size_t)rank[o] * T;
    uint64_t m = mask;
    for (int j = 0; m; j++, m &= m - 1) out[(size_t)j * n + o] = row[__ffsll((long long)m) - 1];
  }
}

__global__ void seed_mark_kernel(int32_t *seed_idx, const int32_t *rows, int count, int set) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
    const int r = rows[j];
    if (!set) {
      seed_idx[r] = -1;
    } else if (j == 0 || rows[j - 1] != r) {
      seed_idx[r] = j;  // first entry of that row
    }
  }
}

// ---------------------------------------------------------------------------
template <int T>
static hipError_t launch_sweep_t(hipStream_t s, const PprArgs &a, int grid) {
  ppr_sweep_kernel<T><<<grid, kBlock, 0, s>>>(a);
  ppr_reduce_kernel<T><<<kTileReduce, kBlock, 0, s>>>(a, grid);
  ppr_finish_kernel<T><<<1, kBlock, 0, s>>>(a, grid);
  return hipGetLastError();
}

hipError_t ppr_launch_sweep(hipStream_t s, int T, const PprArgs &a, int grid) {
  switch (T) {
    case 8: return launch_sweep_t<8>(s, a, grid);
    case 16: return launch_sweep_t<16>(s, a, grid);
    case 32: return launch_sweep_t<32>(s, a, grid);
    default: return launch_sweep_t<64>(s, a, grid);
  }
}

template <int T>
static hipError_t launch_admit_t(hipStream_t s, const PprArgs &a, uint64_t clear_mask,
                                 uint64_t admit_mask, const int32_t *zrows, int nz) {
  const int rows_per_block = kBlock / (T / kLaneQ);
  int grid = (std::max(a.n_in, nz) + rows_per_block - 1) / rows_per_block;
  grid = std::max(1, std::min(grid, 4096));
  ppr_clear_kernel<T><<<grid, kBlock, 0, s>>>(a, clear_mask, zrows, nz);
  ppr_admit_kernel<T><<<1, kBlock, 0, s>>>(a, admit_mask);
  return hipGetLastError();
}

hipError_t ppr_launch_admit(hipStream_t s, int T, const PprArgs &a, uint64_t clear_mask,
                            uint64_t admit_mask, const int32_t *zrows, int nz) {
  switch (T) {
    case 8: return launch_admit_t<8>(s, a, clear_mask, admit_mask, zrows, nz);
    case 16: return launch_admit_t<16>(s, a, clear_mask, admit_mask, zrows, nz);
    case 32: return launch_admit_t<32>(s, a, clear_mask, admit_mask, zrows, nz);
    default: return launch_admit_t<64>(s, a, clear_mask, admit_mask, zrows, nz);
  }
}

hipError_t ppr_launch_gather(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank, int n,
                             int q, double *out) {
  const int grid = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
  switch (T) {
    case 8: ppr_gather_kernel<8><<<grid, kBlock, 0, s>>>(a, rank, n, q, out); break;
    case 16: ppr_gather_kernel<16><<<grid, kBlock, 0, s>>>(a, rank, n, q, out); break;
    case 32: ppr_gather_kernel<32><<<grid, kBlock, 0, s>>>(a, rank, n, q, out); break;
    default: ppr_gather_kernel<64><<<grid, kBlock, 0, s>>>(a, rank, n, q, out); break;
  }
  return hipGetLastError();
}

hipError_t ppr_launch_gather_mask(hipStream_t s, int T, const PprArgs &a, const uint32_t *rank,
                                  int n, uint64_t mask, double *out) {
  const int grid = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
  switch (T) {
    case 8: ppr_gather_mask_kernel<8><<<grid, kBlock, 0, s>>>(a, rank, n, mask, out); break;
    case 16: ppr_gather_mask_kernel<16><<<grid, kBlock, 0, s>>>(a, rank, n, mask, out); break;
    case 32: ppr_gather_mask_kernel<32><<<grid, kBlock, 0, s>>>(a, rank, n, mask, out); break;
    default: ppr_gather_mask_kernel<64><<<grid, kBlock, 0, s>>>(a, rank, n, mask, out); break;
  }
  return hipGetLastError();
}

hipError_t ppr_launch_seed_mark(hipStream_t s, int32_t *seed_idx, const int32_t *rows, int count,
                                int set) {
  if (count <= 0) return hipSuccess;
  const int grid = std::max(1, std::min((count + 255) / 256, 1024));
  seed_mark_kernel<<<grid, 256, 0, s>>>(seed_idx, rows, count, set);
  return hipGetLastError();
}

}  // namespace gops
