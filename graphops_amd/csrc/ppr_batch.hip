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

// Per (row, query) epilogue.  mode: kModeDelta / kModeDeltaFold / kModeAnchor.
template <int T>
__device__ __forceinline__ void ppr_epilogue(const PprArgs &a, const TileState *ts, int mode,
                                             int r, int q, double acc, Red &red) {
  if (!ts->live[q]) return;
  const double d = ts->damping[q];
  double seed_p = 0.0;
  const int sidx = a.seed_idx[r];
  if (sidx >= 0) {
    for (int j = sidx; j < a.num_seed_entries && a.seed_row[j] == r; j++)
      if (a.seed_q[j] == q) seed_p += a.seed_p[j];
  }
  const size_t at = (size_t)r * T + q;
  const int c64 = ts->y64_cur[q];
  // every path sets all five contributions; one update of red at the end
  double delta, dang_x = 0.0, sum_x = 0.0;
  bool dangling;
  if (mode == kModeAnchor) {
    const double sc = ts->x_scale[q];
    const double dsum = ts->dsum_x[q] * sc;
    const double t = sc * acc + ((1.0 - d) * seed_p + (d * dsum) * seed_p);  // ppr.rs:126-127
    delta = t - sc * a.x[at];
    a.x[at] = t;
    const double inv = a.inv64[r];
    a.yout32[at] = (float)((d * delta) * inv);
    if (ts->f64_phase[q]) (c64 ? a.y64a : a.y64b)[at] = (d * t) * inv;  // next anchor's input
    dangling = inv == 0.0;
    dang_x = dangling ? t : 0.0;
  } else {
    // delta' = A delta: the dangling and (removed) total sums go to the seeds
    delta = acc + d * (ts->dsum_in[q] - ts->mass_in[q]) * seed_p;
    const double x = __builtin_nontemporal_load(a.x + at) + delta;
    __builtin_nontemporal_store(x, a.x + at);
    const float inv = a.inv32[r];
    a.yout32[at] = (float)(d * delta * (double)inv);
    dangling = inv == 0.0f;
    if (mode == kModeDeltaFold) {
      const double inv64 = a.inv64[r];
      (c64 ? a.y64b : a.y64a)[at] = (d * x) * inv64;
      dang_x = inv64 == 0.0 ? x : 0.0;
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

// Sum of v over the block's lanes with the same query (lane % T), for
// block-wide rows; valid in threads < T.
template <int T>
__device__ __forceinline__ double tile_query_sum(double v) {
  __shared__ double sh[kBlock / 64][T];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int lb = 0;
  while ((1 << lb) < T) lb++;
  const double s = lane_bits_sum(v, lb, 6);
  if (lane < T) sh[wave][lane] = s;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < T) r = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
  __syncthreads();
  return r;
}

// Gather-sum of one query's slice of a row: f64 iterate for a query in an
// anchor sweep, f32 delta otherwise.
template <int T>
__device__ __forceinline__ double tile_row_sum(const PprArgs &a, const double *y64, bool anchor,
                                               int b, int e, int step, int q) {
  // four edges per iteration: their column loads, then their gathers, are in
  // flight together
  double acc = 0.0;
  int k = b;
  if (anchor) {
    for (; k + 3 * step < e; k += 4 * step) {
      const int c0 = a.col[k], c1 = a.col[k + step], c2 = a.col[k + 2 * step],
                c3 = a.col[k + 3 * step];
      const double v0 = y64[(size_t)c0 * T + q], v1 = y64[(size_t)c1 * T + q],
                   v2 = y64[(size_t)c2 * T + q], v3 = y64[(size_t)c3 * T + q];
      acc += (v0 + v1) + (v2 + v3);
    }
    for (; k < e; k += step) acc += y64[(size_t)a.col[k] * T + q];
  } else {
    for (; k + 3 * step < e; k += 4 * step) {
      const int c0 = a.col[k], c1 = a.col[k + step], c2 = a.col[k + 2 * step],
                c3 = a.col[k + 3 * step];
      const float v0 = a.yin32[(size_t)c0 * T + q], v1 = a.yin32[(size_t)c1 * T + q],
                  v2 = a.yin32[(size_t)c2 * T + q], v3 = a.yin32[(size_t)c3 * T + q];
      acc += ((double)v0 + (double)v1) + ((double)v2 + (double)v3);
    }
    for (; k < e; k += step) acc += (double)a.yin32[(size_t)a.col[k] * T + q];
  }
  return acc;
}

template <int T>
__global__ __launch_bounds__(kBlock) void ppr_sweep_kernel(PprArgs a) {
  const TileState *ts = a.ts;
  if (ts->all_done) return;
  const int ycur = ts->yd_cur;
  a.yin32 = ycur ? a.yd1 : a.yd0;
  a.yout32 = ycur ? a.yd0 : a.yd1;
  const int q = threadIdx.x % T;
  const int mode = ts->mode[q];  // every query has its own sweep mode
  const bool anchor = mode == kModeAnchor;
  const double *y64 = ts->y64_cur[q] ? a.y64b : a.y64a;
  Red red;
  for (int di = blockIdx.x; di < a.num_desc; di += gridDim.x) {
    const BlockDesc bd = a.desc[di];
    if ((bd.x & (kHubFlag | kRowFlag)) == 0) {
      // a run of rows, L = T * 2^s lanes per row (s slices of the edges)
      const int ls = bd.x;
      const int L = T << ls;
      const int g = threadIdx.x / L;
      const int slice = (threadIdx.x % L) / T;
      const int ngrp = kBlock / L;
      int lb = 0;
      while ((1 << lb) < T) lb++;
      for (int base = bd.y; base < bd.z; base += ngrp) {
        const int r = base + g;
        double acc = 0.0;
        if (r < bd.z)
          acc = tile_row_sum<T>(a, y64, anchor, a.row_ptr[r] + slice, a.row_ptr[r + 1], 1 << ls, q);
        acc = lane_bits_sum(acc, lb, lb + ls < 6 ? lb + ls : 6);  // T << ls <= 64
        if (r < bd.z && slice == 0) ppr_epilogue<T>(a, ts, mode, r, q, acc, red);
      }
    } else {
      // one long row (or one chunk of a hub row) over the whole block
      const int slice = threadIdx.x / T;
      int b, e, r = -1;
      if (bd.x & kRowFlag) {
        r = bd.y;
        b = a.row_ptr[r];
        e = a.row_ptr[r + 1];
      } else {
        b = bd.z;
        e = bd.w;
      }
      double acc = tile_row_sum<T>(a, y64, anchor, b + slice, e, kBlock / T, q);
      acc = tile_query_sum<T>(acc);
      if (threadIdx.x < T) {
        if (r >= 0) {
          ppr_epilogue<T>(a, ts, mode, r, q, acc, red);
        } else {
          const int h = bd.y, chunk = bd.x & (kHubFlag - 1);
          a.hub_partial[(size_t)(a.hub_first[h] + chunk) * T + q] = acc;
        }
      }
    }
  }
  tile_block_sum<T>(red, a.partials + (size_t)blockIdx.x * T * 5);
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
  Red red;
  // hub rows: (hub, query) pairs; thread t always gets query t % T
  for (int p = threadIdx.x; p < a.num_hubs * T; p += kBlock) {
    const int h = p / T;
    const int first = a.hub_first[h], nch = a.hub_nchunks[h];
    double s = 0.0;
    for (int c = 0; c < nch; c++) s += a.hub_partial[(size_t)(first + c) * T + q];
    ppr_epilogue<T>(a, ts, mode, a.hub_row[h], q, s, red);
  }
  for (int p = threadIdx.x; p < sweep_blocks * T; p += kBlock) {
    const int b = p / T;
    {
      const double *pp = a.partials + ((size_t)b * T + q) * 5;
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
        ts->y64_cur[qq] ^= 1;  // this anchor wrote the next one's input
        next = kModeAnchor;
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
      const bool anchor = diff <= ts->anchor_drop * ts->diff_anchor[qq] ||
                          (since + 1 >= ts->anchor_period && diff > ts->anchor_min * ts->tol[qq]);
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
    int live = 0;
    for (int i = 0; i < T; i++) live += ts->live[i];
    ts->yd_cur = ycur ^ 1;
    ts->sweep += 1;
    ts->all_done = live == 0;
    if (a.mirror)
      __hip_atomic_store(a.mirror, (ts->sweep << 1) | (uint64_t)(live == 0), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Admission, step 1: clear the admitted queries' columns (x = 0, y64 = 0).
// The tile's other queries are not touched.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_clear_kernel(PprArgs a, int n, uint64_t clear_mask) {
  const int q = threadIdx.x % T;
  if (!((clear_mask >> q) & 1ull)) return;
  const int rows_per_block = kBlock / T;
  for (int r = blockIdx.x * rows_per_block + threadIdx.x / T; r < n;
       r += gridDim.x * rows_per_block) {
    const size_t at = (size_t)r * T + q;
    a.x[at] = 0.0;
    a.y64a[at] = 0.0;  // the admitted query's first anchor gathers y64a
  }
}

// Admission, step 2: the admitted queries start from their personalization
// (x = p-hat on the seeds, ppr.rs:110) with an f64 anchor sweep.
template <int T>
__global__ __launch_bounds__(kBlock) void ppr_admit_kernel(PprArgs a, uint64_t admit_mask) {
  TileState *ts = a.ts;
  for (int j = threadIdx.x; j < a.num_seed_entries; j += kBlock) {
    const int qq = a.seed_q[j];
    if (!((admit_mask >> qq) & 1ull)) continue;
    const int r = a.seed_row[j];
    const size_t at = (size_t)r * T + qq;
    const double p = a.seed_p[j];
    a.x[at] = p;  // one entry per (row, query): the host merges duplicate seeds
    a.y64a[at] = (ts->damping[qq] * p) * a.inv64[r];
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
      ts->mode[qq] = kModeAnchor;
      ts->since_anchor[qq] = 0;
      ts->y64_cur[qq] = 0;
      ts->f64_phase[qq] = ts->f32_below > 0.0;
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
static hipError_t launch_admit_t(hipStream_t s, const PprArgs &a, int n, uint64_t clear_mask,
                                 uint64_t admit_mask) {
  const int rows_per_block = kBlock / T;
  int grid = (n + rows_per_block - 1) / rows_per_block;
  grid = std::max(1, std::min(grid, 4096));
  ppr_clear_kernel<T><<<grid, kBlock, 0, s>>>(a, n, clear_mask);
  ppr_admit_kernel<T><<<1, kBlock, 0, s>>>(a, admit_mask);
  return hipGetLastError();
}

hipError_t ppr_launch_admit(hipStream_t s, int T, const PprArgs &a, int n, uint64_t clear_mask,
                            uint64_t admit_mask) {
  switch (T) {
    case 8: return launch_admit_t<8>(s, a, n, clear_mask, admit_mask);
    case 16: return launch_admit_t<16>(s, a, n, clear_mask, admit_mask);
    case 32: return launch_admit_t<32>(s, a, n, clear_mask, admit_mask);
    default: return launch_admit_t<64>(s, a, n, clear_mask, admit_mask);
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

hipError_t ppr_launch_seed_mark(hipStream_t s, int32_t *seed_idx, const int32_t *rows, int count,
                                int set) {
  if (count <= 0) return hipSuccess;
  const int grid = std::max(1, std::min((count + 255) / 256, 1024));
  seed_mark_kernel<<<grid, 256, 0, s>>>(seed_idx, rows, count, set);
  return hipGetLastError();
}

}  // namespace gops
