// pagerank.hip -- gfx950 sweep kernels of the PageRank family.
//
// One sweep = one iteration of pagerank.rs:142-180 / 239-281 or ppr.rs:117-149,
// written as a pull over the CSR of in-edges (rows renumbered by in-degree).
// Each sweep is one pass over HBM that fuses: the gather-multiply-reduce of
// the row, the teleport and dangling-redistribution terms, the L1 residual
// (f64) and the dangling sum the next sweep needs (f64).  The grid-wide f64
// sums are made by the last workgroup to finish, in a fixed order, so the
// results do not depend on scheduling.
//
// Two sweep kinds (mixed precision, "delta" scheme):
//   f64 anchor sweep:  t = fill + A x  in f64 (x = anchor iterate xbar),
//                      delta = t - xbar, xbar <- t.
//   f32 delta sweep:   delta' = A delta in f32 (A = d * P^T plus dangling
//                      redistribution, a linear map), e += delta'.
// The iterate is x = xbar + e and delta is exactly x_{k+1} - x_k, so the L1
// residual sum |delta| tracks the f64 iteration's residual without the f32
// rounding floor of storing x itself in f32.  A periodic f64 anchor sweep
// (after folding e into xbar) recomputes delta from the iterate and bounds
// the drift of the f32 chain.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace gops {

struct NoW {};
struct Bf16W {};

struct SweepArgs {
  const BlockDesc *desc;
  const int32_t *row_ptr;
  const int32_t *col;
  const void *w;
  const float *inv32;
  const double *inv64;
  const float *yin32;
  float *yout32;
  const double *yin64;
  double *yout64;
  float *e;
  double *xbar;
  const float *q32;
  const double *q64;
  double *partials;
  double *hub_partial;
  uint32_t *hub_count;
  const int32_t *hub_row;
  const int32_t *hub_nchunks;
  const int32_t *hub_first;
  RunState *st;
};

template <typename WT>
__device__ __forceinline__ float wload32(const void *w, int k) {
  if constexpr (std::is_same<WT, float>::value) {
    return __builtin_nontemporal_load(static_cast<const float *>(w) + k);
  } else if constexpr (std::is_same<WT, Bf16W>::value) {
    const uint32_t h = __builtin_nontemporal_load(static_cast<const uint16_t *>(w) + k);
    return __uint_as_float(h << 16);
  } else if constexpr (std::is_same<WT, double>::value) {
    return (float)__builtin_nontemporal_load(static_cast<const double *>(w) + k);
  } else {
    return 1.0f;
  }
}

template <typename WT>
__device__ __forceinline__ double wload64(const void *w, int k) {
  if constexpr (std::is_same<WT, float>::value) {
    return (double)__builtin_nontemporal_load(static_cast<const float *>(w) + k);
  } else if constexpr (std::is_same<WT, Bf16W>::value) {
    const uint32_t h = __builtin_nontemporal_load(static_cast<const uint16_t *>(w) + k);
    return (double)__uint_as_float(h << 16);
  } else if constexpr (std::is_same<WT, double>::value) {
    return __builtin_nontemporal_load(static_cast<const double *>(w) + k);
  } else {
    return 1.0;
  }
}

// ---- cross-lane sums within a group of 2^lg lanes (all lanes get the sum) --
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = dpp_i<CTRL>(__double2loint(v));
  const int hi = dpp_i<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
// ds_swizzle bit mode: and_mask 0x1f, or_mask 0, xor_mask 0x10 -> lane ^ 16 within 32
constexpr int kSwzXor16 = 0x1F | (0x10 << 10);

__device__ __forceinline__ float group_sum(float v, int lg) {
  if (lg >= 1) v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  if (lg >= 2) v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  if (lg >= 3) v += dpp_f<0x141>(v);  // row_half_mirror
  if (lg >= 4) v += dpp_f<0x140>(v);  // row_mirror
  if (lg >= 5) v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), kSwzXor16));
  if (lg >= 6) v += __shfl_xor(v, 32);
  return v;
}
__device__ __forceinline__ double group_sum(double v, int lg) {
  if (lg >= 1) v += dpp_d<0xB1>(v);
  if (lg >= 2) v += dpp_d<0x4E>(v);
  if (lg >= 3) v += dpp_d<0x141>(v);
  if (lg >= 4) v += dpp_d<0x140>(v);
  if (lg >= 5) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), kSwzXor16);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), kSwzXor16);
    v += __hiloint2double(hi, lo);
  }
  if (lg >= 6) v += __shfl_xor(v, 32);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  return group_sum(v, 6);
}

// Sum of `v` over the workgroup, valid in thread 0.  Fixed order.
template <typename T, int N>
__device__ __forceinline__ void block_sum(T (&v)[N]) {
  __shared__ T sh[N][kBlock / 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < N; i++) {
    T s = wave_sum(v[i]);
    if (lane == 0) sh[i][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = (sh[i][0] + sh[i][1]) + (sh[i][2] + sh[i][3]);
  }
  __syncthreads();
}

// Writes this workgroup's partial sums; the last workgroup to arrive reduces
// all of them (fixed order) into out[0..N) and returns true in thread 0.
template <int N>
__device__ bool grid_reduce(double (&v)[N], double *partials, uint32_t *ticket,
                            double (&out)[N]) {
  block_sum<double, N>(v);
  __shared__ int last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) partials[(size_t)blockIdx.x * N + i] = v[i];
    __threadfence();
    const uint32_t t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double s[N];
#pragma unroll
  for (int i = 0; i < N; i++) s[i] = 0.0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) {
#pragma unroll
    for (int i = 0; i < N; i++) s[i] += __builtin_nontemporal_load(partials + (size_t)b * N + i);
  }
  block_sum<double, N>(s);
  if (threadIdx.x == 0) {
    *ticket = 0;
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = s[i];
    return true;
  }
  return false;
}

// Bookkeeping of pagerank.rs:163-179: iterations counts the sweep, diff_l1 is
// its residual, stop on diff < tolerance (strict) or at the iteration cap.
__device__ __forceinline__ void finish_sweep(RunState *st, double diff) {
  const uint64_t it = st->iter + 1;
  st->iter = it;
  st->diff = diff;
  int conv = 0, done = 0;
  if (diff < st->tolerance) {
    conv = 1;
    done = 1;
  } else if (it >= st->max_iter) {
    done = 1;
  }
  st->converged = conv;
  st->done = done;
  if (st->mirror) {
    __hip_atomic_store(st->mirror, (it << 2) | (uint64_t)(conv << 1) | (uint64_t)done,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
  }
}

// ---------------------------------------------------------------------------
// f32 delta sweep:  delta'_v = sum_{u->v} w_uv * yin_u + d * dsum(delta) * q_v
//                   e_v += delta'_v;  yout_v = d * delta'_v * inv_v
template <typename WT>
__global__ __launch_bounds__(kBlock) void sweep_f32_kernel(SweepArgs a) {
  RunState *st = a.st;
  if (st->done) return;
  const float df = (float)st->damping;
  const double ddsum = st->damping * st->dsum_in;
  const float uni = (float)(ddsum * st->inv_n);
  const float ddsum_f = (float)ddsum;
  const BlockDesc bd = a.desc[blockIdx.x];
  double red[2] = {0.0, 0.0};  // |delta| sum, dangling delta sum

  auto epilogue = [&](int r, float acc) {
    const float delta = acc + (a.q32 ? ddsum_f * a.q32[r] : uni);
    const float inv = __builtin_nontemporal_load(a.inv32 + r);
    const float e_old = __builtin_nontemporal_load(a.e + r);
    __builtin_nontemporal_store(e_old + delta, a.e + r);
    a.yout32[r] = (df * delta) * inv;
    red[0] += fabs((double)delta);
    if (inv == 0.0f) red[1] += (double)delta;
  };

  if ((bd.x & kHubFlag) == 0) {
    const int lg = bd.x;
    const int G = 1 << lg;
    const int lane = threadIdx.x & (G - 1);
    const int grp = threadIdx.x >> lg;
    const int ngrp = kBlock >> lg;
    for (int base = bd.y; base < bd.z; base += ngrp) {
      const int r = base + grp;
      float acc = 0.0f;
      if (r < bd.z) {
        const int b = __builtin_nontemporal_load(a.row_ptr + r);
        const int e = __builtin_nontemporal_load(a.row_ptr + r + 1);
        for (int k = b + lane; k < e; k += G) {
          const int c = __builtin_nontemporal_load(a.col + k);
          acc += wload32<WT>(a.w, k) * a.yin32[c];
        }
      }
      acc = group_sum(acc, lg);
      if (r < bd.z && lane == 0) epilogue(r, acc);
    }
  } else {
    const int h = bd.y;
    const int chunk = bd.x & (kHubFlag - 1);
    float acc = 0.0f;
    for (int k = bd.z + (int)threadIdx.x; k < bd.w; k += kBlock) {
      const int c = __builtin_nontemporal_load(a.col + k);
      acc += wload32<WT>(a.w, k) * a.yin32[c];
    }
    float tot[1] = {acc};
    block_sum<float, 1>(tot);
    if (threadIdx.x == 0) {
      const int first = a.hub_first[h], nch = a.hub_nchunks[h];
      a.hub_partial[first + chunk] = (double)tot[0];
      __threadfence();
      const uint32_t prev = atomicAdd(a.hub_count + h, 1u);
      if (prev == (uint32_t)(nch - 1)) {
        __threadfence();
        double s = 0.0;
        for (int c = 0; c < nch; c++) s += __builtin_nontemporal_load(a.hub_partial + first + c);
        a.hub_count[h] = 0;
        epilogue(a.hub_row[h], (float)s);
      }
    }
  }

  double out[2];
  if (grid_reduce<2>(red, a.partials, &st->ticket, out)) {
    st->dsum_in = out[1];
    finish_sweep(st, out[0]);
  }
}

// ---------------------------------------------------------------------------
// f64 anchor sweep:  t_v = fill_v + sum_{u->v} w_uv * yin64_u
//   fill_v = (1-d)/n + (d*dsum)/n (pagerank.rs:150-152) or
//            (1-d)*q_v + (d*dsum)*q_v (ppr.rs:126-127)
//   delta = t - xbar_v; xbar_v = t; optional f32 yout = d*delta*inv (for a
//   following delta sweep) and f64 y64out = d*t*inv (for a following anchor).
template <typename WT, bool WRITE_YD, bool WRITE_Y64>
__global__ __launch_bounds__(kBlock) void sweep_f64_kernel(SweepArgs a) {
  RunState *st = a.st;
  if (st->done) return;
  const double d = st->damping;
  const double nf = st->n_f64;
  const double fill_u = (1.0 - d) / nf + d * st->dsum_x / nf;
  const double ddsum = d * st->dsum_x;
  const BlockDesc bd = a.desc[blockIdx.x];
  double red[3] = {0.0, 0.0, 0.0};  // |delta|, dangling delta, dangling t

  auto epilogue = [&](int r, double acc) {
    double fill;
    if (a.q64) {
      const double q = a.q64[r];
      fill = (1.0 - d) * q + ddsum * q;
    } else {
      fill = fill_u;
    }
    const double t = acc + fill;
    const double xo = a.xbar[r];
    const double delta = t - xo;
    a.xbar[r] = t;
    const double inv = a.inv64[r];
    if (WRITE_YD) a.yout32[r] = (float)((d * delta) * inv);
    if (WRITE_Y64) a.yout64[r] = (d * t) * inv;
    red[0] += fabs(delta);
    if (inv == 0.0) {
      red[1] += delta;
      red[2] += t;
    }
  };

  if ((bd.x & kHubFlag) == 0) {
    const int lg = bd.x;
    const int G = 1 << lg;
    const int lane = threadIdx.x & (G - 1);
    const int grp = threadIdx.x >> lg;
    const int ngrp = kBlock >> lg;
    for (int base = bd.y; base < bd.z; base += ngrp) {
      const int r = base + grp;
      double acc = 0.0;
      if (r < bd.z) {
        const int b = a.row_ptr[r], e = a.row_ptr[r + 1];
        for (int k = b + lane; k < e; k += G) {
          const int c = __builtin_nontemporal_load(a.col + k);
          acc += wload64<WT>(a.w, k) * a.yin64[c];
        }
      }
      acc = group_sum(acc, lg);
      if (r < bd.z && lane == 0) epilogue(r, acc);
    }
  } else {
    const int h = bd.y;
    const int chunk = bd.x & (kHubFlag - 1);
    double acc = 0.0;
    for (int k = bd.z + (int)threadIdx.x; k < bd.w; k += kBlock) {
      const int c = __builtin_nontemporal_load(a.col + k);
      acc += wload64<WT>(a.w, k) * a.yin64[c];
    }
    double tot[1] = {acc};
    block_sum<double, 1>(tot);
    if (threadIdx.x == 0) {
      const int first = a.hub_first[h], nch = a.hub_nchunks[h];
      a.hub_partial[first + chunk] = tot[0];
      __threadfence();
      const uint32_t prev = atomicAdd(a.hub_count + h, 1u);
      if (prev == (uint32_t)(nch - 1)) {
        __threadfence();
        double s = 0.0;
        for (int c = 0; c < nch; c++) s += __builtin_nontemporal_load(a.hub_partial + first + c);
        a.hub_count[h] = 0;
        epilogue(a.hub_row[h], s);
      }
    }
  }

  double out[3];
  if (grid_reduce<3>(red, a.partials, &st->ticket, out)) {
    st->dsum_in = out[1];
    st->dsum_x = out[2];
    finish_sweep(st, out[0]);
  }
}

// ---------------------------------------------------------------------------
// init (x0 = 1/n, pagerank.rs:133, or the normalised personalization,
// ppr.rs:104-110) and fold (xbar += e, e = 0) both end with
// y64 = d*x*inv and the dangling sum of x.
template <bool INIT>
__global__ __launch_bounds__(kBlock) void fold_kernel(SweepArgs a, int n) {
  RunState *st = a.st;
  if (!INIT && st->done) return;
  const double d = st->damping;
  double red[1] = {0.0};
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < n; v += gridDim.x * kBlock) {
    double x;
    if (INIT) {
      x = a.q64 ? a.q64[v] : st->inv_n;
    } else {
      x = a.xbar[v] + (double)a.e[v];
    }
    a.xbar[v] = x;
    a.e[v] = 0.0f;
    const double inv = a.inv64[v];
    a.yout64[v] = (d * x) * inv;
    if (inv == 0.0) red[0] += x;
  }
  double out[1];
  if (grid_reduce<1>(red, a.partials, &st->ticket, out)) st->dsum_x = out[0];
}

// out[o] = xbar[rank[o]] + e[rank[o]] in the caller's ids; the last workgroup
// leaves the sum in st->dsum_x for the optional normalisation.
__global__ __launch_bounds__(kBlock) void finalize_kernel(SweepArgs a, const uint32_t *rank,
                                                          double *out, int n) {
  double red[1] = {0.0};
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) {
    const uint32_t r = rank[o];
    const double x = a.xbar[r] + (double)a.e[r];
    out[o] = x;
    red[0] += x;
  }
  double s[1];
  if (grid_reduce<1>(red, a.partials, &a.st->ticket, s)) a.st->dsum_x = s[0];
}

// reference normalize() semantics (topk.rs:54-61): divide by the sum if > 0
__global__ __launch_bounds__(kBlock) void scale_kernel(const RunState *st, double *out, int n) {
  const double s = st->dsum_x;
  if (!(s > 0.0)) return;
  for (int o = blockIdx.x * kBlock + threadIdx.x; o < n; o += gridDim.x * kBlock) out[o] /= s;
}

// ---------------------------------------------------------------------------
static SweepArgs make_args(gops_graph *g, const SweepSetup &s) {
  SweepArgs a{};
  DevGraph &dg = g->dev;
  RunBuffers &rb = g->run;
  a.desc = dg.desc;
  a.row_ptr = dg.row_ptr;
  a.col = dg.col;
  a.w = s.weighted ? dg.w : nullptr;
  a.inv32 = s.weighted ? dg.inv_ws32 : dg.inv_deg32;
  a.inv64 = s.weighted ? dg.inv_ws64 : dg.inv_deg64;
  a.e = rb.e;
  a.xbar = rb.xbar;
  a.q32 = s.personalized ? rb.q32 : nullptr;
  a.q64 = s.personalized ? rb.q64 : nullptr;
  a.partials = rb.partials;
  a.hub_partial = rb.hub_partial;
  a.hub_count = rb.hub_count;
  a.hub_row = dg.hub_row;
  a.hub_nchunks = dg.hub_nchunks;
  a.hub_first = dg.hub_first;
  a.st = rb.state;
  return a;
}

static int elementwise_grid(uint64_t n) {
  uint64_t blocks = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  return (int)blocks;
}

hipError_t launch_init(gops_context *ctx, gops_graph *g, const SweepSetup &s) {
  SweepArgs a = make_args(g, s);
  a.yout64 = g->run.y64[0];
  const int n = (int)g->dev.n;
  fold_kernel<true><<<elementwise_grid(n), kBlock, 0, ctx->stream>>>(a, n);
  return hipGetLastError();
}

hipError_t launch_fold(gops_context *ctx, gops_graph *g, const SweepSetup &s) {
  SweepArgs a = make_args(g, s);
  a.yout64 = g->run.y64[0];
  const int n = (int)g->dev.n;
  fold_kernel<false><<<elementwise_grid(n), kBlock, 0, ctx->stream>>>(a, n);
  return hipGetLastError();
}

template <typename WT>
static void launch_f32_wt(gops_context *ctx, const SweepArgs &a, int blocks) {
  sweep_f32_kernel<WT><<<blocks, kBlock, 0, ctx->stream>>>(a);
}

hipError_t launch_sweep_f32(gops_context *ctx, gops_graph *g, const SweepSetup &s, int parity) {
  SweepArgs a = make_args(g, s);
  a.yin32 = g->run.yd[parity];
  a.yout32 = g->run.yd[parity ^ 1];
  const int blocks = g->dev.num_desc;
  if (!s.weighted) {
    launch_f32_wt<NoW>(ctx, a, blocks);
  } else {
    switch (g->dev.weight_storage) {
      case GOPS_WEIGHT_F32: launch_f32_wt<float>(ctx, a, blocks); break;
      case GOPS_WEIGHT_BF16: launch_f32_wt<Bf16W>(ctx, a, blocks); break;
      default: launch_f32_wt<double>(ctx, a, blocks); break;
    }
  }
  return hipGetLastError();
}

template <typename WT>
static void launch_f64_wt(gops_context *ctx, const SweepArgs &a, int blocks, bool yd, bool y64) {
  if (yd && y64)
    sweep_f64_kernel<WT, true, true><<<blocks, kBlock, 0, ctx->stream>>>(a);
  else if (yd)
    sweep_f64_kernel<WT, true, false><<<blocks, kBlock, 0, ctx->stream>>>(a);
  else
    sweep_f64_kernel<WT, false, true><<<blocks, kBlock, 0, ctx->stream>>>(a);
}

// gathers y64[y64_in]; writes y64[y64_in ^ 1] (write_y64) and yd[yd_out]
// (write_yd).
hipError_t launch_sweep_f64(gops_context *ctx, gops_graph *g, const SweepSetup &s, int y64_in,
                            int yd_out, bool write_y64, bool write_yd) {
  SweepArgs a = make_args(g, s);
  a.yin64 = g->run.y64[y64_in];
  a.yout64 = g->run.y64[y64_in ^ 1];
  a.yout32 = g->run.yd[yd_out];
  const int blocks = g->dev.num_desc;
  if (!s.weighted) {
    launch_f64_wt<NoW>(ctx, a, blocks, write_yd, write_y64);
  } else {
    switch (g->dev.weight_storage) {
      case GOPS_WEIGHT_F32: launch_f64_wt<float>(ctx, a, blocks, write_yd, write_y64); break;
      case GOPS_WEIGHT_BF16: launch_f64_wt<Bf16W>(ctx, a, blocks, write_yd, write_y64); break;
      default: launch_f64_wt<double>(ctx, a, blocks, write_yd, write_y64); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_finalize(gops_context *ctx, gops_graph *g, int normalize) {
  SweepSetup s{0, 0};
  SweepArgs a = make_args(g, s);
  const int n = (int)g->dev.n;
  finalize_kernel<<<elementwise_grid(n), kBlock, 0, ctx->stream>>>(a, g->dev.rank, g->run.out, n);
  if (normalize)
    scale_kernel<<<elementwise_grid(n), kBlock, 0, ctx->stream>>>(g->run.state, g->run.out, n);
  return hipGetLastError();
}

}  // namespace gops
