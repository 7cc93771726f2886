// sweep_common.h -- device building blocks shared by the sweep kernels:
// weight loads, cross-lane and workgroup sums, and the load-balanced pull of
// one work descriptor (a run of short rows staged in LDS, one long row, or
// one chunk of a hub row) over the CSR of a graph.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"

namespace gops {

struct NoW {};
struct Bf16W {};

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
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = dpp_i<CTRL>(__double2loint(v));
  const int hi = dpp_i<CTRL>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
// ds_swizzle bit mode: and_mask 0x1f, or_mask 0, xor_mask 0x10 -> lane ^ 16 within 32
constexpr int kSwzXor16 = 0x1F | (0x10 << 10);

__device__ __forceinline__ double group_sum(double v, int lg) {
  if (lg >= 1) v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  if (lg >= 2) v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  if (lg >= 3) v += dpp_d<0x141>(v);  // row_half_mirror
  if (lg >= 4) v += dpp_d<0x140>(v);  // row_mirror
  if (lg >= 5) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), kSwzXor16);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), kSwzXor16);
    v += __hiloint2double(hi, lo);
  }
  if (lg >= 6) v += __shfl_xor(v, 32);
  return v;
}

// Sum of `v` over the workgroup, valid in thread 0.  Fixed order.  scratch:
// N * kBlock / 64 doubles of LDS that no other part of the workgroup uses
// meanwhile (a sweep passes its idle stage buffer, so that the kernel's LDS
// stays at the stage's size).
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *scratch) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int W = kBlock / 64;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double s = group_sum(v[i], 6);
    if (lane == 0) scratch[i * W + wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++)
      v[i] = (scratch[i * W] + scratch[i * W + 1]) + (scratch[i * W + 2] + scratch[i * W + 3]);
  }
  __syncthreads();
}

// block_sum for a workgroup of W waves (fixed order: wave 0, 1, ...).
template <int N, int W>
__device__ __forceinline__ void block_sum_waves(double (&v)[N], double *scratch) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double s = group_sum(v[i], 6);
    if (lane == 0) scratch[i * W + wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      double t = scratch[i * W];
      for (int w = 1; w < W; w++) t += scratch[i * W + w];
      v[i] = t;
    }
  }
  __syncthreads();
}

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
  __shared__ double sh[N * (kBlock / 64)];
  block_sum<N>(v, sh);
}

// Sum over edges [e0, e1) by the whole workgroup, kUnroll independent
// loads in flight per lane (out-of-range slots re-read edge e0 and count 0).
// Waves start on a multiple of 64 edges, so that each wave's loads of the
// col and weight streams cover whole 128 B lines: from an unaligned start a
// wave touches one more line per load, and the line two waves share is
// fetched by both.
template <typename LOAD>
__device__ __forceinline__ double stream_sum(const LOAD &load, int e0, int e1) {
  double acc = 0.0;
  for (int k = (e0 & ~63) + (int)threadIdx.x; k < e1; k += kUnroll * kBlock) {
    double v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const int kk = k + u * kBlock;
      const bool ok = kk >= e0 && kk < e1;
      const double x = load(ok ? kk : e0);
      v[u] = ok ? x : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc += v[u];
  }
  return acc;
}

// One work descriptor: a run of rows (2^lg lanes per row), or one chunk of a
// hub row whose sum goes to hub_partial for the finish kernel.
// A: any argument struct with row_ptr, col, hub_partial and hub_first.
template <typename A, typename LOAD, typename EPI, int R>
__device__ __forceinline__ void do_desc(const A &a, const BlockDesc &bd, const LOAD &load,
                                        const EPI &epi, double (&red)[R], double *stage) {
  if ((bd.x & (kHubFlag | kRowFlag)) == 0) {
    const int lg = bd.x;
    const int G = 1 << lg;
    const int lane = threadIdx.x & (G - 1);
    const int grp = threadIdx.x >> lg;
    const int ngrp = kBlock >> lg;
    // CSR-stream: the run's edges (<= kStageNnz) are streamed coalesced,
    // several independent loads in flight per lane, and their products
    // staged in LDS (f32 or f64, one buffer); then each group of G lanes
    // reduces one row from LDS, in the order of a direct pull (lane, lane +
    // G, ...; then across the group).  stage: kStageNnz doubles of LDS.
    // Lanes are laid over [e0 - off, e0 + nnz) with off = e0 mod 64, so that
    // waves start on whole lines (as in stream_sum); the planner keeps off +
    // nnz <= kStageNnz for runs of several rows, so that an f32 run still
    // takes one pass of U loads per lane.
    using P = typename LOAD::P;
    constexpr int U = LOAD::kStageUnroll;
    P *prod = reinterpret_cast<P *>(stage);
    const int e0 = a.row_ptr[bd.y];
    const int nnz = a.row_ptr[bd.z] - e0;
    const int off = e0 & 63;
    for (int i = threadIdx.x; i < off + nnz; i += U * kBlock) {
      int c[U];
      P wv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int ii = i + u * kBlock - off;
        const int k = e0 + ((unsigned)ii < (unsigned)nnz ? ii : 0);
        c[u] = __builtin_nontemporal_load(a.col + k);
        wv[u] = load.stage_weight(k);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int ii = i + u * kBlock - off;
        const P v = wv[u] * load.stage_gather(c[u]);
        if ((unsigned)ii < (unsigned)nnz) prod[ii] = v;
      }
    }
    __syncthreads();
    for (int base = bd.y; base < bd.z; base += ngrp) {
      const int r = base + grp;
      double acc = 0.0;
      if (r < bd.z) {
        const int b = a.row_ptr[r] - e0;
        const int e = a.row_ptr[r + 1] - e0;
        for (int k = b + lane; k < e; k += G) acc += (double)prod[k];
      }
      acc = group_sum(acc, lg);
      if (r < bd.z && lane == 0) epi(r, acc, red);
    }
    __syncthreads();
  } else if (bd.x & kRowFlag) {
    // one long row, the whole workgroup
    const int r = bd.y;
    double acc[1] = {stream_sum(load, a.row_ptr[r], a.row_ptr[r + 1])};
    block_sum<1>(acc, stage);
    if (threadIdx.x == 0) epi(r, acc[0], red);
  } else {
    // one chunk of a split hub row; the finish kernel adds the chunks
    const int h = bd.y;
    const int chunk = bd.x & (kHubFlag - 1);
    double acc[1] = {stream_sum(load, bd.z, bd.w)};
    block_sum<1>(acc, stage);
    if (threadIdx.x == 0) a.hub_partial[a.hub_first[h] + chunk] = acc[0];
  }
}

}  // namespace gops
