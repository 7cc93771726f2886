// blocked_common.h -- device helpers of the blocked passes (pagerank.hip:
// PageRank's delta and anchor sweeps; centrality.hip: Katz): wave scans, the
// lanes' weight loads, the f64 scatter of one item's windows, and the 128-bit
// fixed-point row sums of the f64 accumulate.  The layout is described in
// internal.h (BlockedLayout) and built in blocked.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "internal.h"
#include "sweep_common.h"

namespace gops {

// An f32 DPP move (row_shr / row_bcast), 0 where the control selects no lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ float pb_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}

// Segmented inclusive sum over a wave: every lane adds the lanes of its run
// from the run's head (start: the highest lane at or below it that holds a
// head).  An inclusive scan within rows of 16 lanes (row_shr 1, 2, 4, 8),
// then across rows (row_bcast 15: rows 1 and 3 add the last lane of the row
// before; row_bcast 31: rows 2 and 3 add lane 31); a lane adds only lanes of
// its run.  The order of the additions depends only on the lanes' positions,
// so a run's sum is the same bits in every sweep.
__device__ __forceinline__ float pb_seg_scan(float v, int start) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  float o;
  o = pb_dpp<0x111, 0xF>(v); if (r >= 1 && lane - 1 >= start) v += o;
  o = pb_dpp<0x112, 0xF>(v); if (r >= 2 && lane - 2 >= start) v += o;
  o = pb_dpp<0x114, 0xF>(v); if (r >= 4 && lane - 4 >= start) v += o;
  o = pb_dpp<0x118, 0xF>(v); if (r >= 8 && lane - 8 >= start) v += o;
  o = pb_dpp<0x142, 0xA>(v); if (((lane >> 4) & 1) && (lane | 15) - 16 >= start) v += o;
  o = pb_dpp<0x143, 0xC>(v); if (lane >= 32 && 31 >= start) v += o;
  return v;
}

// Inclusive sum over a wave of a 32-bit integer (the same DPP steps, no
// segments: lanes outside a row read 0).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t pb_dppu(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, true);
}
__device__ __forceinline__ uint32_t pb_int_scan(uint32_t v) {
  v += pb_dppu<0x111, 0xF>(v);
  v += pb_dppu<0x112, 0xF>(v);
  v += pb_dppu<0x114, 0xF>(v);
  v += pb_dppu<0x118, 0xF>(v);
  v += pb_dppu<0x142, 0xA>(v);
  v += pb_dppu<0x143, 0xC>(v);
  return v;
}

typedef uint32_t pb_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t pb_u32x2 __attribute__((ext_vector_type(2)));

// A lane's kPbLane consecutive weights (padding entries carry 0), as f32.
template <typename WT>
__device__ __forceinline__ void pb_load_w8(const void *w, size_t e, float (&out)[kPbLane]) {
  static_assert(kPbLane == 8, "the loads below take eight entries");
  if constexpr (std::is_same<WT, Bf16W>::value) {
    const pb_u32x4 q = __builtin_nontemporal_load(
        reinterpret_cast<const pb_u32x4 *>(static_cast<const uint16_t *>(w) + e));
#pragma unroll
    for (int i = 0; i < 4; i++) {
      out[2 * i] = __uint_as_float(q[i] << 16);
      out[2 * i + 1] = __uint_as_float(q[i] & 0xffff0000u);
    }
  } else if constexpr (std::is_same<WT, float>::value) {
    const pb_u32x4 *p = reinterpret_cast<const pb_u32x4 *>(static_cast<const float *>(w) + e);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const pb_u32x4 q = __builtin_nontemporal_load(p + h);
#pragma unroll
      for (int i = 0; i < 4; i++) out[4 * h + i] = __uint_as_float(q[i]);
    }
  } else if constexpr (std::is_same<WT, double>::value) {
    const pb_u32x4 *p = reinterpret_cast<const pb_u32x4 *>(static_cast<const double *>(w) + e);
#pragma unroll
    for (int h = 0; h < 4; h++) {
      const pb_u32x4 q = __builtin_nontemporal_load(p + h);
      out[2 * h] = (float)__hiloint2double((int)q[1], (int)q[0]);
      out[2 * h + 1] = (float)__hiloint2double((int)q[3], (int)q[2]);
    }
  } else {  // unweighted: a byte 1 per edge, 0 per padding entry
    const pb_u32x2 q = __builtin_nontemporal_load(
        reinterpret_cast<const pb_u32x2 *>(static_cast<const uint8_t *>(w) + e));
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = (float)((q[i / 4] >> (8 * (i % 4))) & 0xffu);
  }
}

// Threads of a scatter workgroup (the f32 scatter holds two per CU, the f64
// scatter one).
constexpr int kScatterBlock = 1024;

constexpr int kAnchorFix = 100;  // fixed-point units of the anchors' row sums: 2^-100

// An f64 DPP move: both halves with the same control (0 where it selects no lane).
template <int CTRL, int ROWS>
__device__ __forceinline__ double pb_dpp64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xF, true);
  return __hiloint2double(hi, lo);
}

// pb_seg_scan in f64 (the same steps and order).
__device__ __forceinline__ double pb_seg_scan64(double v, int start) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  double o;
  o = pb_dpp64<0x111, 0xF>(v); if (r >= 1 && lane - 1 >= start) v += o;
  o = pb_dpp64<0x112, 0xF>(v); if (r >= 2 && lane - 2 >= start) v += o;
  o = pb_dpp64<0x114, 0xF>(v); if (r >= 4 && lane - 4 >= start) v += o;
  o = pb_dpp64<0x118, 0xF>(v); if (r >= 8 && lane - 8 >= start) v += o;
  o = pb_dpp64<0x142, 0xA>(v); if (((lane >> 4) & 1) && (lane | 15) - 16 >= start) v += o;
  o = pb_dpp64<0x143, 0xC>(v); if (lane >= 32 && 31 >= start) v += o;
  return v;
}

// A lane's kPbLane consecutive f64 weights (for the anchors' products; the
// other storage types are exact in f32 and take pb_load_w8).
__device__ __forceinline__ void pb_load_w8(const void *w, size_t e, double (&out)[kPbLane]) {
  const pb_u32x4 *p = reinterpret_cast<const pb_u32x4 *>(static_cast<const double *>(w) + e);
#pragma unroll
  for (int h = 0; h < 4; h++) {
    const pb_u32x4 q = __builtin_nontemporal_load(p + h);
    out[2 * h] = __hiloint2double((int)q[1], (int)q[0]);
    out[2 * h + 1] = __hiloint2double((int)q[3], (int)q[2]);
  }
}

// c * 2^fix as a signed 128-bit integer (lo, hi), truncated toward zero.
// Returns false for |c| >= 2^(126 - fix) (the anchors, fix = kAnchorFix: 2^26).
// Finite c.
__device__ __forceinline__ bool pb_fixed128(double c, int fix, unsigned long long &lo,
                                            unsigned long long &hi) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(c);
  const int ex = (int)((b >> 52) & 0x7ffu);
  lo = 0ull;
  hi = 0ull;
  if (ex == 0) return true;  // zero or subnormal: below one unit
  const unsigned long long mant = (b & 0xfffffffffffffull) | (1ull << 52);
  const int sh = ex - 1075 + fix;  // c = mant * 2^(ex - 1075)
  if (sh > 126 - 53) return false;
  if (sh >= 64) {
    hi = mant << (sh - 64);
  } else if (sh > 0) {
    lo = mant << sh;
    hi = mant >> (64 - sh);
  } else if (sh > -53) {
    lo = mant >> -sh;
  }
  if (b >> 63) {  // two's complement
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0ull ? 1ull : 0ull);
  }
  return true;
}

// The 128-bit sum (lo, hi) in units of 2^-fix as a double.
__device__ __forceinline__ double pb_value128(unsigned long long lo, unsigned long long hi, int fix) {
  const bool neg = (long long)hi < 0;
  if (neg) {
    lo = ~lo + 1ull;
    hi = ~hi + (lo == 0ull ? 1ull : 0ull);
  }
  const double v = ldexp((double)hi, 64 - fix) + ldexp((double)lo, -fix);
  return neg ? -v : v;
}

__device__ __forceinline__ double pb_row_value128(unsigned long long lo, unsigned long long hi,
                                                  uint32_t flags, int fix) {
  if (flags) {
    if ((flags & 1u) || (flags & 6u) == 6u) return __builtin_nan("");
    return (flags & 2u) ? __builtin_inf() : -__builtin_inf();
  }
  return pb_value128(lo, hi, fix);
}


// The f64 values of an item's chunk of sources into LDS z (kChunk of them).
__device__ __forceinline__ void pb_load_chunk64(const double *yin, int n, const PbItem &it,
                                                double *z) {
  constexpr int kChunk = 1 << kPbChunkShift;
  const int base = it.x << kPbChunkShift;
  const int len = min(kChunk, n - base);
  if (len == kChunk && (base & 1) == 0) {
    constexpr int V = kChunk / 2 / kScatterBlock;
    double2 v[V];
#pragma unroll
    for (int j = 0; j < V; j++)
      v[j] = *reinterpret_cast<const double2 *>(yin + base + 2 * ((int)threadIdx.x + j * kScatterBlock));
#pragma unroll
    for (int j = 0; j < V; j++)
      *reinterpret_cast<double2 *>(z + 2 * ((int)threadIdx.x + j * kScatterBlock)) = v[j];
  } else {
    for (int i = threadIdx.x; i < len; i += kScatterBlock) z[i] = yin[base + i];
  }
  __syncthreads();
}

// An f64 scatter item (z: its chunk's values in LDS): pb_scatter_kernel's
// windows and run logic on f64 products (f64 weights as stored, the others
// converted exactly) with f64 run sums, one per run to pb.rsum64.
template <typename WT>
__device__ __forceinline__ void pb_scatter64_windows(const BlockedLayout &pb, const PbItem &it,
                                                     const double *z) {
  const int lane = threadIdx.x & 63;
  const uint32_t nruns = (uint32_t)pb.num_runs, nseg = (uint32_t)pb.num_segs;
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
#ifndef GOPS_TUNE_S64U
#define GOPS_TUNE_S64U 4
#endif
  // windows' loads in flight per wave (f64 weights take twice the registers)
  constexpr int U = std::is_same<WT, double>::value ? 4 : GOPS_TUNE_S64U;
  constexpr int kWaves = kScatterBlock / 64;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int w0 = it.y + wave; w0 < it.z; w0 += U * kWaves) {
    using WV = typename std::conditional<std::is_same<WT, double>::value, double, float>::type;
    pb_u32x4 sq[U];
    WV wv[U][kPbLane];
    int run0[U], seg0[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int wi = min(w0 + u * kWaves, it.z - 1);  // clamped: every load in bounds
      const size_t e = (size_t)wi * kPbWin + (size_t)lane * kPbLane;
      sq[u] = __builtin_nontemporal_load(reinterpret_cast<const pb_u32x4 *>(pb.src + e));
      if constexpr (std::is_same<WT, double>::value)
        pb_load_w8(pb.w, e, wv[u]);
      else
        pb_load_w8<WT>(std::is_same<WT, NoW>::value ? (const void *)pb.mask : pb.w, e, wv[u]);
      run0[u] = pb.win[2 * wi];
      seg0[u] = pb.win[2 * wi + 1];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (w0 + u * kWaves >= it.z) break;  // the same in every lane
      uint32_t sw[kPbLane];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        sw[2 * i] = sq[u][i] & 0xffffu;
        sw[2 * i + 1] = sq[u][i] >> 16;
      }
      double q[kPbLane];
      uint32_t hm = 0;
#pragma unroll
      for (int j = 0; j < kPbLane; j++) {
        const double pj = (double)wv[u][j] * z[sw[j] & kPbSrcMask];
        const bool h = (sw[j] & kPbRunHead) != 0;
        hm |= (h ? 1u : 0u) << j;
        q[j] = (j == 0 || h) ? pj : q[j - 1] + pj;
      }
      const unsigned long long lheads = __ballot(hm != 0);
      const int start = 63 - __clzll(lheads & upto);
      const double S = pb_seg_scan64(q[kPbLane - 1], start);
      const double X = __shfl_up(S, 1);
      const uint32_t sh = (sw[0] & kPbSegHead) ? 1u : 0u;
      const uint32_t cnt = (uint32_t)__popc(hm) | (sh << 16);
      const uint32_t before = pb_int_scan(cnt) - cnt;
      const uint32_t sh0 = __builtin_amdgcn_readfirstlane(sh);
      const uint32_t run = (uint32_t)run0[u] + (before & 0xffffu) + (hm & 1u) - 1u;
      const uint32_t seg = (uint32_t)seg0[u] + (before >> 16) + sh - sh0;
      const int32_t off = pb.seg_off[seg < nseg ? seg : 0];
      const unsigned long long first_heads = __ballot((hm & 1u) != 0);
      const bool next_head = lane == 63 || ((first_heads >> (lane + 1)) & 1ull);
      uint32_t place = run + (uint32_t)off;
#pragma unroll
      for (int j = 0; j < kPbLane; j++) {
        const bool tail = j + 1 < kPbLane ? ((hm >> (j + 1)) & 1u) != 0 : next_head;
        const double t = (hm & ((2u << j) - 1u)) ? q[j] : q[j] + X;
        if (tail) {
          if (seg < nseg && place < nruns) pb.rsum64[place] = t;
          place++;
        }
      }
    }
  }
}

// The runs of accumulate item it added to their rows' 128-bit fixed-point
// sums in LDS (units 2^-fix): an atomic on the low word, whose returned old
// value gives the carry into an atomic on the high word, so a row's sum is
// exact whatever the order.  A non-finite run sum, or one beyond the format,
// sets the row's flag byte in nonfin (1 NaN or beyond, 2 +inf, 4 -inf).
template <int BLOCK>
__device__ __forceinline__ void pb_accumulate128_runs(const BlockedLayout &pb, const PbItem &it,
                                                      int fix, unsigned long long *acc_lo,
                                                      unsigned long long *acc_hi,
                                                      uint32_t *nonfin) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int U = 8;
  if (it.y < it.z) {
    for (int f0 = it.y & ~63; f0 < it.z; f0 += U * BLOCK) {
      double cv[U];
      int key[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int f = f0 + u * BLOCK + wave * 64 + lane;
        const bool ok = f >= it.y && f < it.z;
        const int ff = ok ? f : it.y;
        cv[u] = __builtin_nontemporal_load(pb.rsum64 + ff);
        const int d = __builtin_nontemporal_load(pb.rrow + ff);
        key[u] = ok ? d : -1;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (key[u] < 0) continue;
        const double c = cv[u];
        unsigned long long lo, hi;
        uint32_t bad = 0u;
        if (!isfinite(c)) bad = isnan(c) ? 1u : (c > 0.0 ? 2u : 4u);
        else if (!pb_fixed128(c, fix, lo, hi)) bad = 1u;
        if (bad) {
          atomicOr(&nonfin[key[u] >> 2], bad << (8 * (key[u] & 3)));
        } else {
          const unsigned long long old = atomicAdd(&acc_lo[key[u]], lo);
          atomicAdd(&acc_hi[key[u]], hi + (old + lo < old ? 1ull : 0ull));
        }
      }
    }
  }
}

// Row i of a split bin: its parts' 128-bit sums (slot order) and flags.
__device__ __forceinline__ void pb_combine128_row(const BlockedLayout &pb, const PbItem &sp, int i,
                                                  unsigned long long &lo, unsigned long long &hi,
                                                  uint32_t &fl) {
  const long long *bp = pb.bin_partial + ((size_t)sp.y << (kPbBinShift + 1)) + 2 * (size_t)i;
  const uint8_t *fp = pb.bin_flags + ((size_t)sp.y << kPbBinShift) + i;
  lo = 0ull;
  hi = 0ull;
  fl = 0;
  int j = 0;
  for (; j + 4 <= sp.z; j += 4) {
    unsigned long long vl[4], vh[4];
    uint32_t f[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const size_t at = (size_t)(j + u) << (kPbBinShift + 1);
      vl[u] = (unsigned long long)bp[at];
      vh[u] = (unsigned long long)bp[at + 1];
      f[u] = fp[(size_t)(j + u) << kPbBinShift];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const unsigned long long old = lo;
      lo += vl[u];
      hi += vh[u] + (lo < old ? 1ull : 0ull);
      fl |= f[u];
    }
  }
  for (; j < sp.z; j++) {
    const size_t at = (size_t)j << (kPbBinShift + 1);
    const unsigned long long old = lo;
    lo += (unsigned long long)bp[at];
    hi += (unsigned long long)bp[at + 1] + (lo < old ? 1ull : 0ull);
    fl |= fp[(size_t)j << kPbBinShift];
  }
}

}  // namespace gops
