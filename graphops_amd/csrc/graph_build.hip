// graph_build.hip -- the preprocessed graph (CSR of in-edges, rows renumbered
// by in-degree) built on the device.  The arrays are those of build_csrt on
// the host, bit for bit, from the adapter's neighbour lists uploaded once.
//
// Steps, each a grid-stride kernel or a hipCUB primitive over the graph:
//   1. the largest id >= n, if any (the host's bounds message);
//   2. per-edge source ids: the first entry of each non-empty neighbour list
//      marked with its node, then an inclusive max-scan;
//   3. per node, 1/out-degree and 1/out_wsum with out_wsum summed in
//      neighbour order (pagerank.rs:137, 227-234), and the first invalid
//      weight in adapter order (pagerank.rs:300-314): a lane per short list,
//      a wave per long one;
//   4. in-degrees: the targets radix-sorted, each id's run measured (no
//      device-wide atomics: on a chip of several L2s they go to memory);
//   5. the order: a stable radix sort of ids by descending in-degree, the
//      ids without in- and out-edges after the other ids without in-edges,
//      so that ties keep ascending ids;
//   6. row pointers: an exclusive scan of the in-degrees in that order;
//   7. the edges: a stable radix sort of (row, new source id) keys carrying
//      max(w, 0) in the storage type, so duplicate edges keep adapter order,
//      the order of the reference's push loop (pagerank.rs:154-162); the
//      sorted values are the graph's weight array;
//   8. col from the sorted keys.
// Only the row pointers come back to the host, for the block plan, with the
// order and rank that map query seeds and results.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace gops {

namespace {

constexpr int kB = 256;

inline int grid_for(uint64_t count) {
  const uint64_t g = (count + kB - 1) / kB;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 16));
}

#define GRID_STRIDE(i, count) \
  for (uint64_t i = blockIdx.x * (uint64_t)kB + threadIdx.x; i < (count); i += (uint64_t)gridDim.x * kB)

// *bad = largest id >= n, plus one (0 when every id is in range)
__global__ void bad_ids_kernel(const uint32_t *nb, uint64_t m, uint64_t n, unsigned long long *bad) {
  GRID_STRIDE(k, m) {
    if (nb[k] >= n) atomicMax(bad, (unsigned long long)nb[k] + 1ull);
  }
}

// marks[off[u]] = u for each u with neighbours (their starts are distinct)
__global__ void mark_starts_kernel(const uint64_t *off, uint64_t n, uint32_t *marks) {
  GRID_STRIDE(u, n) {
    if (off[u + 1] > off[u]) marks[off[u]] = (uint32_t)u;
  }
}

// per node (original ids): 1/out-degree, and for a list of at most
// kShortList entries 1/out_wsum (0 when 0) and the smallest adapter position
// of a weight that is negative or not finite; longer lists are flagged for
// long_list_kernel.
constexpr uint64_t kShortList = 64;
__global__ void node_kernel(const uint64_t *off, const double *ew, uint64_t n, double *inv_deg_o,
                            double *inv_ws_o, unsigned long long *first_bad, uint8_t *is_long) {
  GRID_STRIDE(u, n) {
    const uint64_t b = off[u], e = off[u + 1];
    inv_deg_o[u] = e > b ? 1.0 / (double)(e - b) : 0.0;
    if (!ew) continue;
    const bool longl = e - b > kShortList;
    is_long[u] = longl;
    if (longl) continue;
    double s = 0.0;
    uint64_t bad = ~0ull;
    for (uint64_t j = b; j < e; j++) {
      const double w = ew[j];
      if ((!isfinite(w) || w < 0.0) && bad == ~0ull) bad = j;
      s += fmax(w, 0.0);
    }
    inv_ws_o[u] = s != 0.0 ? 1.0 / s : 0.0;
    if (bad != ~0ull) atomicMin(first_bad, (unsigned long long)bad);
  }
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// the flagged lists, one wave each (hubs have ids close together -- R-MAT
// puts them at the low ids -- so one wave per list keeps them apart).  The
// wave loads 256 entries per step, the next step's while it adds the
// current one's in list order (one broadcast each), so the sum is the
// sequential one and the chain of additions is the only wait.
constexpr int kLongChunks = 4;
__global__ void long_list_kernel(const uint64_t *off, const double *ew, const uint32_t *nodes,
                                 const uint32_t *count, double *inv_ws_o,
                                 unsigned long long *first_bad) {
  const int lane = threadIdx.x & 63;
  const uint32_t num = *count;
  const uint32_t waves = gridDim.x * (kB / 64);
  const uint32_t first = __builtin_amdgcn_readfirstlane((blockIdx.x * kB + threadIdx.x) / 64);
  for (uint32_t i = first; i < num; i += waves) {
    const uint32_t u = __builtin_amdgcn_readfirstlane(nodes[i]);
    const uint64_t b = uniform_u64(off[u]), e = uniform_u64(off[u + 1]);
    double acc = 0.0;
    uint64_t bad = ~0ull;
    double cur[kLongChunks], nxt[kLongChunks];
    // unconditional loads (indices clamped to the list) keep the waits exact:
    // the adds wait for the current step's loads only
#pragma unroll
    for (int q = 0; q < kLongChunks; q++) cur[q] = ew[min(b + 64 * q + lane, e - 1)];
    for (uint64_t j0 = b; j0 < e; j0 += 64 * kLongChunks) {
      const uint64_t jn = j0 + 64 * kLongChunks;
#pragma unroll
      for (int q = 0; q < kLongChunks; q++) nxt[q] = ew[min(jn + 64 * q + lane, e - 1)];
#pragma unroll
      for (int q = 0; q < kLongChunks; q++) {
        const uint64_t c0 = j0 + 64 * q;
        if (c0 >= e) break;
        const double w = cur[q];
        const uint64_t bm = __ballot(c0 + lane < e && (!isfinite(w) || w < 0.0));
        if (bm && bad == ~0ull) bad = c0 + __ffsll((unsigned long long)bm) - 1;
        const double wc = fmax(w, 0.0);
        if (e - c0 >= 64) {
#pragma unroll
          for (int k = 0; k < 64; k++) acc += readlane_f64(wc, k);
        } else {
          const int cnt = (int)(e - c0);
          for (int k = 0; k < cnt; k++) acc += readlane_f64(wc, k);
        }
      }
#pragma unroll
      for (int q = 0; q < kLongChunks; q++) cur[q] = nxt[q];
    }
    if (lane == 0) {
      inv_ws_o[u] = acc != 0.0 ? 1.0 / acc : 0.0;
      if (bad != ~0ull) atomicMin(first_bad, (unsigned long long)bad);
    }
  }
}

// run [beg, end) of each id present in the sorted targets (absent ids keep
// beg = end = 0)
__global__ void runs_kernel(const uint32_t *t, uint64_t m, uint32_t *beg, uint32_t *end) {
  GRID_STRIDE(i, m) {
    const uint32_t v = t[i];
    if (i == 0 || t[i - 1] != v) beg[v] = (uint32_t)i;
    if (i + 1 == m || t[i + 1] != v) end[v] = (uint32_t)(i + 1);
  }
}

// indeg = end - beg; ascending key = descending in-degree, and among ids
// without in-edges the isolated ones (no out-edges either) last: key =
// 2^32 - 1 - (2 indeg + 1), or 2^32 - 1 for an isolated id (indeg < 2^31);
// value = the id
__global__ void relabel_keys_kernel(const uint32_t *beg, const uint32_t *end, const uint64_t *off,
                                    uint64_t n, uint32_t *indeg, uint32_t *key, uint32_t *val) {
  GRID_STRIDE(v, n) {
    const uint32_t d = end[v] - beg[v];
    indeg[v] = d;
    const bool isolated = d == 0 && off[v + 1] == off[v];
    key[v] = isolated ? 0xFFFFFFFFu : 0xFFFFFFFFu - (2u * d + 1u);
    val[v] = (uint32_t)v;
  }
}

__global__ void rank_kernel(const uint32_t *order, const uint32_t *indeg, uint64_t n, uint32_t *rank,
                            int32_t *cnt) {
  GRID_STRIDE(r, n) {
    const uint32_t v = order[r];
    rank[v] = (uint32_t)r;
    cnt[r] = (int32_t)indeg[v];
  }
}

// f32 -> bf16, round to nearest even, NaN kept quiet (as f32_to_bf16)
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) {
    uint16_t h = (uint16_t)(u >> 16);
    if (u & 0x007fffffu) h |= 0x40;
    return h;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <typename V>
__device__ __forceinline__ V stored(double x);
template <>
__device__ __forceinline__ double stored<double>(double x) { return x; }
template <>
__device__ __forceinline__ float stored<float>(double x) { return (float)x; }
template <>
__device__ __forceinline__ uint16_t stored<uint16_t>(double x) { return to_bf16((float)x); }

// key = rank[target] << cb | rank[source]; value = max(w, 0) as stored
template <typename V>
__global__ void edge_keys_kernel(const uint32_t *nb, const uint32_t *src, const uint32_t *rank,
                                 const double *ew, uint64_t m, int cb, uint64_t *key, V *val) {
  GRID_STRIDE(k, m) {
    key[k] = ((uint64_t)rank[nb[k]] << cb) | (uint64_t)rank[src[k]];
    if (ew) val[k] = stored<V>(fmax(ew[k], 0.0));
  }
}

__global__ void col_kernel(const uint64_t *key, uint64_t m, uint64_t col_mask, int32_t *col) {
  GRID_STRIDE(i, m) { col[i] = (int32_t)(key[i] & col_mask); }
}

// per row (new ids): the f64 value and its f32 copy; a non-zero value that
// underflows in f32 keeps the smallest subnormal so it stays non-dangling
__global__ void inv_kernel(const uint32_t *order, const double *src, uint64_t n, double *d64,
                           float *d32) {
  GRID_STRIDE(r, n) {
    const double x = src[order[r]];
    d64[r] = x;
    float f = (float)x;
    if (x != 0.0 && f == 0.0f) f = std::numeric_limits<float>::denorm_min();
    d32[r] = f;
  }
}

// marks[row_ptr[r]] = r for each non-empty row (their starts are distinct)
__global__ void mark_rows_kernel(const int32_t *row_ptr, uint64_t n, uint32_t *marks) {
  GRID_STRIDE(r, n) {
    if (row_ptr[r + 1] > row_ptr[r]) marks[row_ptr[r]] = (uint32_t)r;
  }
}

// out-degree of each row's node from its 1/out-degree (RN(1/RN(1/d)) = d
// for every d < 2^51); ascending key = descending out-degree; value = row
__global__ void outdeg_keys_kernel(const double *inv_deg, uint64_t n, uint32_t *deg,
                                   uint32_t *key, uint32_t *val) {
  GRID_STRIDE(r, n) {
    const double inv = inv_deg[r];
    const uint32_t d = inv != 0.0 ? (uint32_t)rint(1.0 / inv) : 0u;
    deg[r] = d;
    key[r] = 0xFFFFFFFFu - d;
    val[r] = (uint32_t)r;
  }
}

// key = out-row of the source << cb | target (k: edge of the in-edge CSR)
__global__ void fwd_keys_kernel(const int32_t *col, const uint32_t *row, const uint32_t *frank,
                                uint64_t m, int cb, uint64_t *key) {
  GRID_STRIDE(k, m) { key[k] = ((uint64_t)frank[col[k]] << cb) | row[k]; }
}

#undef GRID_STRIDE

// device allocations of one build; whatever is not released is freed
struct Pool {
  std::vector<void *> ptrs;
  hipError_t err = hipSuccess;
  template <typename T>
  T *get(size_t count) {
    if (err != hipSuccess) return nullptr;
    void *p = nullptr;
    err = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (err != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
  void release(void *p) { ptrs.erase(std::remove(ptrs.begin(), ptrs.end(), p), ptrs.end()); }
  ~Pool() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

// bits that hold every value in [0, x]
int bits_for(uint64_t x) {
  int b = 1;
  while (b < 63 && (1ull << b) <= x) b++;
  return b;
}

// the edge sort and the kernel that feeds it, for one weight type (V unused
// when unweighted: keys only)
template <typename V>
hipError_t sort_edges(void *tmp, size_t &bytes, const uint32_t *nb, const uint32_t *src,
                      const uint32_t *rank, const double *ew, uint64_t m, int cb, uint64_t *k_in,
                      uint64_t *k_out, V *v_in, V *v_out, hipStream_t s) {
  if (tmp) {
    edge_keys_kernel<V><<<grid_for(m), kB, 0, s>>>(nb, src, rank, ew, m, cb, k_in, v_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || m == 0) return e;
  }
  if (ew)
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k_in, k_out, v_in, v_out, (int)m, 0,
                                              2 * cb, s);
  return hipcub::DeviceRadixSort::SortKeys(tmp, bytes, k_in, k_out, (int)m, 0, 2 * cb, s);
}

hipError_t sort_edges_for(int storage, void *tmp, size_t &bytes, const uint32_t *nb,
                          const uint32_t *src, const uint32_t *rank, const double *ew, uint64_t m,
                          int cb, uint64_t *k_in, uint64_t *k_out, void *v_in, void *v_out,
                          hipStream_t s) {
  switch (storage) {
    case GOPS_WEIGHT_F64:
      return sort_edges(tmp, bytes, nb, src, rank, ew, m, cb, k_in, k_out,
                        static_cast<double *>(v_in), static_cast<double *>(v_out), s);
    case GOPS_WEIGHT_F32:
      return sort_edges(tmp, bytes, nb, src, rank, ew, m, cb, k_in, k_out,
                        static_cast<float *>(v_in), static_cast<float *>(v_out), s);
    default:  // bf16, or unweighted (keys only)
      return sort_edges(tmp, bytes, nb, src, rank, ew, m, cb, k_in, k_out,
                        static_cast<uint16_t *>(v_in), static_cast<uint16_t *>(v_out), s);
  }
}

}  // namespace

gops_status build_on_device(gops_context *ctx, uint64_t n, const uint64_t *off, const uint32_t *nb,
                            const double *ew, int weight_storage, gops_graph *g, Plan &plan,
                            std::string &err) {
  // offset checks, as build_csrt
  if (n >= (1ull << 31)) {
    err = "node count must be < 2^31";
    return GOPS_NOT_SUPPORTED;
  }
  const uint64_t m = off[n];
  if (m >= (1ull << 31)) {
    err = "edge count must be < 2^31";
    return GOPS_NOT_SUPPORTED;
  }
  if (off[0] != 0) {
    err = "offsets[0] must be 0";
    return GOPS_INVALID_PARAMETER;
  }
  for (uint64_t u = 0; u < n; u++) {
    if (off[u + 1] < off[u]) {
      err = "offsets must be non-decreasing";
      return GOPS_INVALID_PARAMETER;
    }
  }
  hipStream_t s = ctx->stream;
  Pool pool;
  hipError_t herr = hipSuccess;
  const char *what = "";
#define DTRY(expr)            \
  do {                        \
    herr = (expr);            \
    if (herr != hipSuccess) { \
      what = #expr;           \
      goto hip_error;         \
    }                         \
  } while (0)
  {
    const int storage = ew ? weight_storage : GOPS_WEIGHT_NONE;
    const size_t wbytes = storage == GOPS_WEIGHT_F64 ? 8 : storage == GOPS_WEIGHT_F32 ? 4 : 2;
    uint64_t *d_off = pool.get<uint64_t>(n + 1);
    uint32_t *d_nb = pool.get<uint32_t>(m);
    double *d_ew = ew ? pool.get<double>(m) : nullptr;
    unsigned long long *d_flags = pool.get<unsigned long long>(2);  // bad id + 1, first bad weight
    unsigned long long flags[2] = {0ull, ~0ull};
    DTRY(pool.err);
    DTRY(staged_upload(ctx, d_off, off, (n + 1) * sizeof(uint64_t), s));
    if (m) DTRY(staged_upload(ctx, d_nb, nb, m * sizeof(uint32_t), s));
    if (ew && m) DTRY(staged_upload(ctx, d_ew, ew, m * sizeof(double), s));
    DTRY(hipMemcpyAsync(d_flags, flags, sizeof(flags), hipMemcpyHostToDevice, s));
    bad_ids_kernel<<<grid_for(m), kB, 0, s>>>(d_nb, m, n, d_flags);
    DTRY(hipGetLastError());
    DTRY(hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    DTRY(hipStreamSynchronize(s));
    if (flags[0]) {
      err = "index out of bounds: " + std::to_string(flags[0] - 1);
      return GOPS_INDEX_OUT_OF_BOUNDS;
    }

    const int cb = bits_for(n ? n - 1 : 0);  // bits of an id
    uint32_t *d_marks = pool.get<uint32_t>(m), *d_src = pool.get<uint32_t>(m);
    double *d_inv_deg_o = pool.get<double>(n);
    double *d_inv_ws_o = ew ? pool.get<double>(n) : nullptr;
    uint32_t *d_indeg = pool.get<uint32_t>(n);
    uint32_t *d_key32 = pool.get<uint32_t>(n), *d_key32b = pool.get<uint32_t>(n);
    uint32_t *d_val32 = pool.get<uint32_t>(n);
    uint32_t *d_order = pool.get<uint32_t>(n), *d_rank = pool.get<uint32_t>(n);
    int32_t *d_cnt = pool.get<int32_t>(n + 1), *d_row_ptr = pool.get<int32_t>(n + 1);
    uint64_t *d_key64 = pool.get<uint64_t>(m), *d_key64b = pool.get<uint64_t>(m);
    void *d_wv = ew ? (void *)pool.get<uint8_t>(m * wbytes) : nullptr;
    void *d_w = ew ? (void *)pool.get<uint8_t>(m * wbytes) : nullptr;  // sorted: the graph's
    DTRY(pool.err);

    // scratch of the hipCUB calls: the largest of their needs
    size_t tmp_bytes = 1, b = 0;
    DTRY(hipcub::DeviceScan::InclusiveScan(nullptr, b, d_marks, d_src, hipcub::Max(), (int)m, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceRadixSort::SortKeys(nullptr, b, d_nb, d_marks, (int)m, 0, cb, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, d_key32, d_key32b, d_val32, d_order, (int)n,
                                            0, 32, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, d_cnt, d_row_ptr, (int)n + 1, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(sort_edges_for(storage, nullptr, b, d_nb, d_src, d_rank, d_ew, m, cb, d_key64, d_key64b,
                        d_wv, d_w, s));
    tmp_bytes = std::max(tmp_bytes, b);
    uint8_t *d_is_long = ew ? pool.get<uint8_t>(n) : nullptr;
    uint32_t *d_long = ew ? pool.get<uint32_t>(n) : nullptr;
    uint32_t *d_num_long = pool.get<uint32_t>(1);
    DTRY(pool.err);
    if (ew) {
      DTRY(hipcub::DeviceSelect::Flagged(nullptr, b, hipcub::CountingInputIterator<uint32_t>(0),
                                         d_is_long, d_long, d_num_long, (int64_t)n, s));
      tmp_bytes = std::max(tmp_bytes, b);
    }
    void *d_tmp = pool.get<uint8_t>(tmp_bytes);
    DTRY(pool.err);

    // 2-3: sources, per-node sums
    if (m) {
      DTRY(hipMemsetAsync(d_marks, 0, m * sizeof(uint32_t), s));
      mark_starts_kernel<<<grid_for(n), kB, 0, s>>>(d_off, n, d_marks);
      b = tmp_bytes;
      DTRY(hipcub::DeviceScan::InclusiveScan(d_tmp, b, d_marks, d_src, hipcub::Max(), (int)m, s));
    }
    node_kernel<<<grid_for(n), kB, 0, s>>>(d_off, d_ew, n, d_inv_deg_o, d_inv_ws_o, d_flags + 1,
                                           d_is_long);
    DTRY(hipGetLastError());
    if (ew && n) {
      b = tmp_bytes;
      DTRY(hipcub::DeviceSelect::Flagged(d_tmp, b, hipcub::CountingInputIterator<uint32_t>(0),
                                         d_is_long, d_long, d_num_long, (int64_t)n, s));
      long_list_kernel<<<8192, kB, 0, s>>>(d_off, d_ew, d_long, d_num_long, d_inv_ws_o,
                                          d_flags + 1);
      DTRY(hipGetLastError());
    }

    // 4: in-degrees (the marks buffer holds the sorted targets, the key
    // buffers the runs)
    if (n) {
      DTRY(hipMemsetAsync(d_key32, 0, n * sizeof(uint32_t), s));
      DTRY(hipMemsetAsync(d_key32b, 0, n * sizeof(uint32_t), s));
    }
    if (m) {
      b = tmp_bytes;
      DTRY(hipcub::DeviceRadixSort::SortKeys(d_tmp, b, d_nb, d_marks, (int)m, 0, cb, s));
      runs_kernel<<<grid_for(m), kB, 0, s>>>(d_marks, m, d_key32, d_key32b);
    }
    relabel_keys_kernel<<<grid_for(n), kB, 0, s>>>(d_key32, d_key32b, d_off, n, d_indeg, d_key32,
                                                   d_val32);
    DTRY(hipGetLastError());

    // 5-6: order, rank, row pointers
    if (n) {
      b = tmp_bytes;
      DTRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, b, d_key32, d_key32b, d_val32, d_order, (int)n,
                                              0, 32, s));
    }
    DTRY(hipMemsetAsync(d_cnt + n, 0, sizeof(int32_t), s));
    rank_kernel<<<grid_for(n), kB, 0, s>>>(d_order, d_indeg, n, d_rank, d_cnt);
    DTRY(hipGetLastError());
    b = tmp_bytes;
    DTRY(hipcub::DeviceScan::ExclusiveSum(d_tmp, b, d_cnt, d_row_ptr, (int)n + 1, s));

    // 7-8: edges by (row, new source id), ties in adapter order; col
    b = tmp_bytes;
    DTRY(sort_edges_for(storage, d_tmp, b, d_nb, d_src, d_rank, d_ew, m, cb, d_key64, d_key64b,
                        d_wv, d_w, s));
    int32_t *d_col = pool.get<int32_t>(m);
    double *d_inv_deg64 = pool.get<double>(n);
    float *d_inv_deg32 = pool.get<float>(n);
    double *d_inv_ws64 = ew ? pool.get<double>(n) : nullptr;
    float *d_inv_ws32 = ew ? pool.get<float>(n) : nullptr;
    DTRY(pool.err);
    col_kernel<<<grid_for(m), kB, 0, s>>>(d_key64b, m, (1ull << cb) - 1, d_col);
    inv_kernel<<<grid_for(n), kB, 0, s>>>(d_order, d_inv_deg_o, n, d_inv_deg64, d_inv_deg32);
    if (ew) inv_kernel<<<grid_for(n), kB, 0, s>>>(d_order, d_inv_ws_o, n, d_inv_ws64, d_inv_ws32);
    DTRY(hipGetLastError());

    // host copies: row pointers for the plan, ids for query seeds and results
    HostCSRT h;
    h.n = n;
    h.m = m;
    h.row_ptr.resize(n + 1);
    std::vector<uint32_t> order(n), rank(n);
    DTRY(staged_download(ctx, h.row_ptr.data(), d_row_ptr, (n + 1) * sizeof(int32_t), s));
    if (n) {
      DTRY(staged_download(ctx, order.data(), d_order, n * sizeof(uint32_t), s));
      DTRY(staged_download(ctx, rank.data(), d_rank, n * sizeof(uint32_t), s));
    }
    DTRY(hipMemcpyAsync(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    DTRY(hipStreamSynchronize(s));
    plan_blocks(h, plan);

    // success: the graph takes the arrays it keeps
    DevGraph &dg = g->dev;
    dg.n = n;
    dg.m = m;
    dg.weight_storage = storage;
    dg.row_ptr = d_row_ptr;
    dg.col = d_col;
    dg.w = d_w;
    dg.rank = d_rank;
    dg.order = d_order;
    dg.inv_deg64 = d_inv_deg64;
    dg.inv_deg32 = d_inv_deg32;
    dg.inv_ws64 = d_inv_ws64;
    dg.inv_ws32 = d_inv_ws32;
    for (void *p : {(void *)d_row_ptr, (void *)d_col, d_w, (void *)d_rank, (void *)d_order,
                    (void *)d_inv_deg64, (void *)d_inv_deg32, (void *)d_inv_ws64,
                    (void *)d_inv_ws32})
      if (p) pool.release(p);
    g->order_host = std::move(order);
    g->rank_host = std::move(rank);
    g->first_bad_weight = 0;
    if (ew && flags[1] != ~0ull) g->first_bad_weight = std::isfinite(ew[flags[1]]) ? 2 : 1;
    return GOPS_OK;
  }
hip_error:
#undef DTRY
  (void)hipStreamSynchronize(s);
  err = std::string("device graph build: ") + what + ": " + hipGetErrorString(herr);
  return (herr == hipErrorOutOfMemory || herr == hipErrorMemoryAllocation) ? GOPS_OUT_OF_MEMORY
                                                                           : GOPS_DEVICE_ERROR;
}

hipError_t upload_plan_to(hipStream_t stream, DevGraph &dg, const Plan &p) {
  auto up = [&](auto **dst, const auto &vec) -> hipError_t {
    using T = typename std::decay_t<decltype(vec)>::value_type;
    *dst = nullptr;
    hipError_t e = hipMalloc((void **)dst, std::max<size_t>(vec.size(), 1) * sizeof(T));
    if (e != hipSuccess || vec.empty()) return e;
    return hipMemcpyAsync(*dst, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, stream);
  };
  hipError_t e;
  if ((e = up(&dg.desc, p.desc)) != hipSuccess) return e;
  dg.num_desc = (int)p.desc.size();
  if ((e = up(&dg.hub_row, p.hub_row)) != hipSuccess) return e;
  if ((e = up(&dg.hub_nchunks, p.hub_nchunks)) != hipSuccess) return e;
  if ((e = up(&dg.hub_first, p.hub_first)) != hipSuccess) return e;
  dg.num_hubs = p.num_hubs;
  dg.num_hub_chunks = 0;
  for (int32_t c : p.hub_nchunks) dg.num_hub_chunks += c;
  return hipStreamSynchronize(stream);
}

// The out-edge CSR: its rows are the nodes (new ids) ordered by out-degree,
// descending, ties by id (so that runs of rows have similar lengths, as in
// the in-edge CSR), fwd.order maps a row to its node; each row lists the
// node's targets (new ids) ascending.  The edges (source = col[k], target =
// row of k) of the in-edge CSR are radix-sorted by (out-row, target).
gops_status build_forward(gops_context *ctx, gops_graph *g, std::string &err) {
  if (g->has_fwd) return GOPS_OK;
  const DevGraph &dg = g->dev;
  const uint64_t n = dg.n, m = dg.m;
  hipStream_t s = ctx->stream;
  Pool pool;
  hipError_t herr = hipSuccess;
  const char *what = "";
#define DTRY(expr)            \
  do {                        \
    herr = (expr);            \
    if (herr != hipSuccess) { \
      what = #expr;           \
      goto hip_error;         \
    }                         \
  } while (0)
  {
    const int cb = bits_for(n ? n - 1 : 0);
    uint32_t *d_marks = pool.get<uint32_t>(m), *d_row = pool.get<uint32_t>(m);
    uint64_t *d_key = pool.get<uint64_t>(m), *d_keyb = pool.get<uint64_t>(m);
    uint32_t *d_deg = pool.get<uint32_t>(n), *d_k32 = pool.get<uint32_t>(n);
    uint32_t *d_k32b = pool.get<uint32_t>(n), *d_v32 = pool.get<uint32_t>(n);
    uint32_t *d_forder = pool.get<uint32_t>(n), *d_frank = pool.get<uint32_t>(n);
    int32_t *d_cnt = pool.get<int32_t>(n + 1);
    int32_t *d_row_ptr = pool.get<int32_t>(n + 1), *d_col = pool.get<int32_t>(m);
    DTRY(pool.err);
    size_t tmp_bytes = 1, b = 0;
    DTRY(hipcub::DeviceScan::InclusiveScan(nullptr, b, d_marks, d_row, hipcub::Max(), (int)m, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceRadixSort::SortKeys(nullptr, b, d_key, d_keyb, (int)m, 0, 2 * cb, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, d_k32, d_k32b, d_v32, d_forder, (int)n, 0,
                                            32, s));
    tmp_bytes = std::max(tmp_bytes, b);
    DTRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, d_cnt, d_row_ptr, (int)n + 1, s));
    tmp_bytes = std::max(tmp_bytes, b);
    void *d_tmp = pool.get<uint8_t>(tmp_bytes);
    DTRY(pool.err);
    // rows: nodes by descending out-degree
    outdeg_keys_kernel<<<grid_for(n), kB, 0, s>>>(dg.inv_deg64, n, d_deg, d_k32, d_v32);
    DTRY(hipGetLastError());
    if (n) {
      b = tmp_bytes;
      DTRY(hipcub::DeviceRadixSort::SortPairs(d_tmp, b, d_k32, d_k32b, d_v32, d_forder, (int)n, 0,
                                              32, s));
    }
    DTRY(hipMemsetAsync(d_cnt + n, 0, sizeof(int32_t), s));
    rank_kernel<<<grid_for(n), kB, 0, s>>>(d_forder, d_deg, n, d_frank, d_cnt);
    DTRY(hipGetLastError());
    b = tmp_bytes;
    DTRY(hipcub::DeviceScan::ExclusiveSum(d_tmp, b, d_cnt, d_row_ptr, (int)n + 1, s));
    // edges by (out-row, target)
    if (m) {
      DTRY(hipMemsetAsync(d_marks, 0, m * sizeof(uint32_t), s));
      mark_rows_kernel<<<grid_for(n), kB, 0, s>>>(dg.row_ptr, n, d_marks);
      b = tmp_bytes;
      DTRY(hipcub::DeviceScan::InclusiveScan(d_tmp, b, d_marks, d_row, hipcub::Max(), (int)m, s));
      fwd_keys_kernel<<<grid_for(m), kB, 0, s>>>(dg.col, d_row, d_frank, m, cb, d_key);
      b = tmp_bytes;
      DTRY(hipcub::DeviceRadixSort::SortKeys(d_tmp, b, d_key, d_keyb, (int)m, 0, 2 * cb, s));
      col_kernel<<<grid_for(m), kB, 0, s>>>(d_keyb, m, (1ull << cb) - 1, d_col);
      DTRY(hipGetLastError());
    }
    HostCSRT h;
    h.n = n;
    h.m = m;
    h.row_ptr.resize(n + 1);
    DTRY(staged_download(ctx, h.row_ptr.data(), d_row_ptr, (n + 1) * sizeof(int32_t), s));
    Plan plan;
    plan_blocks(h, plan);
    DevGraph f;
    f.n = n;
    f.m = m;
    f.row_ptr = d_row_ptr;
    f.col = d_col;
    f.order = d_forder;
    pool.release(d_row_ptr);
    pool.release(d_col);
    pool.release(d_forder);
    herr = upload_plan_to(s, f, plan);
    if (herr != hipSuccess) {
      for (void *p : {(void *)f.row_ptr, (void *)f.col, (void *)f.order, (void *)f.desc,
                      (void *)f.hub_row, (void *)f.hub_nchunks, (void *)f.hub_first})
        (void)hipFree(p);
      what = "upload of the out-edge plan";
      goto hip_error;
    }
    g->fwd = f;
    g->has_fwd = true;
    return GOPS_OK;
  }
hip_error:
#undef DTRY
  (void)hipStreamSynchronize(s);
  err = std::string("out-edge CSR build: ") + what + ": " + hipGetErrorString(herr);
  return (herr == hipErrorOutOfMemory || herr == hipErrorMemoryAllocation) ? GOPS_OUT_OF_MEMORY
                                                                           : GOPS_DEVICE_ERROR;
}

}  // namespace gops
