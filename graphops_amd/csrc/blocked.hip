// blocked.hip -- the layout of the blocked f32 delta sweep, built once per
// graph on the device from the CSR of in-edges.
//
// A pull sweep gathers one source value per edge from a vector that no L2
// holds (67 MB of f32 deltas at R-MAT scale 24): the L1's miss handling, not
// HBM bandwidth, sets its rate.  The blocked sweep (kernels in pagerank.hip)
// replaces the gathers by two passes that stream:
//   scatter     a workgroup holds the deltas of one chunk of sources in LDS
//               and walks that chunk's edges (chunk-major order), writing each
//               product to the entry's place in the bin-major order;
//   accumulate  a workgroup holds fixed-point sums of one bin of rows in LDS
//               and walks the bin's products in order.
// Both orders keep the CSR's order within a row, and the sums are integers,
// so every result is independent of scheduling.
//
// Entry order:
//   chunk-major: (source >> kPbChunkShift, row, source)
//   bin-major:   (row >> kPbBinShift, source >> kPbChunkShift, row, source)
// Each is a stable sort of the CSR edge index by a key of at most 49 bits.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "internal.h"

namespace gops {

namespace {

constexpr int kB = 256;

int grid_for(uint64_t count) {
  const uint64_t g = (count + kB - 1) / kB;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 65535 * 8));
}

int bits_of(uint64_t x) {  // bits that hold every value in [0, x]
  int b = 1;
  while (b < 63 && (1ull << b) <= x) b++;
  return b;
}

// marks[row_ptr[v]] = v for rows with edges (marks zeroed): an inclusive max
// scan then gives every edge its row
__global__ void pb_mark_kernel(const int32_t *row_ptr, int n, int32_t *marks) {
  for (int v = blockIdx.x * kB + threadIdx.x; v < n; v += gridDim.x * kB)
    if (row_ptr[v] < row_ptr[v + 1]) marks[row_ptr[v]] = v;
}

// key of edge k in one of the two orders; value = k
__global__ void pb_keys_kernel(const int32_t *col, const int32_t *row, uint32_t m, int bin_major,
                               int sbits, uint64_t *keys, uint32_t *vals) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < m; k += gridDim.x * kB) {
    const uint64_t u = (uint32_t)col[k], v = (uint32_t)row[k];
    const uint64_t s = u >> kPbChunkShift;
    uint64_t key;
    if (bin_major) {
      key = ((v >> kPbBinShift) << (sbits + kPbBinShift)) | (s << kPbBinShift) |
            (v & ((1u << kPbBinShift) - 1));
    } else {
      key = (s << 31) | v;
    }
    keys[k] = key;
    vals[k] = k;
  }
}

// bin-major: position_of[edge] = its index; dst = row - bin base
__global__ void pb_binmajor_kernel(const uint64_t *keys, const uint32_t *vals, uint32_t m,
                                   uint32_t *position_of, uint16_t *dst) {
  for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < m; i += gridDim.x * kB) {
    position_of[vals[i]] = i;
    dst[i] = (uint16_t)(keys[i] & ((1u << kPbBinShift) - 1));
  }
}

// chunk-major: the entry's place in bin-major order, source - chunk base, weight
template <typename W>
__global__ void pb_chunkmajor_kernel(const uint32_t *vals, uint32_t m, const int32_t *col,
                                     const uint32_t *position_of, const W *w, uint32_t *pos,
                                     uint16_t *src, W *w_out) {
  for (uint32_t j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB) {
    const uint32_t k = vals[j];
    pos[j] = position_of[k];
    src[j] = (uint16_t)((uint32_t)col[k] & ((1u << kPbChunkShift) - 1));
    if (w) w_out[j] = w[k];
  }
}

// Segments of a scatter item: maximal runs of entries whose bin-major place
// is the chunk-major index plus one constant (a run of one chunk into one
// bin is one).  Thread t of an item's workgroup takes kPer consecutive entries.
constexpr int kPer = kPbItem1 / kB;

__device__ __forceinline__ bool seg_head(const PbItem &it, const uint32_t *pos, int j) {
  return j == it.y || (int64_t)pos[j] - j != (int64_t)pos[j - 1] - (j - 1);
}

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int *total) {
  sh[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < kB; i++) {
      const int x = sh[i];
      sh[i] = acc;
      acc += x;
    }
    sh[kB] = acc;
  }
  __syncthreads();
  const int r = sh[threadIdx.x];
  *total = sh[kB];
  __syncthreads();
  return r;
}

__global__ void pb_seg_count_kernel(const PbItem *items, const uint32_t *pos, int *count) {
  __shared__ int sh[kB + 1];
  const PbItem it = items[blockIdx.x];
  const int j0 = it.y + (int)threadIdx.x * kPer;
  int c = 0;
  for (int q = 0; q < kPer; q++) {
    const int j = j0 + q;
    if (j < it.z) c += seg_head(it, pos, j);
  }
  int total;
  (void)block_excl_scan(c, sh, &total);
  if (threadIdx.x == 0) count[blockIdx.x] = total;
}

__global__ void pb_seg_write_kernel(const PbItem *items, const uint32_t *pos, uint16_t *seg,
                                    int32_t *seg_off) {
  __shared__ int sh[kB + 1];
  const PbItem it = items[blockIdx.x];
  const int j0 = it.y + (int)threadIdx.x * kPer;
  int c = 0;
  for (int q = 0; q < kPer; q++) {
    const int j = j0 + q;
    if (j < it.z) c += seg_head(it, pos, j);
  }
  int total;
  int r = block_excl_scan(c, sh, &total) - 1;  // the segment before this thread's first entry
  for (int q = 0; q < kPer; q++) {
    const int j = j0 + q;
    if (j >= it.z) break;
    if (seg_head(it, pos, j)) {
      r++;
      seg_off[it.w + r] = (int32_t)((int64_t)pos[j] - j);
    }
    seg[j] = (uint16_t)r;
  }
}

// out[j] = first index i with keys[i] >= j << shift, for j in [0, count]
__global__ void pb_bounds_kernel(const uint64_t *keys, uint32_t m, int count, int shift,
                                 uint32_t *out) {
  for (int j = blockIdx.x * kB + threadIdx.x; j <= count; j += gridDim.x * kB) {
    const uint64_t t = (uint64_t)j << shift;
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (keys[mid] < t) lo = mid + 1;
      else hi = mid;
    }
    out[j] = lo;
  }
}

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
  void keep(void *p) { ptrs.erase(std::remove(ptrs.begin(), ptrs.end(), p), ptrs.end()); }
  ~Pool() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

size_t weight_bytes(int storage) {
  return storage == GOPS_WEIGHT_F64 ? 8 : storage == GOPS_WEIGHT_F32 ? 4
                                      : storage == GOPS_WEIGHT_BF16 ? 2 : 0;
}

hipError_t launch_chunkmajor(int storage, const uint32_t *vals, uint32_t m, const int32_t *col,
                             const uint32_t *position_of, const void *w, uint32_t *pos,
                             uint16_t *src, void *w_out, hipStream_t s) {
  const int g = grid_for(m);
  switch (storage) {
    case GOPS_WEIGHT_F64:
      pb_chunkmajor_kernel<double><<<g, kB, 0, s>>>(vals, m, col, position_of,
                                                    static_cast<const double *>(w), pos, src,
                                                    static_cast<double *>(w_out));
      break;
    case GOPS_WEIGHT_F32:
      pb_chunkmajor_kernel<float><<<g, kB, 0, s>>>(vals, m, col, position_of,
                                                   static_cast<const float *>(w), pos, src,
                                                   static_cast<float *>(w_out));
      break;
    default:
      pb_chunkmajor_kernel<uint16_t><<<g, kB, 0, s>>>(vals, m, col, position_of,
                                                      static_cast<const uint16_t *>(w), pos, src,
                                                      static_cast<uint16_t *>(w_out));
      break;
  }
  return hipGetLastError();
}

// ranges [off[j], off[j + 1]) split into items of at most `most` entries;
// with splits: a split range (a bin) gets one combine entry per kBlock rows,
// {bin, first slot, slots, first row}
void split_items(const std::vector<uint32_t> &off, uint32_t most, std::vector<PbItem> &items,
                 std::vector<PbItem> *splits, int *slots) {
  for (size_t j = 0; j + 1 < off.size(); j++) {
    const uint32_t b = off[j], e = off[j + 1];
    const uint32_t pieces = std::max<uint32_t>(1, (e - b + most - 1) / most);
    const bool split = splits && pieces > 1;
    if (split)
      for (int r = 0; r < (1 << kPbBinShift); r += kBlock)
        splits->push_back({(int32_t)j, *slots, (int32_t)pieces, r});
    for (uint32_t p = 0; p < pieces; p++) {
      const uint32_t lo = b + (uint32_t)(((uint64_t)(e - b) * p) / pieces);
      const uint32_t hi = b + (uint32_t)(((uint64_t)(e - b) * (p + 1)) / pieces);
      items.push_back({(int32_t)j, (int32_t)lo, (int32_t)hi, split ? (*slots)++ : -1});
    }
  }
}

}  // namespace

void free_blocked(BlockedLayout &pb) {
  (void)hipFree(pb.src);
  (void)hipFree(pb.w);
  (void)hipFree(pb.seg);
  (void)hipFree(pb.seg_off);
  (void)hipFree(pb.dst);
  (void)hipFree(pb.contrib);
  (void)hipFree(pb.items1);
  (void)hipFree(pb.items2);
  (void)hipFree(pb.splits);
  (void)hipFree(pb.bin_partial);
  (void)hipFree(pb.partials);
  pb = BlockedLayout{};
}

gops_status build_blocked(gops_context *ctx, gops_graph *g, std::string &err) {
  if (g->has_pb) return GOPS_OK;
  const DevGraph &dg = g->dev;
  const uint64_t n = dg.n, m = dg.m;
  const hipStream_t s = ctx->stream;
  const int chunks = (int)std::max<uint64_t>(1, (n + (1u << kPbChunkShift) - 1) >> kPbChunkShift);
  const int bins = (int)std::max<uint64_t>(1, (n + (1u << kPbBinShift) - 1) >> kPbBinShift);
  const int sbits = bits_of((uint64_t)chunks);
  const size_t wb = weight_bytes(dg.w ? dg.weight_storage : GOPS_WEIGHT_NONE);
  Pool pool;
  BlockedLayout pb;
  std::vector<uint32_t> chunk_off(chunks + 1, 0), bin_off(bins + 1, 0);
  hipError_t e = hipSuccess;
  const char *what = "";
#define BTRY(expr)        \
  do {                    \
    e = (expr);           \
    if (e != hipSuccess) { \
      what = #expr;       \
      goto fail;          \
    }                     \
  } while (0)
  {
    pb.src = pool.get<uint16_t>(m);
    pb.seg = pool.get<uint16_t>(m);
    uint32_t *pos = pool.get<uint32_t>(m);  // bin-major place of each chunk-major entry
    pb.dst = pool.get<uint16_t>(m);
    pb.contrib = pool.get<float>(m);
    if (wb) pb.w = pool.get<uint8_t>(m * wb);
    int32_t *row = pool.get<int32_t>(m);
    uint64_t *keys = pool.get<uint64_t>(m), *keys_b = pool.get<uint64_t>(m);
    uint32_t *vals = pool.get<uint32_t>(m), *vals_b = pool.get<uint32_t>(m);
    uint32_t *position_of = pool.get<uint32_t>(m);
    uint32_t *d_chunk_off = pool.get<uint32_t>(chunks + 1);
    uint32_t *d_bin_off = pool.get<uint32_t>(bins + 1);
    BTRY(pool.err);
    if (m) {
      // rows of the edges
      BTRY(hipMemsetAsync(vals, 0, m * sizeof(int32_t), s));
      pb_mark_kernel<<<grid_for(n), kB, 0, s>>>(dg.row_ptr, (int)n, (int32_t *)vals);
      BTRY(hipGetLastError());
      size_t tmp_bytes = 0, b = 0;
      BTRY(hipcub::DeviceScan::InclusiveScan(nullptr, b, (int32_t *)vals, row, hipcub::Max(),
                                             (int)m, s));
      tmp_bytes = std::max(tmp_bytes, b);
      const int bits_bin = bits_of((uint64_t)bins) + sbits + kPbBinShift;
      const int bits_chunk = sbits + 31;
      BTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, keys, keys_b, vals, vals_b, (int)m, 0,
                                              std::max(bits_bin, bits_chunk), s));
      tmp_bytes = std::max(tmp_bytes, b);
      void *tmp = pool.get<uint8_t>(tmp_bytes);
      BTRY(pool.err);
      b = tmp_bytes;
      BTRY(hipcub::DeviceScan::InclusiveScan(tmp, b, (int32_t *)vals, row, hipcub::Max(), (int)m,
                                             s));
      // bin-major order
      pb_keys_kernel<<<grid_for(m), kB, 0, s>>>(dg.col, row, (uint32_t)m, 1, sbits, keys, vals);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys_b, vals, vals_b, (int)m, 0,
                                              bits_bin, s));
      pb_binmajor_kernel<<<grid_for(m), kB, 0, s>>>(keys_b, vals_b, (uint32_t)m, position_of,
                                                    pb.dst);
      BTRY(hipGetLastError());
      pb_bounds_kernel<<<grid_for(bins + 1), kB, 0, s>>>(keys_b, (uint32_t)m, bins,
                                                         sbits + kPbBinShift, d_bin_off);
      BTRY(hipGetLastError());
      // chunk-major order
      pb_keys_kernel<<<grid_for(m), kB, 0, s>>>(dg.col, row, (uint32_t)m, 0, sbits, keys, vals);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys_b, vals, vals_b, (int)m, 0,
                                              bits_chunk, s));
      BTRY(launch_chunkmajor(dg.weight_storage, vals_b, (uint32_t)m, dg.col, position_of,
                             wb ? dg.w : nullptr, pos, pb.src, pb.w, s));
      pb_bounds_kernel<<<grid_for(chunks + 1), kB, 0, s>>>(keys_b, (uint32_t)m, chunks, 31,
                                                           d_chunk_off);
      BTRY(hipGetLastError());
      BTRY(hipMemcpyAsync(chunk_off.data(), d_chunk_off, (chunks + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
      BTRY(hipMemcpyAsync(bin_off.data(), d_bin_off, (bins + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    }
    BTRY(hipStreamSynchronize(s));
    // work items: every bin gets one, also without entries (its rows still
    // take the epilogue); chunks without entries get none
    std::vector<PbItem> items1, items2, splits;
    int slots = 0;
    {
      std::vector<uint32_t> off;
      for (int c = 0; c < chunks; c++) {
        if (chunk_off[c + 1] == chunk_off[c]) continue;
        off.assign({chunk_off[c], chunk_off[c + 1]});
        const size_t before = items1.size();
        split_items(off, kPbItem1, items1, nullptr, nullptr);
        for (size_t i = before; i < items1.size(); i++) items1[i].x = c;
      }
    }
    split_items(bin_off, kPbItem2, items2, &splits, &slots);
    pb.num_items1 = (int)items1.size();
    pb.num_items2 = (int)items2.size();
    pb.num_splits = (int)splits.size();
    pb.items1 = pool.get<PbItem>(items1.size());
    // the scatter items' segments: counted per item, then written from each
    // item's first segment on
    if (!items1.empty()) {
      int *d_count = pool.get<int>(items1.size());
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(pb.items1, items1.data(), items1.size() * sizeof(PbItem),
                          hipMemcpyHostToDevice, s));
      pb_seg_count_kernel<<<(int)items1.size(), kB, 0, s>>>(pb.items1, pos, d_count);
      BTRY(hipGetLastError());
      std::vector<int> count(items1.size());
      BTRY(hipMemcpyAsync(count.data(), d_count, count.size() * sizeof(int),
                          hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      int64_t total = 0;
      for (size_t i = 0; i < items1.size(); i++) {
        items1[i].w = (int32_t)total;
        total += count[i];
      }
      pb.seg_off = pool.get<int32_t>((size_t)total);
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(pb.items1, items1.data(), items1.size() * sizeof(PbItem),
                          hipMemcpyHostToDevice, s));
      pb_seg_write_kernel<<<(int)items1.size(), kB, 0, s>>>(pb.items1, pos, pb.seg, pb.seg_off);
      BTRY(hipGetLastError());
    }
    pb.items2 = pool.get<PbItem>(items2.size());
    pb.splits = pool.get<PbItem>(splits.size());
    pb.bin_partial = pool.get<long long>((size_t)std::max(slots, 1) << kPbBinShift);
    pb.partials = pool.get<double>((items2.size() + splits.size()) * kMaxReduce);
    BTRY(pool.err);
    BTRY(hipMemcpyAsync(pb.items2, items2.data(), items2.size() * sizeof(PbItem),
                        hipMemcpyHostToDevice, s));
    if (!splits.empty())
      BTRY(hipMemcpyAsync(pb.splits, splits.data(), splits.size() * sizeof(PbItem),
                          hipMemcpyHostToDevice, s));
    BTRY(hipStreamSynchronize(s));
  }
#undef BTRY
  for (void *p : {(void *)pb.src, (void *)pb.seg, (void *)pb.seg_off, (void *)pb.dst,
                  (void *)pb.contrib, pb.w,
                  (void *)pb.items1, (void *)pb.items2, (void *)pb.splits,
                  (void *)pb.bin_partial, (void *)pb.partials})
    if (p) pool.keep(p);
  g->pb = pb;
  g->has_pb = true;
  return GOPS_OK;
fail:
  err = std::string("blocked layout: ") + what + ": " + hipGetErrorString(e);
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GOPS_OUT_OF_MEMORY
                                                                     : GOPS_DEVICE_ERROR;
}

}  // namespace gops
