// blocked.hip -- the layout of the blocked f32 delta sweep, built once per
// graph on the device from the CSR of in-edges.
//
// A pull sweep gathers one source value per edge from a vector that no L2
// holds (67 MB of f32 deltas at R-MAT scale 24): the L1's miss handling, not
// HBM bandwidth, sets its rate.  The blocked sweep (kernels in pagerank.hip)
// replaces the gathers by two passes that stream:
//   scatter     a workgroup holds the deltas of one chunk of sources in LDS
//               and walks that chunk's edges in (row, source) order, eight
//               consecutive entries per lane; the products of each run of one
//               row are summed and one f32 per run goes to the run's place in
//               bin-major order;
//   accumulate  a workgroup holds fixed-point sums of one bin of rows in LDS
//               and walks the bin's run sums in order.
// The summation order within a run is fixed by the layout and the row sums
// are integers, so every result is independent of scheduling.
//
// Orders:
//   entries, chunk-major:  (source >> kPbChunkShift, row, source), padded so
//                          that each segment (one chunk into one bin) starts at
//                          a multiple of kPbLane entries and each chunk at a
//                          multiple of kPbWin (a window, one wave's entries)
//   runs, bin-major:       (bin of the row, source chunk, row, run number)
// Bins are ranges of consecutive rows, at most 2^kPbBinShift of them, cut so
// that the runs of a bin stay within one accumulate workgroup's share
// (pb_group_marks_kernel bounds every row's runs before the layout exists).
// A run is a maximal sequence of entries of one row within one window.  A
// segment's runs are contiguous in both orders, so a run's bin-major place is
// its chunk-major number plus its segment's offset.  The scatter finds an
// entry's run and segment from the head bits in the source words and the
// numbers of its window's first run and segment: no per-entry index is stored.
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
// count[0] += rows without in-edges and out-edges; count[1] = max over the
// other rows of (row + 1): the isolated rows are the tail [count[1], n) when
// count[0] = n - count[1]
__global__ void pb_iso_kernel(const int32_t *row_ptr, const float *inv_deg, int n,
                              unsigned long long *count) {
  unsigned long long iso = 0, last = 0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    if (row_ptr[r + 1] == row_ptr[r] && inv_deg[r] == 0.0f) iso++;
    else last = (unsigned long long)r + 1 > last ? (unsigned long long)r + 1 : last;
  }
  if (iso) atomicAdd(&count[0], iso);
  if (last) atomicMax(&count[1], last);
}

__global__ void pb_mark_kernel(const int32_t *row_ptr, int n, int32_t *marks) {
  for (int v = blockIdx.x * kB + threadIdx.x; v < n; v += gridDim.x * kB)
    if (row_ptr[v] < row_ptr[v + 1]) marks[row_ptr[v]] = v;
}

// Per edge k of the CSR of in-edges (row[k] its row, sources ascending): 1
// if k starts a group (the row's edges from one source chunk), 1 if it is a
// group's second edge, else 0.  A group of c entries is contiguous in the
// padded layout and meets at most 1 + floor((c + 510) / kPbWin) windows: 1
// for c = 1, at most 2 + (c - 2) / kPbWin for c >= 2.  So a row's marks plus
// ceil(edges / kPbWin) bound its runs, whatever the padding.
__global__ void pb_group_marks_kernel(const int32_t *row_ptr, const int32_t *col,
                                      const int32_t *row, uint32_t m, int32_t *mark) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < m; k += gridDim.x * kB) {
    const int32_t s = row_ptr[row[k]];
    const uint32_t c = (uint32_t)col[k] >> kPbChunkShift;
    const bool head = (int32_t)k == s || ((uint32_t)col[k - 1] >> kPbChunkShift) != c;
    const bool second =
        !head && ((int32_t)k - 1 == s || ((uint32_t)col[k - 2] >> kPbChunkShift) != c);
    mark[k] = (head || second) ? 1 : 0;
  }
}

// bound[r] = the runs row r can have (above), from the inclusive sum of marks
__global__ void pb_row_bound_kernel(const int32_t *row_ptr, const int32_t *msum, int n,
                                    uint32_t *bound) {
  for (int r = blockIdx.x * kB + threadIdx.x; r < n; r += gridDim.x * kB) {
    const int32_t s = row_ptr[r], e = row_ptr[r + 1];
    bound[r] = e > s ? (uint32_t)(msum[e - 1] - (s > 0 ? msum[s - 1] : 0)) +
                           (uint32_t)((e - s + kPbWin - 1) / kPbWin)
                     : 0u;
  }
}

// bin_of[r]: the bin whose rows [row0[b], row0[b + 1]) hold r
__global__ void pb_bin_of_kernel(const int32_t *row0, int bins, int n, int32_t *bin_of) {
  for (int r = blockIdx.x * kB + threadIdx.x; r < n; r += gridDim.x * kB) {
    int lo = 0, hi = bins - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (row0[mid] <= r) lo = mid;
      else hi = mid - 1;
    }
    bin_of[r] = lo;
  }
}

// chunk-major key of edge k (chunk of the source, row); value = k.  A stable
// sort keeps the CSR's ascending sources within a row.
__global__ void pb_keys_kernel(const int32_t *col, const int32_t *row, uint32_t m, uint64_t *keys,
                               uint32_t *vals) {
  for (uint32_t k = blockIdx.x * kB + threadIdx.x; k < m; k += gridDim.x * kB) {
    const uint64_t u = (uint32_t)col[k], v = (uint32_t)row[k];
    keys[k] = ((u >> kPbChunkShift) << 31) | v;
    vals[k] = k;
  }
}

__device__ __forceinline__ uint32_t key_chunk(uint64_t key) { return (uint32_t)(key >> 31); }
__device__ __forceinline__ uint32_t key_row(uint64_t key) { return (uint32_t)(key & 0x7fffffffu); }

// segment starts among the sorted entries: a new chunk or a new bin of rows
__global__ void pb_seg_flag_kernel(const uint64_t *keys, const int32_t *bin_of, uint32_t m,
                                   int32_t *flag) {
  for (uint32_t j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB) {
    const uint64_t k = keys[j];
    bool f = j == 0;
    if (!f) {
      const uint64_t q = keys[j - 1];
      f = key_chunk(k) != key_chunk(q) || bin_of[key_row(k)] != bin_of[key_row(q)];
    }
    flag[j] = f;
  }
}

// the first entry of every segment (sg: inclusive count of segment starts)
__global__ void pb_seg_first_kernel(const int32_t *flag, const int32_t *sg, uint32_t m,
                                    int32_t *seg_first) {
  for (uint32_t j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB)
    if (flag[j]) seg_first[sg[j] - 1] = (int32_t)j;
}

// padded length of every segment (a multiple of kPbLane); 0 at [nseg]
__global__ void pb_seg_len_kernel(const int32_t *seg_first, uint32_t nseg, int32_t *plen) {
  for (uint32_t q = blockIdx.x * kB + threadIdx.x; q <= nseg; q += gridDim.x * kB)
    plen[q] = q < nseg ? ((seg_first[q + 1] - seg_first[q] + kPbLane - 1) / kPbLane) * kPbLane : 0;
}

// out[i] = in[idx[i]]
__global__ void pb_gather_kernel(const int32_t *in, const int32_t *idx, int count, int32_t *out) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < count; i += gridDim.x * kB) out[i] = in[idx[i]];
}

// Every real entry to its padded place p: source word (run and segment head
// bits), weight; posof[j] = p.  shift[c] = padded start of chunk c minus the
// padded start (G) of its first segment.
template <typename W>
__global__ void pb_place_entries_kernel(const uint64_t *keys, const uint32_t *vals, const int32_t *flag,
                                        const int32_t *sg, const int32_t *seg_first, const int32_t *G,
                                        const int32_t *shift, const int32_t *col, const W *w,
                                        uint32_t m, uint16_t *src, W *w_out, uint8_t *one_out,
                                        int32_t *posof) {
  for (uint32_t j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB) {
    const uint64_t key = keys[j];
    const int sgi = sg[j] - 1;
    const int32_t p = shift[key_chunk(key)] + G[sgi] + ((int32_t)j - seg_first[sgi]);
    const bool seg_head = flag[j] != 0;
    const bool head = seg_head || (p % kPbWin) == 0 || key_row(keys[j - 1]) != key_row(key);
    const uint32_t k = vals[j];
    src[p] = (uint16_t)(((uint32_t)col[k] & kPbSrcMask) | (head ? kPbRunHead : 0u) |
                        (seg_head ? kPbSegHead : 0u));
    if (w) w_out[p] = w[k];
    if (one_out) one_out[p] = 1;
    posof[j] = p;
  }
}

// Padding after every segment's entries: up to the next segment's padded
// start, or for a chunk's last segment up to the chunk's padded end; each
// padding entry repeats the source of the segment's last entry (no head
// bits, so it continues that run) with weight 0.
template <typename W>
__global__ void pb_pad_kernel(const uint64_t *keys, const int32_t *seg_first, const int32_t *G,
                              const int32_t *shift, const int32_t *chunk_end, uint32_t nseg,
                              uint16_t *src, W *w_out, uint8_t *one_out) {
  for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < nseg; q += gridDim.x * kB) {
    const uint32_t c = key_chunk(keys[seg_first[q]]);
    const int32_t e = shift[c] + G[q] + (seg_first[q + 1] - seg_first[q]);
    const bool last = q + 1 == nseg || key_chunk(keys[seg_first[q + 1]]) != c;
    const int32_t pend = last ? chunk_end[c] : shift[c] + G[q + 1];
    const uint16_t s = (uint16_t)(src[e - 1] & kPbSrcMask);
    for (int32_t p = e; p < pend; p++) {
      src[p] = s;
      if (w_out) w_out[p] = W(0);
      if (one_out) one_out[p] = 0;
    }
  }
}

// run-head flags of the padded entries, for the scan of run numbers
__global__ void pb_is_head_kernel(const uint16_t *src, uint32_t mp, int32_t *flag) {
  for (uint32_t p = blockIdx.x * kB + threadIdx.x; p < mp; p += gridDim.x * kB)
    flag[p] = (src[p] & kPbRunHead) != 0;
}

// Per run (at its head, a real entry; runid = inclusive count of heads): the
// bin-major key (bin, chunk, row in bin), value = run number; every
// segment's first run; every window's first run and segment.
__global__ void pb_run_keys_kernel(const uint64_t *keys, const int32_t *posof, const uint16_t *src,
                                   const int32_t *runid, const int32_t *flag, const int32_t *sg,
                                   const int32_t *bin_of, const int32_t *bin_row0, uint32_t m,
                                   int sbits, uint64_t *rkey, uint32_t *rval, int32_t *seg_run,
                                   int32_t *win) {
  for (uint32_t j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB) {
    const int32_t p = posof[j];
    if (!(src[p] & kPbRunHead)) continue;
    const uint32_t r = (uint32_t)runid[p] - 1;
    const uint64_t key = keys[j];
    const uint64_t v = key_row(key), c = key_chunk(key);
    const uint64_t bin = (uint64_t)bin_of[v];
    rkey[r] = (bin << (sbits + kPbBinShift)) | (c << kPbBinShift) | (v - (uint64_t)bin_row0[bin]);
    rval[r] = r;
    if (flag[j]) seg_run[sg[j] - 1] = (int32_t)r;
    if (p % kPbWin == 0) {
      win[2 * (p / kPbWin)] = (int32_t)r;
      win[2 * (p / kPbWin) + 1] = sg[j] - 1;
    }
  }
}

// bin-major place of every run, and its row within the bin
__global__ void pb_place_kernel(const uint64_t *rkey_sorted, const uint32_t *rval_sorted,
                                uint32_t nr, int32_t *place_of, uint16_t *rrow) {
  for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < nr; i += gridDim.x * kB) {
    place_of[rval_sorted[i]] = (int32_t)i;
    rrow[i] = (uint16_t)(rkey_sorted[i] & ((1u << kPbBinShift) - 1));
  }
}

// seg_off per segment: place of its first run - that run's number
__global__ void pb_seg_off_kernel(const int32_t *place_of, const int32_t *seg_run, uint32_t nseg,
                                  int32_t *seg_off) {
  for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < nseg; q += gridDim.x * kB)
    seg_off[q] = place_of[seg_run[q]] - seg_run[q];
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
  void release(void *p) {
    if (!p) return;
    keep(p);
    (void)hipFree(p);
  }
  ~Pool() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

size_t weight_bytes(int storage) {
  return storage == GOPS_WEIGHT_F64 ? 8 : storage == GOPS_WEIGHT_F32 ? 4
                                      : storage == GOPS_WEIGHT_BF16 ? 2 : 0;
}

template <typename W>
hipError_t place_and_pad(const DevGraph &dg, const void *w, uint32_t m, uint32_t nseg,
                         const uint64_t *keys, const uint32_t *vals, const int32_t *flag,
                         const int32_t *sg, const int32_t *seg_first, const int32_t *G,
                         const int32_t *shift, const int32_t *chunk_end, uint16_t *src,
                         void *w_out, uint8_t *one_out, int32_t *posof, hipStream_t s) {
  pb_place_entries_kernel<W><<<grid_for(m), kB, 0, s>>>(
      keys, vals, flag, sg, seg_first, G, shift, dg.col, static_cast<const W *>(w), m, src,
      static_cast<W *>(w_out), one_out, posof);
  pb_pad_kernel<W><<<grid_for(nseg), kB, 0, s>>>(keys, seg_first, G, shift, chunk_end, nseg, src,
                                                  static_cast<W *>(w_out), one_out);
  return hipGetLastError();
}

hipError_t launch_place_and_pad(int storage, const DevGraph &dg, uint32_t m, uint32_t nseg,
                                const uint64_t *keys, const uint32_t *vals, const int32_t *flag,
                                const int32_t *sg, const int32_t *seg_first, const int32_t *G,
                                const int32_t *shift, const int32_t *chunk_end, uint16_t *src,
                                void *w_out, uint8_t *mask, int32_t *posof, hipStream_t s) {
  switch (storage) {
    case GOPS_WEIGHT_F64:
      return place_and_pad<double>(dg, dg.w, m, nseg, keys, vals, flag, sg, seg_first, G, shift,
                                   chunk_end, src, w_out, mask, posof, s);
    case GOPS_WEIGHT_F32:
      return place_and_pad<float>(dg, dg.w, m, nseg, keys, vals, flag, sg, seg_first, G, shift,
                                  chunk_end, src, w_out, mask, posof, s);
    case GOPS_WEIGHT_BF16:
      return place_and_pad<uint16_t>(dg, dg.w, m, nseg, keys, vals, flag, sg, seg_first, G,
                                     shift, chunk_end, src, w_out, mask, posof, s);
    default:  // unweighted graph: the mask alone
      return place_and_pad<uint8_t>(dg, nullptr, m, nseg, keys, vals, flag, sg, seg_first, G,
                                    shift, chunk_end, src, nullptr, mask, posof, s);
  }
}

// ranges [off[j], off[j + 1]) split into items of at most `most` entries;
// with splits: a split range (bin j, of row0[j + 1] - row0[j] rows) gets one
// combine entry per kBlock rows, {bin, first slot, slots, first row}
void split_items(const std::vector<uint32_t> &off, uint32_t most, std::vector<PbItem> &items,
                 std::vector<PbItem> *splits, int *slots, const std::vector<int32_t> *row0) {
  for (size_t j = 0; j + 1 < off.size(); j++) {
    const uint32_t b = off[j], e = off[j + 1];
    const uint32_t pieces = std::max<uint32_t>(1, (e - b + most - 1) / most);
    const bool split = splits && pieces > 1;
    if (split)
      for (int r = 0; r < (*row0)[j + 1] - (*row0)[j]; r += kBlock)
        splits->push_back({(int32_t)j, *slots, (int32_t)pieces, r});
    for (uint32_t p = 0; p < pieces; p++) {
      const uint32_t lo = b + (uint32_t)(((uint64_t)(e - b) * p) / pieces);
      const uint32_t hi = b + (uint32_t)(((uint64_t)(e - b) * (p + 1)) / pieces);
      items.push_back({(int32_t)j, (int32_t)lo, (int32_t)hi, split ? (*slots)++ : -1});
    }
  }
}

// First row of every bin, then n: a bin ends at 2^kPbBinShift rows, or (cap >
// 0) before the row whose run bound would take it past cap runs.  A row whose
// bound alone passes cap gets a bin of its own, split over several
// accumulate items.
void cut_bins(const std::vector<uint32_t> &bound, uint64_t n, int cap, std::vector<int32_t> &row0) {
  row0.assign(1, 0);
  uint64_t acc = 0;
  int rows = 0;
  for (uint64_t r = 0; r < n; r++) {
    const uint64_t b = bound.empty() ? 0 : bound[r];
    if (rows == (1 << kPbBinShift) || (cap > 0 && rows > 0 && acc + b > (uint64_t)cap)) {
      row0.push_back((int32_t)r);
      acc = 0;
      rows = 0;
    }
    acc += b;
    rows++;
  }
  row0.push_back((int32_t)n);
}

}  // namespace

void free_blocked(BlockedLayout &pb) {
  for (void *p : {(void *)pb.src, pb.w, (void *)pb.mask, (void *)pb.win, (void *)pb.seg_off,
                  (void *)pb.rrow, (void *)pb.rsum, (void *)pb.items1, (void *)pb.items2,
                  (void *)pb.splits, (void *)pb.bin_partial, (void *)pb.bin_flags,
                  (void *)pb.partials, (void *)pb.bin_row0})
    (void)hipFree(p);
  pb = BlockedLayout{};
}

namespace {

// The blocked layout of dg's rows (the CSR of in-edges, or of out-edges for
// the power iterations that sum over a node's own neighbours) into out.
// with_iso: set the isolated tail apart (rows without in-edges and without
// out-edges, from the in-edge CSR's inverse degrees).
gops_status build_layout(gops_context *ctx, const DevGraph &dg, bool with_iso, BlockedLayout &out,
                         std::string &err) {
  const uint64_t n = dg.n, m = dg.m;
  const hipStream_t s = ctx->stream;
  const int chunks = (int)std::max<uint64_t>(1, (n + (1u << kPbChunkShift) - 1) >> kPbChunkShift);
  const int sbits = bits_of((uint64_t)chunks);
  const size_t wb = dg.w ? weight_bytes(dg.weight_storage) : 0;  // weight bytes per entry
  const int storage = dg.w ? dg.weight_storage : GOPS_WEIGHT_NONE;
  Pool pool;
  BlockedLayout pb;
  std::vector<uint32_t> chunk_off(chunks + 1, 0), bin_off;
  std::vector<int32_t> row0s;  // first row of every bin, then n
  int bins = 0;
  std::vector<PbItem> items1, items2, splits;
  int slots = 0;
  uint32_t nr = 0, nseg = 0;
  uint64_t mp = 0;  // padded entries
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
    if (m) {
      int32_t *row = pool.get<int32_t>(m);
      uint64_t *keys = pool.get<uint64_t>(m), *keys_b = pool.get<uint64_t>(m);
      uint32_t *vals = pool.get<uint32_t>(m), *vals_b = pool.get<uint32_t>(m);
      uint32_t *d_chunk_off = pool.get<uint32_t>(chunks + 1);
      BTRY(pool.err);
      // rows of the edges
      BTRY(hipMemsetAsync(vals, 0, m * sizeof(int32_t), s));
      pb_mark_kernel<<<grid_for(n), kB, 0, s>>>(dg.row_ptr, (int)n, (int32_t *)vals);
      BTRY(hipGetLastError());
      size_t tmp_bytes = 0, b = 0;
      BTRY(hipcub::DeviceScan::InclusiveScan(nullptr, b, (int32_t *)vals, row, hipcub::Max(),
                                             (int)m, s));
      tmp_bytes = std::max(tmp_bytes, b);
      // sized for the largest sort (entries; runs are fewer) and scans (the
      // padded entries are the most)
      BTRY(hipcub::DeviceRadixSort::SortPairs(nullptr, b, keys, keys_b, vals, vals_b, (int)m, 0,
                                              63, s));
      tmp_bytes = std::max(tmp_bytes, b);
      const uint64_t mp_max = m + (uint64_t)(kPbLane - 1) * m + (uint64_t)kPbWin * chunks;
      if (mp_max < (1ull << 31)) {
        BTRY(hipcub::DeviceScan::InclusiveSum(nullptr, b, (int32_t *)vals, (int32_t *)vals_b,
                                              (int)mp_max, s));
        tmp_bytes = std::max(tmp_bytes, b);
      }
      BTRY(hipcub::DeviceScan::InclusiveSum(nullptr, b, (int32_t *)vals, (int32_t *)vals_b,
                                            (int)std::min<uint64_t>(mp_max, (1ull << 31) - 1), s));
      tmp_bytes = std::max(tmp_bytes, b);
      BTRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int32_t *)vals, (int32_t *)vals_b,
                                            (int)m + 1, s));
      tmp_bytes = std::max(tmp_bytes, b);
      void *tmp = pool.get<uint8_t>(tmp_bytes);
      BTRY(pool.err);
      b = tmp_bytes;
      BTRY(hipcub::DeviceScan::InclusiveScan(tmp, b, (int32_t *)vals, row, hipcub::Max(), (int)m,
                                             s));
      // bins of consecutive rows, cut by the rows' run bounds (marks and their
      // sum in keys_b, free until the sort)
      int32_t *bin_of = pool.get<int32_t>(n);
      {
        int32_t *mark = reinterpret_cast<int32_t *>(keys_b), *msum = mark + m;
        uint32_t *bound = pool.get<uint32_t>(n);
        BTRY(pool.err);
        pb_group_marks_kernel<<<grid_for(m), kB, 0, s>>>(dg.row_ptr, dg.col, row, (uint32_t)m,
                                                         mark);
        BTRY(hipGetLastError());
        b = tmp_bytes;
        BTRY(hipcub::DeviceScan::InclusiveSum(tmp, b, mark, msum, (int)m, s));
        pb_row_bound_kernel<<<grid_for(n), kB, 0, s>>>(dg.row_ptr, msum, (int)n, bound);
        BTRY(hipGetLastError());
        std::vector<uint32_t> hb(n);
        BTRY(hipMemcpyAsync(hb.data(), bound, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        BTRY(hipStreamSynchronize(s));
        pool.release(bound);
        cut_bins(hb, n, ctx->pb_bin_runs, row0s);
      }
      bins = (int)row0s.size() - 1;
      pb.bin_row0 = pool.get<int32_t>(row0s.size());
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(pb.bin_row0, row0s.data(), row0s.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
      pb_bin_of_kernel<<<grid_for(n), kB, 0, s>>>(pb.bin_row0, bins, (int)n, bin_of);
      BTRY(hipGetLastError());
      // chunk-major order of the edges
      pb_keys_kernel<<<grid_for(m), kB, 0, s>>>(dg.col, row, (uint32_t)m, keys, vals);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys_b, vals, vals_b, (int)m, 0,
                                              sbits + 31, s));
      pool.release(row);
      pool.release(keys);
      keys = nullptr;
      pb_bounds_kernel<<<grid_for(chunks + 1), kB, 0, s>>>(keys_b, (uint32_t)m, chunks, 31,
                                                           d_chunk_off);
      BTRY(hipGetLastError());
      BTRY(hipMemcpyAsync(chunk_off.data(), d_chunk_off, (chunks + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
      // segments: starts, numbers (sg, from 1), first entries, padded starts G
      int32_t *flag = pool.get<int32_t>(m);
      int32_t *sg = reinterpret_cast<int32_t *>(vals);  // vals (unsorted) is free now
      BTRY(pool.err);
      pb_seg_flag_kernel<<<grid_for(m), kB, 0, s>>>(keys_b, bin_of, (uint32_t)m, flag);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceScan::InclusiveSum(tmp, b, flag, sg, (int)m, s));
      int32_t last = 0;
      BTRY(hipMemcpyAsync(&last, sg + (m - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      nseg = (uint32_t)last;
      pb.num_segs = nseg;
      int32_t *seg_first = pool.get<int32_t>((size_t)nseg + 1);
      int32_t *plen = pool.get<int32_t>((size_t)nseg + 1);
      int32_t *G = pool.get<int32_t>((size_t)nseg + 1);
      BTRY(pool.err);
      pb_seg_first_kernel<<<grid_for(m), kB, 0, s>>>(flag, sg, (uint32_t)m, seg_first);
      BTRY(hipGetLastError());
      const int32_t m32 = (int32_t)m;
      BTRY(hipMemcpyAsync(seg_first + nseg, &m32, sizeof(int32_t), hipMemcpyHostToDevice, s));
      pb_seg_len_kernel<<<grid_for((uint64_t)nseg + 1), kB, 0, s>>>(seg_first, nseg, plen);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceScan::ExclusiveSum(tmp, b, plen, G, (int)nseg + 1, s));
      // per chunk with edges: its first segment and G there, then its padded
      // start and end (chunks start at multiples of kPbWin)
      std::vector<int32_t> nonempty, first_entry;
      for (int c = 0; c < chunks; c++)
        if (chunk_off[c + 1] > chunk_off[c]) {
          nonempty.push_back(c);
          first_entry.push_back((int32_t)chunk_off[c]);
        }
      const int ne = (int)nonempty.size();
      int32_t *d_idx = pool.get<int32_t>((size_t)ne + 1);
      int32_t *d_out = pool.get<int32_t>((size_t)ne + 1);
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(d_idx, first_entry.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice,
                          s));
      pb_gather_kernel<<<grid_for(ne), kB, 0, s>>>(sg, d_idx, ne, d_out);
      BTRY(hipGetLastError());
      std::vector<int32_t> first_seg(ne + 1), g_at(ne + 1);
      BTRY(hipMemcpyAsync(first_seg.data(), d_out, ne * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      for (int i = 0; i < ne; i++) first_seg[i] -= 1;
      first_seg[ne] = (int32_t)nseg;
      BTRY(hipMemcpyAsync(d_idx, first_seg.data(), (ne + 1) * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
      pb_gather_kernel<<<grid_for(ne + 1), kB, 0, s>>>(G, d_idx, ne + 1, d_out);
      BTRY(hipGetLastError());
      BTRY(hipMemcpyAsync(g_at.data(), d_out, (ne + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      std::vector<int32_t> shift(chunks, 0), chunk_end(chunks, 0);
      uint64_t at = 0;
      for (int i = 0; i < ne; i++) {
        const uint64_t total = (uint64_t)(g_at[i + 1] - g_at[i]);
        const uint64_t padded = (total + kPbWin - 1) / kPbWin * kPbWin;
        shift[nonempty[i]] = (int32_t)((int64_t)at - g_at[i]);
        chunk_end[nonempty[i]] = (int32_t)(at + padded);
        // scatter items: the chunk's windows
        const uint32_t w0 = (uint32_t)(at / kPbWin), w1 = (uint32_t)((at + padded) / kPbWin);
        std::vector<uint32_t> off{w0, w1};
        const size_t before = items1.size();
        split_items(off, kPbItem1, items1, nullptr, nullptr, nullptr);
        for (size_t t = before; t < items1.size(); t++) items1[t] = {nonempty[i], items1[t].y, items1[t].z, 0};
        at += padded;
        if (at >= (1ull << 31)) {
          e = hipErrorOutOfMemory;
          what = "padded entries beyond 2^31";
          goto fail;
        }
      }
      mp = at;
      pb.m = mp;
      int32_t *d_shift = pool.get<int32_t>(chunks), *d_end = pool.get<int32_t>(chunks);
      pb.src = pool.get<uint16_t>(mp);
      if (wb) pb.w = pool.get<uint8_t>(mp * wb);
      pb.mask = pool.get<uint8_t>(mp);
      pb.win = pool.get<int32_t>(2 * (mp / kPbWin));
      int32_t *posof = pool.get<int32_t>(m);
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(d_shift, shift.data(), chunks * sizeof(int32_t), hipMemcpyHostToDevice, s));
      BTRY(hipMemcpyAsync(d_end, chunk_end.data(), chunks * sizeof(int32_t), hipMemcpyHostToDevice,
                          s));
      BTRY(launch_place_and_pad(storage, dg, (uint32_t)m, nseg, keys_b, vals_b, flag, sg, seg_first,
                                G, d_shift, d_end, pb.src, pb.w, pb.mask, posof, s));
      // run numbers over the padded entries
      int32_t *hflag = pool.get<int32_t>(mp), *runid = pool.get<int32_t>(mp);
      BTRY(pool.err);
      pb_is_head_kernel<<<grid_for(mp), kB, 0, s>>>(pb.src, (uint32_t)mp, hflag);
      BTRY(hipGetLastError());
      b = tmp_bytes;
      BTRY(hipcub::DeviceScan::InclusiveSum(tmp, b, hflag, runid, (int)mp, s));
      BTRY(hipMemcpyAsync(&last, runid + (mp - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      pool.release(hflag);
      nr = (uint32_t)last;
      pb.num_runs = nr;
      // runs: bin-major keys, sorted; places and rows; segment offsets
      uint64_t *rkey = pool.get<uint64_t>(nr), *rkey_b = pool.get<uint64_t>(nr);
      uint32_t *rval = pool.get<uint32_t>(nr), *rval_b = pool.get<uint32_t>(nr);
      int32_t *place_of = pool.get<int32_t>(nr);
      int32_t *seg_run = pool.get<int32_t>(nseg);
      uint32_t *d_bin_off = pool.get<uint32_t>(bins + 1);
      pb.rrow = pool.get<uint16_t>(nr);
      pb.rsum64 = pool.get<double>(nr);
      pb.rsum = reinterpret_cast<float *>(pb.rsum64);
      pb.seg_off = pool.get<int32_t>(nseg);
      BTRY(pool.err);
      pb_run_keys_kernel<<<grid_for(m), kB, 0, s>>>(keys_b, posof, pb.src, runid, flag, sg,
                                                    bin_of, pb.bin_row0, (uint32_t)m, sbits,
                                                    rkey, rval, seg_run, pb.win);
      BTRY(hipGetLastError());
      BTRY(hipStreamSynchronize(s));
      pool.release(bin_of);
      const int bits_bin = bits_of((uint64_t)bins) + sbits + kPbBinShift;
      b = tmp_bytes;
      BTRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, rkey, rkey_b, rval, rval_b, (int)nr, 0,
                                              bits_bin, s));
      pb_place_kernel<<<grid_for(nr), kB, 0, s>>>(rkey_b, rval_b, nr, place_of, pb.rrow);
      BTRY(hipGetLastError());
      pb_bounds_kernel<<<grid_for(bins + 1), kB, 0, s>>>(rkey_b, nr, bins, sbits + kPbBinShift,
                                                         d_bin_off);
      BTRY(hipGetLastError());
      pb_seg_off_kernel<<<grid_for(nseg), kB, 0, s>>>(place_of, seg_run, nseg, pb.seg_off);
      BTRY(hipGetLastError());
      bin_off.assign(bins + 1, 0);
      BTRY(hipMemcpyAsync(bin_off.data(), d_bin_off, (bins + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
    } else {  // no edges: bins of 2^kPbBinShift rows, no runs
      cut_bins({}, n, 0, row0s);
      bins = (int)row0s.size() - 1;
      bin_off.assign(bins + 1, 0);
      pb.bin_row0 = pool.get<int32_t>(row0s.size());
      BTRY(pool.err);
      BTRY(hipMemcpyAsync(pb.bin_row0, row0s.data(), row0s.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
    }
    pb.num_bins = bins;
    pb.num_items1 = (int)items1.size();
    pb.items1 = pool.get<PbItem>(items1.size());
    // accumulate items: every bin gets one, also without runs (its rows still
    // take the epilogue)
    split_items(bin_off, kPbItem2, items2, &splits, &slots, &row0s);
    for (size_t j = 0; j < splits.size(); j++)
      pb.num_split_bins += j == 0 || splits[j].x != splits[j - 1].x;
    pb.num_items2 = (int)items2.size();
    pb.num_splits = (int)splits.size();
    pb.items2 = pool.get<PbItem>(items2.size());
    pb.splits = pool.get<PbItem>(splits.size());
    pb.bin_partial = pool.get<long long>((size_t)std::max(slots, 1) << (kPbBinShift + 1));
    pb.bin_flags = pool.get<uint8_t>((size_t)std::max(slots, 1) << kPbBinShift);
    pb.partials = pool.get<double>((items2.size() + splits.size()) * kMaxReduce);
    if (!pb.src) pb.src = pool.get<uint16_t>(1);
    if (!pb.mask) pb.mask = pool.get<uint8_t>(8);
    if (!pb.win) pb.win = pool.get<int32_t>(2);
    if (!pb.rrow) pb.rrow = pool.get<uint16_t>(1);
    if (!pb.rsum) {
      pb.rsum64 = pool.get<double>(1);
      pb.rsum = reinterpret_cast<float *>(pb.rsum64);
    }
    if (!pb.seg_off) pb.seg_off = pool.get<int32_t>(1);
    unsigned long long *iso_count = pool.get<unsigned long long>(2);
    BTRY(pool.err);
    if (n && with_iso) {
      // the renumbering puts the isolated rows last; if they were not the
      // tail, none is set apart (n_iso = 0)
      unsigned long long h_count[2] = {0, 0};
      BTRY(hipMemsetAsync(iso_count, 0, 2 * sizeof(unsigned long long), s));
      pb_iso_kernel<<<(int)std::min<uint64_t>((n + 255) / 256, 2048), 256, 0, s>>>(
          dg.row_ptr, dg.inv_deg32, (int)n, iso_count);
      BTRY(hipGetLastError());
      BTRY(hipMemcpyAsync(h_count, iso_count, sizeof(h_count), hipMemcpyDeviceToHost, s));
      BTRY(hipStreamSynchronize(s));
      pb.n_iso = (h_count[0] == n - h_count[1]) ? h_count[0] : 0;
    }
    pool.release(iso_count);
    if (!items1.empty())
      BTRY(hipMemcpyAsync(pb.items1, items1.data(), items1.size() * sizeof(PbItem),
                          hipMemcpyHostToDevice, s));
    BTRY(hipMemcpyAsync(pb.items2, items2.data(), items2.size() * sizeof(PbItem),
                        hipMemcpyHostToDevice, s));
    if (!splits.empty())
      BTRY(hipMemcpyAsync(pb.splits, splits.data(), splits.size() * sizeof(PbItem),
                          hipMemcpyHostToDevice, s));
    BTRY(hipStreamSynchronize(s));
  }
#undef BTRY
  for (void *p : {(void *)pb.src, pb.w, (void *)pb.mask, (void *)pb.win, (void *)pb.seg_off,
                  (void *)pb.rrow, (void *)pb.rsum, (void *)pb.items1, (void *)pb.items2,
                  (void *)pb.splits, (void *)pb.bin_partial, (void *)pb.bin_flags,
                  (void *)pb.partials, (void *)pb.bin_row0})
    if (p) pool.keep(p);
  out = pb;
  return GOPS_OK;
fail:
  err = std::string("blocked layout: ") + what + ": " + hipGetErrorString(e);
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GOPS_OUT_OF_MEMORY
                                                                     : GOPS_DEVICE_ERROR;
}

}  // namespace

gops_status build_blocked(gops_context *ctx, gops_graph *g, std::string &err) {
  if (g->has_pb) return GOPS_OK;
  const gops_status st = build_layout(ctx, g->dev, true, g->pb, err);
  g->has_pb = st == GOPS_OK;
  return st;
}

gops_status build_blocked_fwd(gops_context *ctx, gops_graph *g, std::string &err) {
  if (g->has_pb_fwd) return GOPS_OK;
  gops_status st = build_forward(ctx, g, err);
  if (st != GOPS_OK) return st;
  st = build_layout(ctx, g->fwd, false, g->pb_fwd, err);
  g->has_pb_fwd = st == GOPS_OK;
  return st;
}

}  // namespace gops

