// topk.hip -- top_k (topk.rs:17-42) on the device: the k largest finite,
// positive scores, descending, with the entry set the reference's heap
// keeps when scores tie at the boundary.
//
// The reference scans i = 0..n, keeps a min-heap of (score, i) of at most k
// entries and replaces its minimum only by a strictly larger score.  Its
// result is therefore: every score above T, the k-th largest, and among the
// entries equal to T: those that entered the heap (an entry t enters while
// fewer than k entries >= T precede it) minus the ones a larger score later
// evicted (the smallest indices first) -- the last c of them in index order,
// c = k - (number above T).  The final sort of the reference is unstable;
// equal scores come out here by ascending index.
//
// On the device: a radix selection of T over the 64-bit patterns of the
// scores (monotone for positive doubles; invalid scores count as 0), eight
// 8-bit passes of workgroup histograms; then ordered passes over contiguous
// chunks (per-chunk counts, a scan, per-element ranks from wave ballots)
// that write the selected entries in index order; then a sort of the k
// entries (one workgroup's LDS for k <= 4096, hipCUB beyond).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.h"

namespace gops {

namespace {

constexpr int kTB = 256;        // threads per workgroup of the passes
constexpr int kTGrid = 1024;    // workgroups (contiguous chunks) of the passes
constexpr int kHB = 1024;       // threads per workgroup of the histogram passes
constexpr int kHGrid = 256;     // workgroups (histogram rows) of the histogram passes
constexpr int kSortMax = 4096;  // entries sorted in one workgroup's LDS
constexpr int kSortThreads = 1024;
constexpr int kPassU = 4;       // loads in flight per thread in the ordered passes
static_assert(kTGrid <= 1024, "tk_scan_kernel scans one count per thread");

struct TkState {
  uint64_t prefix;  // radix passes: the bits of T found so far; then T
  uint64_t k_rem;   // keys still to take within the prefix range; then c
  uint64_t above;   // keys above the prefix range; then keys above T
  uint64_t valid;   // finite, positive scores
  uint64_t ties_in; // entries equal to T that entered the heap (j)
  uint32_t all;     // valid <= k: every valid score is selected
  uint32_t resolved;  // radix passes stopped early: T is the lower bound of a bucket
                      // whose keys are all taken
};

// A batch of selections runs one per blockIdx.y: query y's scores are
// y * score doubles in, its selected keys and ids y * sel entries in, its
// results y * out entries in; state, histograms, counts and offsets are
// per query at their own sizes.  A single selection has y = 0.
struct TkBatch {
  uint64_t score, sel, out;
};

__device__ __forceinline__ uint64_t skey(double s) {
  return (isfinite(s) && s > 0.0) ? (uint64_t)__double_as_longlong(s) : 0ull;
}

// digit histogram of the valid keys whose bits above `shift + 8` match the
// prefix: one row of 256 bins per workgroup
__global__ __launch_bounds__(kHB) void tk_hist_kernel(const double *s, uint64_t n, int shift,
                                                       const TkState *st, uint32_t *part,
                                                       TkBatch bt) {
  s += blockIdx.y * bt.score;
  st += blockIdx.y;
  part += (size_t)blockIdx.y * kHGrid * 256;
  if (st->all || st->resolved) return;  // the pick kernel reads no histogram then
  __shared__ uint32_t h[256];
  if (threadIdx.x < 256) h[threadIdx.x] = 0;
  __syncthreads();
  {
    const int hi = shift + 8;
    const uint64_t pre = hi < 64 ? (st->prefix >> hi) : 0;
    const int lane = threadIdx.x & 63;
    constexpr int U = 4;  // loads in flight per lane
    const uint64_t step = (uint64_t)gridDim.x * kHB * U;
    for (uint64_t i0 = blockIdx.x * (uint64_t)kHB * U + threadIdx.x; i0 < n; i0 += step) {
      uint64_t key[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint64_t i = i0 + (uint64_t)u * kHB;
        key[u] = i < n ? skey(s[i]) : 0;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const bool in = key[u] && (hi >= 64 || (key[u] >> hi) == pre);
        const int dg = (int)((key[u] >> shift) & 255);
        // Scores crowd into few digits (PageRank's share their top bits), so
        // the lanes with the first counted lane's digit add once, by
        // popcount; the others add one each.
        const uint64_t act = __ballot(in);
        if (act) {
          const int lead = __ffsll((unsigned long long)act) - 1;
          const int d0 = __shfl(dg, lead);
          const uint64_t same = __ballot(in && dg == d0);
          if (lane == lead) atomicAdd(&h[d0], (uint32_t)__popcll(same));
          if (in && dg != d0) atomicAdd(&h[dg], 1u);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < 256) part[(size_t)blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// one workgroup: sum the kHGrid histogram rows (a thread per four bins and
// sixteenth of the rows, its loads all in flight at once), take the digit of
// the k_rem-th largest
constexpr int kPickThreads = 1024;
constexpr int kPickSlices = kPickThreads / 64;  // row slices (64 threads x 4 bins = 256)
static_assert(kHGrid % kPickSlices == 0, "rows split evenly over the slices");
__global__ __launch_bounds__(kPickThreads) void tk_pick_kernel(TkState *st, const uint32_t *part,
                                                                int shift, uint64_t k) {
  st += blockIdx.y;
  part += (size_t)blockIdx.y * kHGrid * 256;
  if (st->all || st->resolved) return;
  __shared__ uint64_t tot[kPickSlices][256];
  const int g = threadIdx.x & 63, sl = threadIdx.x >> 6;
  constexpr int kRows = kHGrid / kPickSlices;
  uint4 v[kRows];
#pragma unroll
  for (int r = 0; r < kRows; r++)
    v[r] = reinterpret_cast<const uint4 *>(part + (size_t)(sl * kRows + r) * 256)[g];
  uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
  for (int r = 0; r < kRows; r++) {
    a0 += v[r].x;
    a1 += v[r].y;
    a2 += v[r].z;
    a3 += v[r].w;
  }
  tot[sl][4 * g] = a0;
  tot[sl][4 * g + 1] = a1;
  tot[sl][4 * g + 2] = a2;
  tot[sl][4 * g + 3] = a3;
  __syncthreads();
  if (threadIdx.x < 256) {
    uint64_t sum = 0;
#pragma unroll
    for (int r = 0; r < kPickSlices; r++) sum += tot[r][threadIdx.x];
    tot[0][threadIdx.x] = sum;
  }
  __syncthreads();
  if (threadIdx.x != 0 || st->all) return;
  const uint64_t *tb = tot[0];
  if (shift == 56) {  // the first pass sees every valid key
    uint64_t v = 0;
    for (int d = 0; d < 256; d++) v += tb[d];
    st->valid = v;
    if (v <= k) {  // all of them; T = 0 selects every valid key as "above"
      st->all = 1;
      st->prefix = 0;
      st->above = v;
      st->k_rem = 0;
      return;
    }
    st->prefix = 0;
    st->above = 0;
    st->k_rem = k;
  }
  const uint64_t need = st->k_rem;
  uint64_t acc = 0;
  int d = 255;
  for (; d > 0; d--) {
    if (acc + tb[d] >= need) break;
    acc += tb[d];
  }
  st->above += acc;
  st->k_rem = need - acc;
  st->prefix |= (uint64_t)d << shift;
  // Every key of the chosen bucket is needed: the keys >= its lower bound
  // (the prefix, remaining bits zero) are exactly the k to select, so T can
  // stay that bound and the remaining passes return at once.  The select
  // kernel then counts the keys equal to the bound as ties that all entered.
  if (shift > 0 && tb[d] == st->k_rem) st->resolved = 1;
}

// exclusive ranks of two flags over a workgroup of kTB threads, and totals
__device__ __forceinline__ void block_rank2(bool fa, bool fb, uint32_t &ra, uint32_t &rb,
                                            uint32_t &ta, uint32_t &tb) {
  __shared__ uint32_t wa[kTB / 64], wb[kTB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t ba = __ballot(fa), bb = __ballot(fb);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (lane == 0) {
    wa[wave] = (uint32_t)__popcll(ba);
    wb[wave] = (uint32_t)__popcll(bb);
  }
  __syncthreads();
  uint32_t oa = 0, ob = 0;
  ta = tb = 0;
#pragma unroll
  for (int w = 0; w < kTB / 64; w++) {
    if (w < wave) {
      oa += wa[w];
      ob += wb[w];
    }
    ta += wa[w];
    tb += wb[w];
  }
  __syncthreads();
  ra = oa + (uint32_t)__popcll(ba & below);
  rb = ob + (uint32_t)__popcll(bb & below);
}

__device__ __forceinline__ void chunk_of(uint64_t n, uint64_t &b, uint64_t &e) {
  const uint64_t c = (n + gridDim.x - 1) / gridDim.x;
  b = std::min<uint64_t>(n, blockIdx.x * c);
  e = std::min<uint64_t>(n, b + c);
}

// per chunk: entries >= T and == T (valid ones)
__global__ __launch_bounds__(kTB) void tk_count_kernel(const double *s, uint64_t n,
                                                        const TkState *st, uint64_t *cnt,
                                                        TkBatch bt) {
  s += blockIdx.y * bt.score;
  st += blockIdx.y;
  cnt += (size_t)blockIdx.y * 2 * kTGrid;
  const uint64_t T = st->prefix;
  uint64_t b, e;
  chunk_of(n, b, e);
  uint32_t ge = 0, eq = 0;
  for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += (uint64_t)kPassU * kTB) {
    uint64_t key[kPassU];
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      const uint64_t i = i0 + (uint64_t)u * kTB;
      key[u] = i < e ? skey(s[i]) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      ge += key[u] != 0 && key[u] >= T;
      eq += key[u] != 0 && key[u] == T;
    }
  }
  __shared__ uint32_t sg[kTB], se[kTB];
  sg[threadIdx.x] = ge;
  se[threadIdx.x] = eq;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t tg = 0, te = 0;
    for (int t = 0; t < kTB; t++) {
      tg += sg[t];
      te += se[t];
    }
    cnt[2 * blockIdx.x] = tg;
    cnt[2 * blockIdx.x + 1] = te;
  }
}

// one workgroup of nblocks (<= 1024) threads: exclusive scans of the chunk counts
__global__ __launch_bounds__(1024) void tk_scan_kernel(const uint64_t *cnt, int nblocks,
                                                        uint64_t *off) {
  cnt += (size_t)blockIdx.y * 2 * kTGrid;
  off += (size_t)blockIdx.y * 2 * kTGrid;
  __shared__ uint64_t sg[1024], sq[1024];
  const int t = threadIdx.x;
  const uint64_t g = t < nblocks ? cnt[2 * t] : 0, q = t < nblocks ? cnt[2 * t + 1] : 0;
  sg[t] = g;
  sq[t] = q;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
    const uint64_t ag = t >= d ? sg[t - d] : 0, aq = t >= d ? sq[t - d] : 0;
    __syncthreads();
    sg[t] += ag;
    sq[t] += aq;
    __syncthreads();
  }
  if (t < nblocks) {
    off[2 * t] = sg[t] - g;
    off[2 * t + 1] = sq[t] - q;
  }
}

// per chunk: entries equal to T that entered the heap (fewer than k entries
// >= T before them)
__global__ __launch_bounds__(kTB) void tk_entered_kernel(const double *s, uint64_t n, uint64_t k,
                                                          const TkState *st, const uint64_t *off,
                                                          uint64_t *part, TkBatch bt) {
  s += blockIdx.y * bt.score;
  st += blockIdx.y;
  off += (size_t)blockIdx.y * 2 * kTGrid;
  part += (size_t)blockIdx.y * kTGrid;
  const uint64_t T = st->prefix;
  uint64_t b, e;
  chunk_of(n, b, e);
  uint64_t run_ge = off[2 * blockIdx.x];
  uint32_t entered = 0;
  for (uint64_t i00 = b; i00 < e; i00 += (uint64_t)kPassU * kTB) {
    uint64_t keys[kPassU];
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      const uint64_t i = i00 + (uint64_t)u * kTB + threadIdx.x;
      keys[u] = i < e ? skey(s[i]) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      const uint64_t key = keys[u];
      const bool ge = key != 0 && key >= T, eq = key != 0 && key == T;
      uint32_t rg, rq, tg, tq;
      block_rank2(ge, eq, rg, rq, tg, tq);
      if (eq && run_ge + rg < k) entered++;
      run_ge += tg;
    }
  }
  __shared__ uint32_t sh[kTB];
  sh[threadIdx.x] = entered;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int x = 0; x < kTB; x++) t += sh[x];
    part[blockIdx.x] = t;
  }
}

// the selected entries in index order: those above T at [0, above), the kept
// ties (the last c of the j that entered) at [above, above + c)
__global__ __launch_bounds__(kTB) void tk_select_kernel(const double *s, uint64_t n, TkState *st,
                                                         const uint64_t *off,
                                                         const uint64_t *entered, int nblocks,
                                                         uint64_t *out_key, uint32_t *out_id,
                                                         TkBatch bt) {
  s += blockIdx.y * bt.score;
  st += blockIdx.y;
  off += (size_t)blockIdx.y * 2 * kTGrid;
  entered += (size_t)blockIdx.y * kTGrid;
  out_key += blockIdx.y * bt.sel;
  out_id += blockIdx.y * bt.sel;
  const uint64_t T = st->prefix;
  __shared__ uint64_t sj[kTB];
  uint64_t tj = 0;
  for (int w = threadIdx.x; w < nblocks; w += kTB) tj += entered[w];
  sj[threadIdx.x] = tj;
  __syncthreads();
  for (int d = kTB / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) sj[threadIdx.x] += sj[threadIdx.x + d];
    __syncthreads();
  }
  const uint64_t j = st->all ? 0 : sj[0];
  // keys above T, and ties kept; stopped early (T a bucket's lower bound,
  // above + k_rem = k keys >= T): every tie entered (j of them) and is kept
  const uint64_t c = st->resolved ? j : st->k_rem;
  const uint64_t above = st->resolved ? st->above + st->k_rem - j : st->above;
  const uint64_t first_kept = j - c;  // tie rank of the first kept tie
  uint64_t b, e;
  chunk_of(n, b, e);
  uint64_t run_ge = off[2 * blockIdx.x], run_eq = off[2 * blockIdx.x + 1];
  for (uint64_t i00 = b; i00 < e; i00 += (uint64_t)kPassU * kTB) {
    uint64_t keys[kPassU];
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      const uint64_t i = i00 + (uint64_t)u * kTB + threadIdx.x;
      keys[u] = i < e ? skey(s[i]) : 0;
    }
#pragma unroll
    for (int u = 0; u < kPassU; u++) {
      const uint64_t i = i00 + (uint64_t)u * kTB + threadIdx.x;
      const uint64_t key = keys[u];
      const bool ge = key != 0 && key >= T, eq = key != 0 && key == T;
      uint32_t rg, rq, tg, tq;
      block_rank2(ge, eq, rg, rq, tg, tq);
      if (ge && !eq) {
        const uint64_t pos = (run_ge + rg) - (run_eq + rq);  // entries above T before i
        out_key[pos] = key;
        out_id[pos] = (uint32_t)i;
      } else if (eq) {
        const uint64_t tr = run_eq + rq;
        if (tr >= first_kept && tr < j) {
          out_key[above + (tr - first_kept)] = key;
          out_id[above + (tr - first_kept)] = (uint32_t)i;
        }
      }
      run_ge += tg;
      run_eq += tq;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->ties_in = j;
}

// one workgroup: sort count (<= kSortMax) entries by descending key, then
// ascending index; write ids and scores
__global__ __launch_bounds__(kSortThreads) void tk_sort_kernel(const uint64_t *key_in,
                                                                const uint32_t *id_in,
                                                                const TkState *st, uint64_t k,
                                                                uint32_t *ids, double *vals,
                                                                uint32_t *count, TkBatch bt) {
  key_in += blockIdx.y * bt.sel;
  id_in += blockIdx.y * bt.sel;
  st += blockIdx.y;
  ids += blockIdx.y * bt.out;
  vals += blockIdx.y * bt.out;
  count += blockIdx.y;
  __shared__ uint64_t sk[kSortMax];
  __shared__ uint32_t si[kSortMax];
  const uint64_t K = st->all ? st->valid : k;
  uint32_t P = 1;
  while (P < K) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) {
    sk[i] = i < K ? key_in[i] : 0ull;
    si[i] = i < K ? id_in[i] : 0xFFFFFFFFu;
  }
  __syncthreads();
  for (uint32_t size = 2; size <= P; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) {
        const uint32_t l = i ^ stride;
        if (l > i) {
          const bool before = sk[i] > sk[l] || (sk[i] == sk[l] && si[i] < si[l]);
          const bool up = (i & size) == 0;  // this run sorted "first before"
          if (before != up) {
            const uint64_t tk = sk[i];
            sk[i] = sk[l];
            sk[l] = tk;
            const uint32_t ti = si[i];
            si[i] = si[l];
            si[l] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < K; i += kSortThreads) {
    ids[i] = si[i];
    vals[i] = __longlong_as_double((long long)sk[i]);
  }
  if (threadIdx.x == 0) *count = (uint32_t)K;
}

// entries [K, k) (fewer valid scores than k) sort last as key 0
__global__ __launch_bounds__(kTB) void tk_pad_kernel(uint64_t *key, const TkState *st, uint64_t k) {
  const uint64_t K = st->all ? st->valid : k;
  for (uint64_t i = K + blockIdx.x * (uint64_t)kTB + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * kTB)
    key[i] = 0;
}

// after a hipCUB sort of count entries: ids and scores, and the count
__global__ __launch_bounds__(kTB) void tk_emit_kernel(const uint64_t *key, const uint32_t *id,
                                                       const TkState *st, uint64_t k,
                                                       uint32_t *ids, double *vals,
                                                       uint32_t *count) {
  const uint64_t K = st->all ? st->valid : k;
  for (uint64_t i = blockIdx.x * (uint64_t)kTB + threadIdx.x; i < K;
       i += (uint64_t)gridDim.x * kTB) {
    ids[i] = id[i];
    vals[i] = __longlong_as_double((long long)key[i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = (uint32_t)K;
}

}  // namespace

// the selection state (and, on request, the result count) set to zero by a
// kernel rather than a memset, so that every step of a selection is a kernel
__global__ void tk_clear_kernel(TkState *st, uint32_t *count) {
  if (threadIdx.x == 0) {
    if (st) st[blockIdx.y] = TkState{};
    if (count) count[blockIdx.y] = 0;
  }
}

struct TopkScratch {
  uint32_t *part = nullptr;     // kHGrid * 256 histogram bins
  uint64_t *cnt = nullptr;      // 2 * kTGrid chunk counts
  uint64_t *off = nullptr;      // 2 * kTGrid chunk offsets
  uint64_t *entered = nullptr;  // kTGrid
  TkState *st = nullptr;
  size_t cap = 0;               // entries of the arrays below
  uint64_t *key[2] = {nullptr, nullptr};
  uint32_t *id[2] = {nullptr, nullptr};
  uint32_t *ids = nullptr;
  double *vals = nullptr;
  uint32_t *count = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
};

void free_topk(TopkScratch *t) {
  if (!t) return;
  for (void *p : {(void *)t->part, (void *)t->cnt, (void *)t->off, (void *)t->entered,
                  (void *)t->st, (void *)t->key[0], (void *)t->key[1], (void *)t->id[0],
                  (void *)t->id[1], (void *)t->ids, (void *)t->vals, (void *)t->count, t->tmp})
    (void)hipFree(p);
  delete t;
}

static hipError_t topk_scratch(gops_context *ctx, uint32_t k) {
  if (!ctx->topk) ctx->topk = new TopkScratch;
  TopkScratch &t = *ctx->topk;
  hipError_t e = hipSuccess;
  auto get = [&](auto **p, size_t count) {
    if (e == hipSuccess && !*p) e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(**p));
  };
  get(&t.part, (size_t)kHGrid * 256);
  get(&t.cnt, 2 * kTGrid);
  get(&t.off, 2 * kTGrid);
  get(&t.entered, kTGrid);
  get(&t.st, 1);
  get(&t.count, 1);
  if (e != hipSuccess) return e;
  if (t.cap < k) {
    for (void *p : {(void *)t.key[0], (void *)t.key[1], (void *)t.id[0], (void *)t.id[1],
                    (void *)t.ids, (void *)t.vals, t.tmp})
      (void)hipFree(p);
    t.key[0] = t.key[1] = nullptr;
    t.id[0] = t.id[1] = nullptr;
    t.ids = nullptr;
    t.vals = nullptr;
    t.tmp = nullptr;
    t.cap = 0;
    t.tmp_bytes = 0;
    get(&t.key[0], k);
    get(&t.key[1], k);
    get(&t.id[0], k);
    get(&t.id[1], k);
    get(&t.ids, k);
    get(&t.vals, k);
    if (e != hipSuccess) return e;
    if (k > (uint32_t)kSortMax) {
      e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t.tmp_bytes, t.key[0], t.key[1],
                                                       t.id[0], t.id[1], (int)k, 0, 64,
                                                       ctx->stream);
      if (e != hipSuccess) return e;
      e = hipMalloc(&t.tmp, std::max<size_t>(t.tmp_bytes, 1));
      if (e != hipSuccess) return e;
    }
    t.cap = k;
  }
  return hipSuccess;
}

// Queues the selection of top_k of n device scores on stream s with the
// given scratch (kk = min(k, n) > 0 entries; kk <= t.cap, and t.tmp sized
// when kk > kSortMax): ids / vals (kk entries) and *count on the device.
static hipError_t launch_top_k_in(hipStream_t s, const double *scores, uint64_t n, uint64_t kk,
                                  uint32_t *ids, double *vals, uint32_t *count,
                                  const TopkScratch &t) {
  const TkBatch one{0, 0, 0};
  tk_clear_kernel<<<1, 64, 0, s>>>(t.st, nullptr);
  for (int shift = 56; shift >= 0; shift -= 8) {
    tk_hist_kernel<<<kHGrid, kHB, 0, s>>>(scores, n, shift, t.st, t.part, one);
    tk_pick_kernel<<<1, kPickThreads, 0, s>>>(t.st, t.part, shift, kk);
  }
  tk_count_kernel<<<kTGrid, kTB, 0, s>>>(scores, n, t.st, t.cnt, one);
  tk_scan_kernel<<<1, 1024, 0, s>>>(t.cnt, kTGrid, t.off);
  tk_entered_kernel<<<kTGrid, kTB, 0, s>>>(scores, n, kk, t.st, t.off, t.entered, one);
  tk_select_kernel<<<kTGrid, kTB, 0, s>>>(scores, n, t.st, t.off, t.entered, kTGrid, t.key[0],
                                         t.id[0], one);
  if (kk <= (uint64_t)kSortMax) {
    tk_sort_kernel<<<1, kSortThreads, 0, s>>>(t.key[0], t.id[0], t.st, kk, ids, vals, count, one);
    return hipGetLastError();
  }
  // the sort is stable and the entries are in index order within equal keys
  const int g = (int)std::min<uint64_t>((kk + kTB - 1) / kTB, 4096);
  tk_pad_kernel<<<g, kTB, 0, s>>>(t.key[0], t.st, kk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t b = t.tmp_bytes;
  e = hipcub::DeviceRadixSort::SortPairsDescending(t.tmp, b, t.key[0], t.key[1], t.id[0], t.id[1],
                                                   (int)kk, 0, 64, s);
  if (e != hipSuccess) return e;
  tk_emit_kernel<<<g, kTB, 0, s>>>(t.key[1], t.id[1], t.st, kk, ids, vals, count);
  return hipGetLastError();
}

// Queues the selection of top_k of n device scores on ctx's stream: ids /
// vals (k entries) and *count on the device.  Calls on one context share its
// scratch and run in stream order.
hipError_t launch_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                        uint32_t *ids, double *vals, uint32_t *count) {
  hipStream_t s = ctx->stream;
  if (k == 0 || n == 0) {
    tk_clear_kernel<<<1, 64, 0, s>>>(nullptr, count);
    return hipGetLastError();
  }
  const uint64_t kk = std::min<uint64_t>(k, n);
  hipError_t e = topk_scratch(ctx, (uint32_t)kk);
  if (e != hipSuccess) return e;
  return launch_top_k_in(s, scores, n, kk, ids, vals, count, *ctx->topk);
}

// Scratch of batched selections: per query a state, histogram rows, chunk
// counts, offsets and tie counts, and k selected keys and ids.
struct TopkBatchScratch {
  uint32_t cap_q = 0, cap_k = 0;
  uint32_t *part = nullptr;
  uint64_t *cnt = nullptr, *off = nullptr, *entered = nullptr;
  TkState *st = nullptr;
  uint64_t *key = nullptr;
  uint32_t *id = nullptr;
};

void free_topk_batch(TopkBatchScratch *t) {
  if (!t) return;
  for (void *p : {(void *)t->part, (void *)t->cnt, (void *)t->off, (void *)t->entered,
                  (void *)t->st, (void *)t->key, (void *)t->id})
    (void)hipFree(p);
  delete t;
}

// Queues top_k of nq score vectors (n doubles each, n apart in scores) on
// ctx's stream, every kernel launched once for all of them (blockIdx.y is
// the query): query y's ids / vals at y * k and its count at count[y].
// k <= kSortMax (the one-workgroup sort).
hipError_t launch_top_k_batch(gops_context *ctx, const double *scores, uint64_t n, uint32_t nq,
                              uint32_t k, uint32_t *ids, double *vals, uint32_t *count) {
  hipStream_t s = ctx->stream;
  if (nq == 0) return hipSuccess;
  if (k > (uint32_t)kSortMax) return hipErrorInvalidValue;
  if (k == 0 || n == 0) {
    tk_clear_kernel<<<dim3(1, nq), 64, 0, s>>>(nullptr, count);
    return hipGetLastError();
  }
  if (!ctx->topk_batch) ctx->topk_batch = new TopkBatchScratch;
  TopkBatchScratch &t = *ctx->topk_batch;
  const uint64_t kk = std::min<uint64_t>(k, n);
  if (t.cap_q < nq || t.cap_k < kk) {
    for (void *p : {(void *)t.part, (void *)t.cnt, (void *)t.off, (void *)t.entered, (void *)t.st,
                    (void *)t.key, (void *)t.id})
      (void)hipFree(p);
    t = TopkBatchScratch{};
    const size_t q = nq;
    hipError_t e = hipMalloc((void **)&t.part, q * kHGrid * 256 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&t.cnt, q * 2 * kTGrid * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&t.off, q * 2 * kTGrid * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&t.entered, q * kTGrid * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&t.st, q * sizeof(TkState));
    if (e == hipSuccess) e = hipMalloc((void **)&t.key, q * kk * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&t.id, q * kk * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    t.cap_q = nq;
    t.cap_k = (uint32_t)kk;
  }
  const TkBatch bt{n, t.cap_k, k};
  tk_clear_kernel<<<dim3(1, nq), 64, 0, s>>>(t.st, nullptr);
  for (int shift = 56; shift >= 0; shift -= 8) {
    tk_hist_kernel<<<dim3(kHGrid, nq), kHB, 0, s>>>(scores, n, shift, t.st, t.part, bt);
    tk_pick_kernel<<<dim3(1, nq), kPickThreads, 0, s>>>(t.st, t.part, shift, kk);
  }
  tk_count_kernel<<<dim3(kTGrid, nq), kTB, 0, s>>>(scores, n, t.st, t.cnt, bt);
  tk_scan_kernel<<<dim3(1, nq), 1024, 0, s>>>(t.cnt, kTGrid, t.off);
  tk_entered_kernel<<<dim3(kTGrid, nq), kTB, 0, s>>>(scores, n, kk, t.st, t.off, t.entered, bt);
  tk_select_kernel<<<dim3(kTGrid, nq), kTB, 0, s>>>(scores, n, t.st, t.off, t.entered, kTGrid,
                                                   t.key, t.id, bt);
  tk_sort_kernel<<<dim3(1, nq), kSortThreads, 0, s>>>(t.key, t.id, t.st, kk, ids, vals, count, bt);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
namespace {

// The allocation-free variant's scratch: one file-scope array, which every
// GPU the library runs on holds (its own copy) until the process ends.  The
// selection's state, histogram rows, chunk counts and offsets, the kNoAllocK
// selected keys and ids, and the results before they cross to the host.
constexpr uint32_t kNoAllocK = kNoAllocTopK;  // k at most: the one-workgroup sort
static_assert(kNoAllocK <= (uint32_t)kSortMax, "the selected entries are sorted in one workgroup");
constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaPart = 0;
constexpr size_t kNaCnt = kNaPart + na_round((size_t)kHGrid * 256 * sizeof(uint32_t));
constexpr size_t kNaOff = kNaCnt + na_round(2 * (size_t)kTGrid * sizeof(uint64_t));
constexpr size_t kNaEntered = kNaOff + na_round(2 * (size_t)kTGrid * sizeof(uint64_t));
constexpr size_t kNaState = kNaEntered + na_round((size_t)kTGrid * sizeof(uint64_t));
constexpr size_t kNaKey = kNaState + na_round(sizeof(TkState));
constexpr size_t kNaId = kNaKey + na_round((size_t)kNoAllocK * sizeof(uint64_t));
constexpr size_t kNaIds = kNaId + na_round((size_t)kNoAllocK * sizeof(uint32_t));
constexpr size_t kNaVals = kNaIds + na_round((size_t)kNoAllocK * sizeof(uint32_t));
constexpr size_t kNaCount = kNaVals + na_round((size_t)kNoAllocK * sizeof(double));
constexpr size_t kNoAllocBytes = kNaCount + na_round(sizeof(uint32_t));  // about 400 KB

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

}  // namespace

// The device address of `scratch` on the current device.  In host code the
// array's name stands for no device address, so kernels and copies take this
// pointer; it is looked up once per device and kept.
hipError_t noalloc_scratch(int device, void **out) {
  static std::mutex mu;
  static void *ptr[kMaxDevices] = {};
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[device] = p;
  }
  *out = ptr[device];
  return hipSuccess;
}

// top_k of n device scores into host ids / vals with the scratch at `base`
// (the array above on ctx's device); kk = min(k, n) <= kNoAllocK.
gops_status top_k_noalloc(gops_context *ctx, unsigned char *base, const double *scores,
                          uint64_t n, uint32_t k, uint32_t *ids, double *vals, uint32_t *count,
                          std::string &err) {
  hipStream_t s = ctx->stream;
  TopkScratch t;
  t.part = reinterpret_cast<uint32_t *>(base + kNaPart);
  t.cnt = reinterpret_cast<uint64_t *>(base + kNaCnt);
  t.off = reinterpret_cast<uint64_t *>(base + kNaOff);
  t.entered = reinterpret_cast<uint64_t *>(base + kNaEntered);
  t.st = reinterpret_cast<TkState *>(base + kNaState);
  t.key[0] = reinterpret_cast<uint64_t *>(base + kNaKey);
  t.id[0] = reinterpret_cast<uint32_t *>(base + kNaId);
  t.cap = kNoAllocK;
  uint32_t *d_ids = reinterpret_cast<uint32_t *>(base + kNaIds);
  double *d_vals = reinterpret_cast<double *>(base + kNaVals);
  uint32_t *d_count = reinterpret_cast<uint32_t *>(base + kNaCount);
  const uint64_t kk = std::min<uint64_t>(k, n);
  uint32_t cnt = 0;
  hipError_t e = launch_top_k_in(s, scores, n, kk, d_ids, d_vals, d_count, t);
  if (e == hipSuccess) e = hipMemcpyAsync(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess && cnt) {
    if (ids) e = hipMemcpyAsync(ids, d_ids, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && vals)
      e = hipMemcpyAsync(vals, d_vals, cnt * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    err = std::string("top_k without allocation: ") + hipGetErrorString(e);
    return GOPS_DEVICE_ERROR;
  }
  *count = cnt;
  return GOPS_OK;
}

// top_k of n device scores into host ids / values (count entries, <= k)
gops_status device_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                         uint32_t *ids, double *vals, uint32_t *count, std::string &err) {
  *count = 0;
  if (k == 0 || n == 0) return GOPS_OK;  // topk.rs:18-20
  if (n >= (1ull << 32)) {
    err = "top_k: at most 2^32 - 1 scores";
    return GOPS_NOT_SUPPORTED;
  }
  hipStream_t s = ctx->stream;
  hipError_t e = topk_scratch(ctx, (uint32_t)std::min<uint64_t>(k, n));
  uint32_t cnt = 0;
  if (e == hipSuccess) {
    TopkScratch &t = *ctx->topk;
    e = launch_top_k(ctx, scores, n, k, t.ids, t.vals, t.count);
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, t.count, sizeof(cnt), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && cnt) {
      if (ids) e = hipMemcpyAsync(ids, t.ids, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess && vals)
        e = hipMemcpyAsync(vals, t.vals, cnt * sizeof(double), hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
  }
  if (e != hipSuccess) {
    err = std::string("top_k: ") + hipGetErrorString(e);
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GOPS_OUT_OF_MEMORY
                                                                       : GOPS_DEVICE_ERROR;
  }
  *count = cnt;
  return GOPS_OK;
}

}  // namespace gops
