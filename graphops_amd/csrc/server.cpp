// server.cpp -- batched personalized PageRank (gops_ppr_batch_run) and the
// serving loop (gops_server_*), both on resident tiles of T queries
// (ppr_batch.hip).
//
// Serving: queries wait in a host queue; between device steps, every slot
// whose query has finished takes the queue head (with a leading f64 phase, a
// group of the slots one lane serves is refilled once all of its queries have
// finished, with queries of one damping where the queue has them).
// Admission clears the slots' columns and starts the new queries with their
// personalization as the input delta while the tile's running queries
// continue, so new and running queries share every later sweep (their work
// overlaps on the device).  Every query keeps its own sweep mode, so its
// arithmetic does not depend on the other queries of its tile.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <new>


#include "internal.h"
#include "ppr_batch.h"

using namespace gops;

#ifndef GOPS_TUNE_TILE_AHEAD
#define GOPS_TUNE_TILE_AHEAD 4
#endif
#ifndef GOPS_TUNE_TILE_REFILL
#define GOPS_TUNE_TILE_REFILL 4
#endif

namespace {

gops_status sfail(gops_context *ctx, gops_status st, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return st;
}

#define STRY(ctx, expr)                                                                  \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return sfail(ctx,                                                                  \
                   (_e == hipErrorOutOfMemory) ? GOPS_OUT_OF_MEMORY : GOPS_DEVICE_ERROR, \
                   std::string(#expr) + ": " + hipGetErrorString(_e));                   \
  } while (0)

inline int lanes_pow2(int want, int cap) {
  int g = 1;
  while (g < want && g < cap) g <<= 1;
  return g;
}

// Work descriptors of the tile sweep for tile width T (rows already sorted
// by in-degree).  One edge's T queries take Q = T / kLaneQ lanes; a row gets
// 2^s slices of Q lanes, about four edges per slice.  Runs of rows share a
// slice count; rows long enough to give every slice of a workgroup four
// edges get the whole workgroup; rows longer than kHubChunk are split into
// the graph's hub chunks.
void plan_tile(const std::vector<int32_t> &rp, int64_t n_in, int T, const DevGraph &dg,
               const std::vector<int32_t> &hub_row, const std::vector<int32_t> &hub_nchunks,
               std::vector<BlockDesc> &desc) {
  desc.clear();
  const int64_t n = std::min<int64_t>((int64_t)rp.size() - 1, n_in);  // rows with in-edges
  const int Q = T / kLaneQ;
  const int smax = 64 / Q;  // slices of a row within one wave
  const int block_row_min = 4 * (kBlock / Q);
  int h = 0;
  int64_t i = 0;
  while (i < n) {
    const int32_t deg = rp[i + 1] - rp[i];
    if (deg > kHubChunk) {
      // graph plan: hub h is this row, same chunking
      for (int32_t c = 0; c < hub_nchunks[h]; c++) {
        const int32_t b = rp[i] + c * kHubChunk;
        desc.push_back({kHubFlag | c, h, b, std::min(b + kHubChunk, rp[i + 1])});
      }
      (void)hub_row;
      h++;
      i++;
      continue;
    }
    if (deg > block_row_min) {
      desc.push_back({kRowFlag, (int32_t)i, 0, 0});
      i++;
      continue;
    }
    const int S = lanes_pow2((deg + 3) / 4, smax);
    int ls = 0;
    while ((1 << ls) < S) ls++;
    const int L = Q << ls;
    const int64_t max_rows = (int64_t)(kBlock / L) * kRowIters;
    const int64_t s = i;
    int64_t nnz = 0;
    while (i < n && i - s < max_rows) {
      const int32_t d = rp[i + 1] - rp[i];
      if (d > block_row_min || lanes_pow2((d + 3) / 4, smax) != S) break;
      if (i > s && nnz + d > 8192) break;
      nnz += d;
      i++;
    }
    desc.push_back({ls, (int32_t)s, (int32_t)i, 0});
  }
  (void)dg;
}

struct Query {
  uint64_t ticket = 0;
  gops_pagerank_config cfg{};
  std::vector<std::pair<uint32_t, double>> pers;  // (new id, normalised value)
  double init_dsum = 0.0;
  double init_mass = 0.0;
  double submit_ms = 0.0;
  bool dense_fallback = false;  // personalization sum <= 0 -> uniform (ppr.rs:105-109)
};

// One resident tile: device buffers for T queries.
struct Tile {
  int T = 16;
  size_t n = 0;
  double *x = nullptr, *y64a = nullptr, *y64b = nullptr, *partials = nullptr,
         *hub_partial = nullptr;
  float *yd0 = nullptr, *yd1 = nullptr, *carry = nullptr;
  int32_t *seed_idx = nullptr;
  int32_t *seed_row = nullptr, *seed_q = nullptr;
  double *seed_p = nullptr;
  int seed_cap = 0;
  int num_seed = 0;
  int zseed_first = 0;  // first seed entry (sorted by row) of a row without in-edges
  std::vector<int32_t> marked_rows;  // rows currently marked in seed_idx
  TileState *ts = nullptr;
  TileState *hts = nullptr;  // pinned host copy
  // GOPS_LAUNCH_GRAPH: kTileChain sweeps captured as one HIP graph for the
  // arguments in chain_args (an admission may move the seed arrays; the
  // chain is captured again when the arguments change)
  hipGraphExec_t chain = nullptr;
  PprArgs chain_args{};
  Query slot[kMaxTile];
  bool used[kMaxTile] = {};
  // per slot, the seed rows without in-edges of the last query admitted into
  // it: its columns may be nonzero there until the slot's next admission
  std::vector<int32_t> zrows[kMaxTile];
  std::vector<int32_t> clear_host;  // the rows the last admission cleared
  int32_t *clear_rows = nullptr;    // device copy
  int clear_cap = 0;
  int grid = 0;
};

struct TilePlan {
  int T = 0;
  BlockDesc *desc = nullptr;
  int num_desc = 0;
};

}  // namespace

struct gops_server {
  gops_context *ctx = nullptr;
  gops_graph *g = nullptr;
  bool holds_graph = false;  // a reference on g (taken once creation succeeds)
  int T = 16;
  std::vector<Tile *> tiles;
  std::deque<Query> queue;
  struct Done {
    Query q;
    gops_run_info info;
    double finish_ms;
    double *scores;  // device copy of the result (n doubles), owned; null with topk
    std::vector<uint32_t> top_ids;  // topk > 0: the result's top_k, made on the device
    std::vector<double> top_vals;
  };
  double *gather = nullptr;  // topk > 0: gather_cap queries' scores, reused by every harvest
  uint32_t gather_cap = 0;
  uint32_t *top_ids = nullptr;  // topk > 0: a harvest's selections on the device,
  double *top_vals = nullptr;   // top_cap results of topk pairs, and their counts
  uint32_t *top_cnt = nullptr;
  uint32_t top_cap = 0;
  std::deque<Done> finished;
  uint64_t next_ticket = 1;
  uint64_t n_finished = 0;
  std::chrono::steady_clock::time_point t0;
  uint32_t topk = 0;
  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace gops {

// Per-graph host data the batched path needs (built on first use).
struct PprGraphCache {
  std::vector<int32_t> row_ptr;
  std::vector<uint8_t> dangling;  // new ids, out-degree 0
  std::vector<int32_t> hub_row, hub_nchunks;
  int32_t n_in = 0;  // rows with in-edges: [0, n_in) (rows are sorted by in-degree, descending)
  TilePlan plans[4];  // per tile width
  int tuned_T = 0;          // autotuned tile width (0: not measured yet)
  double tile_cost[4] = {}; // measured ms per query-sweep of widths 8/16/32/64
};

}  // namespace gops

namespace {

gops_status graph_cache(gops_context *ctx, gops_graph *g, PprGraphCache **out) {
  if (g->ppr_cache) {
    *out = g->ppr_cache;
    return GOPS_OK;
  }
  auto *c = new PprGraphCache;
  const size_t n = g->dev.n;
  c->row_ptr.resize(n + 1);
  std::vector<double> inv(n);
  hipError_t e = hipMemcpy(c->row_ptr.data(), g->dev.row_ptr, (n + 1) * sizeof(int32_t),
                           hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(inv.data(), g->dev.inv_deg64, n * sizeof(double), hipMemcpyDeviceToHost);
  c->dangling.resize(n);
  for (size_t i = 0; i < n; i++) c->dangling[i] = inv[i] == 0.0;
  // the rows without in-edges come last; if they ever did not, every row
  // is swept (n_in = n)
  {
    size_t r = n;
    while (r > 0 && c->row_ptr[r] == c->row_ptr[r - 1]) r--;
    bool tail = true;
    for (size_t i = 0; i < r && tail; i++) tail = c->row_ptr[i + 1] > c->row_ptr[i];
    c->n_in = (int32_t)(tail ? r : n);
  }
  c->hub_row.resize(g->dev.num_hubs);
  c->hub_nchunks.resize(g->dev.num_hubs);
  if (e == hipSuccess && g->dev.num_hubs)
    e = hipMemcpy(c->hub_row.data(), g->dev.hub_row, g->dev.num_hubs * sizeof(int32_t),
                  hipMemcpyDeviceToHost);
  if (e == hipSuccess && g->dev.num_hubs)
    e = hipMemcpy(c->hub_nchunks.data(), g->dev.hub_nchunks, g->dev.num_hubs * sizeof(int32_t),
                  hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    delete c;
    return sfail(ctx, GOPS_DEVICE_ERROR, std::string("graph cache: ") + hipGetErrorString(e));
  }
  g->ppr_cache = c;
  *out = c;
  return GOPS_OK;
}

int plan_slot(int T) { return T == 8 ? 0 : T == 16 ? 1 : T == 32 ? 2 : 3; }

gops_status tile_plan(gops_context *ctx, gops_graph *g, PprGraphCache *c, int T, TilePlan **out) {
  TilePlan &p = c->plans[plan_slot(T)];
  if (!p.desc) {
    std::vector<BlockDesc> desc;
    plan_tile(c->row_ptr, c->n_in, T, g->dev, c->hub_row, c->hub_nchunks, desc);
    STRY(ctx, hipMalloc((void **)&p.desc, std::max<size_t>(desc.size(), 1) * sizeof(BlockDesc)));
    STRY(ctx, hipMemcpy(p.desc, desc.data(), desc.size() * sizeof(BlockDesc),
                        hipMemcpyHostToDevice));
    p.num_desc = (int)desc.size();
    p.T = T;
  }
  *out = &p;
  return GOPS_OK;
}

void free_tile(Tile *t) {
  if (!t) return;
  if (t->chain) (void)hipGraphExecDestroy(t->chain);
  hipFree(t->x);
  hipFree(t->y64a);
  hipFree(t->y64b);
  hipFree(t->partials);
  hipFree(t->hub_partial);
  hipFree(t->yd0);
  hipFree(t->yd1);
  hipFree(t->carry);
  hipFree(t->seed_idx);
  hipFree(t->seed_row);
  hipFree(t->seed_q);
  hipFree(t->seed_p);
  hipFree(t->clear_rows);
  hipFree(t->ts);
  hipHostFree(t->hts);
  delete t;
}

// x, y64a, y64b (f64), the two stored deltas and the carry (f32) per
// (node, query), plus seed_idx per node
size_t tile_bytes(size_t n, int T) { return n * (size_t)T * (8 + 8 + 8 + 4 + 4 + 4) + n * 4; }

gops_status alloc_tile(gops_context *ctx, gops_graph *g, int T, int grid, Tile **out) {
  auto *t = new (std::nothrow) Tile;
  if (!t) return sfail(ctx, GOPS_OUT_OF_MEMORY, "host allocation failed");
  t->T = T;
  t->n = g->dev.n;
  t->grid = grid;
  const size_t nt = t->n * (size_t)T;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (r != hipSuccess && e == hipSuccess) e = r;
  };
  chk(hipMalloc((void **)&t->x, std::max<size_t>(nt, 1) * 8));
  chk(hipMalloc((void **)&t->y64a, std::max<size_t>(nt, 1) * 8));
  chk(hipMalloc((void **)&t->y64b, std::max<size_t>(nt, 1) * 8));
  chk(hipMalloc((void **)&t->yd0, std::max<size_t>(nt, 1) * 4));
  chk(hipMalloc((void **)&t->yd1, std::max<size_t>(nt, 1) * 4));
  chk(hipMalloc((void **)&t->carry, std::max<size_t>(nt, 1) * 4));
  chk(hipMalloc((void **)&t->partials, (size_t)(kMaxSweepBlocks + kTileReduce) * T * 5 * 8));
  chk(hipMalloc((void **)&t->hub_partial, (size_t)std::max(g->dev.num_hub_chunks, 1) * T * 8));
  chk(hipMalloc((void **)&t->seed_idx, std::max<size_t>(t->n, 1) * 4));
  chk(hipMalloc((void **)&t->ts, sizeof(TileState)));
  chk(hipHostMalloc((void **)&t->hts, sizeof(TileState), hipHostMallocDefault));
  // every vector starts at 0: an admission clears the admitted columns only
  // where an earlier query may have left values (ppr_clear_kernel)
  if (e == hipSuccess) chk(hipMemsetAsync(t->seed_idx, 0xff, std::max<size_t>(t->n, 1) * 4, ctx->stream));
  for (double *v : {t->x, t->y64a, t->y64b})
    if (e == hipSuccess) chk(hipMemsetAsync(v, 0, std::max<size_t>(nt, 1) * 8, ctx->stream));
  for (float *v : {t->yd0, t->yd1, t->carry})
    if (e == hipSuccess) chk(hipMemsetAsync(v, 0, std::max<size_t>(nt, 1) * 4, ctx->stream));
  if (e == hipSuccess) chk(hipStreamSynchronize(ctx->stream));
  if (e != hipSuccess) {
    free_tile(t);
    return sfail(ctx, e == hipErrorOutOfMemory ? GOPS_OUT_OF_MEMORY : GOPS_DEVICE_ERROR,
                 std::string("tile allocation: ") + hipGetErrorString(e));
  }
  std::memset(t->hts, 0, sizeof(TileState));
  t->hts->all_done = 1;
  // the device copy must be valid before any admission: a server reads every
  // tile's state back after each step (sync_state), idle tiles included, and
  // uploads that copy again at the tile's first admission
  e = hipMemcpyAsync(t->ts, t->hts, sizeof(TileState), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free_tile(t);
    return sfail(ctx, GOPS_DEVICE_ERROR, std::string("tile state: ") + hipGetErrorString(e));
  }
  *out = t;
  return GOPS_OK;
}

PprArgs tile_args(gops_context *ctx, gops_graph *g, Tile *t, const TilePlan *p, bool mirror) {
  PprArgs a{};
  a.ts = t->ts;
  a.desc = p->desc;
  a.num_desc = p->num_desc;
  a.row_ptr = g->dev.row_ptr;
  a.col = g->dev.col;
  a.inv32 = g->dev.inv_deg32;
  a.inv64 = g->dev.inv_deg64;
  a.x = t->x;
  a.yd0 = t->yd0;
  a.yd1 = t->yd1;
  a.carry = t->carry;
  a.y64a = t->y64a;
  a.y64b = t->y64b;
  a.partials = t->partials;
  a.hub_partial = t->hub_partial;
  a.hub_row = g->dev.hub_row;
  a.hub_nchunks = g->dev.hub_nchunks;
  a.hub_first = g->dev.hub_first;
  a.num_hubs = g->dev.num_hubs;
  a.seed_idx = t->seed_idx;
  a.seed_row = t->seed_row;
  a.seed_q = t->seed_q;
  a.seed_p = t->seed_p;
  a.num_seed_entries = t->num_seed;
  a.zseed_first = t->zseed_first;
  a.n_in = g->ppr_cache ? g->ppr_cache->n_in : (int)g->dev.n;
  a.mirror = mirror ? ctx->mirror_dev : nullptr;
  return a;
}

// Queue `count` sweeps of tile t with arguments a: kernel by kernel, or
// (GOPS_LAUNCH_GRAPH) as replays of a captured chain of kTileChain sweeps,
// rounded up to whole chains.  Every sweep of a chain has the same
// arguments (the device picks buffers, modes and which queries are live),
// so one capture serves until an admission changes them; sweeps past a
// query's stop or cap return at once.  Returns the number queued.
constexpr uint32_t kTileChain = 4;

hipError_t queue_tile_sweeps(gops_context *ctx, Tile *t, const PprArgs &a, uint32_t count,
                             uint32_t *queued) {
  *queued = 0;
  if (ctx->launch != GOPS_LAUNCH_GRAPH) {
    for (uint32_t k = 0; k < count; k++) {
      const hipError_t e = ppr_launch_sweep(ctx->stream, t->T, a, t->grid);
      if (e != hipSuccess) return e;
      ++*queued;
    }
    return hipSuccess;
  }
  if (t->chain && std::memcmp(&t->chain_args, &a, sizeof(PprArgs)) != 0) {
    (void)hipGraphExecDestroy(t->chain);
    t->chain = nullptr;
  }
  if (!t->chain) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t ce = hipSuccess;
    for (uint32_t k = 0; k < kTileChain && ce == hipSuccess; k++)
      ce = ppr_launch_sweep(ctx->stream, t->T, a, t->grid);
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (ce != hipSuccess || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return ce != hipSuccess ? ce : e;
    }
    e = hipGraphInstantiate(&t->chain, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      t->chain = nullptr;
      return e;
    }
    t->chain_args = a;
  }
  for (uint32_t k = 0; k < count; k += kTileChain) {
    const hipError_t e = hipGraphLaunch(t->chain, ctx->stream);
    if (e != hipSuccess) return e;
    *queued += kTileChain;
  }
  return hipSuccess;
}

// Normalised sparse personalization of one query (ppr.rs:104-109, summed in
// node order like the reference's dense sum); duplicates merged.
gops_status make_query(gops_context *ctx, gops_graph *g, PprGraphCache *c,
                       const gops_pagerank_config *cfg, const uint32_t *seeds, uint64_t ns,
                       const double *w, Query &q) {
  const uint64_t n = g->dev.n;
  std::map<uint32_t, double> acc;  // original id -> weight, node order
  for (uint64_t j = 0; j < ns; j++) {
    if (seeds[j] >= n)
      return sfail(ctx, GOPS_INDEX_OUT_OF_BOUNDS, "index out of bounds: " + std::to_string(seeds[j]));
    acc[seeds[j]] += w ? w[j] : 1.0;
  }
  double psum = 0.0;
  for (auto &kv : acc) psum += kv.second;
  q.cfg = *cfg;
  q.pers.clear();
  q.init_dsum = 0.0;
  q.init_mass = 0.0;
  q.dense_fallback = !(psum > 0.0);
  if (q.dense_fallback) return GOPS_OK;
  for (auto &kv : acc) {
    const uint32_t r = g->rank_host[kv.first];
    const double v = kv.second / psum;
    q.pers.push_back({r, v});
    if (c->dangling[r]) q.init_dsum += v;
    q.init_mass += v;
  }
  return GOPS_OK;
}

// Upload the seed entries of every used slot and re-mark seed_idx.
gops_status upload_seeds(gops_context *ctx, Tile *t, int32_t g_n_in) {
  struct E {
    int32_t row, q;
    double p;
  };
  std::vector<E> ent;
  for (int s = 0; s < t->T; s++)
    if (t->used[s])
      for (auto &pv : t->slot[s].pers) ent.push_back({(int32_t)pv.first, s, pv.second});
  std::sort(ent.begin(), ent.end(), [](const E &a, const E &b) {
    return a.row != b.row ? a.row < b.row : a.q < b.q;
  });
  const int cnt = (int)ent.size();
  if (cnt > t->seed_cap) {
    hipFree(t->seed_row);
    hipFree(t->seed_q);
    hipFree(t->seed_p);
    const int cap = std::max(cnt, 1024);
    STRY(ctx, hipMalloc((void **)&t->seed_row, cap * 4));
    STRY(ctx, hipMalloc((void **)&t->seed_q, cap * 4));
    STRY(ctx, hipMalloc((void **)&t->seed_p, cap * 8));
    t->seed_cap = cap;
  }
  // unmark the previous rows
  if (!t->marked_rows.empty()) {
    STRY(ctx, hipMemcpyAsync(t->seed_row, t->marked_rows.data(), t->marked_rows.size() * 4,
                             hipMemcpyHostToDevice, ctx->stream));
    STRY(ctx, ppr_launch_seed_mark(ctx->stream, t->seed_idx, t->seed_row,
                                   (int)t->marked_rows.size(), 0));
    STRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  std::vector<int32_t> rows(cnt), qs(cnt);
  std::vector<double> ps(cnt);
  for (int j = 0; j < cnt; j++) {
    rows[j] = ent[j].row;
    qs[j] = ent[j].q;
    ps[j] = ent[j].p;
  }
  if (cnt) {
    STRY(ctx, hipMemcpyAsync(t->seed_row, rows.data(), cnt * 4, hipMemcpyHostToDevice, ctx->stream));
    STRY(ctx, hipMemcpyAsync(t->seed_q, qs.data(), cnt * 4, hipMemcpyHostToDevice, ctx->stream));
    STRY(ctx, hipMemcpyAsync(t->seed_p, ps.data(), cnt * 8, hipMemcpyHostToDevice, ctx->stream));
    STRY(ctx, ppr_launch_seed_mark(ctx->stream, t->seed_idx, t->seed_row, cnt, 1));
    STRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  t->num_seed = cnt;
  t->zseed_first = cnt;
  const int32_t n_in = g_n_in;
  for (int j = 0; j < cnt; j++)
    if (rows[j] >= n_in) {
      t->zseed_first = j;
      break;
    }
  t->marked_rows = rows;
  return GOPS_OK;
}

void set_slot_params(Tile *t, int s, const Query &q, double f32_below, uint64_t period) {
  TileState *h = t->hts;
  h->damping[s] = q.cfg.damping;
  h->tol[s] = q.cfg.tolerance;
  h->max_iter[s] = q.cfg.max_iterations;
  h->init_dsum[s] = q.init_dsum;
  h->init_mass[s] = q.init_mass;
  h->dsum_in[s] = 0.0;
  h->mass_in[s] = 0.0;
  h->diff_anchor[s] = 0.0;
  h->anchor_drop = 0.0;  // no residual-triggered anchors: the carried f32 rounding does not
                         // build up (ppr_epilogue)
  h->f32_below = f32_below;
  h->anchor_min = 1e3;
  h->anchor_period = period;
}

// Admit the queries in `mask` (slot mask) into tile t: upload the tile's
// seed entries, clear the admitted columns and start the new queries; the
// running queries keep their state.
gops_status admit(gops_context *ctx, gops_graph *g, Tile *t, const TilePlan *p, uint64_t mask) {
  const int32_t n_in = g->ppr_cache ? g->ppr_cache->n_in : (int32_t)g->dev.n;
  gops_status st = upload_seeds(ctx, t, n_in);
  if (st != GOPS_OK) return st;
  // rows without in-edges to clear in the admitted columns: the departing
  // queries' seed rows there; then the new queries' take their place
  t->clear_host.clear();
  for (int s = 0; s < t->T; s++) {
    if (!((mask >> s) & 1ull)) continue;
    t->clear_host.insert(t->clear_host.end(), t->zrows[s].begin(), t->zrows[s].end());
    t->zrows[s].clear();
    for (auto &pv : t->slot[s].pers)
      if ((int32_t)pv.first >= n_in) t->zrows[s].push_back((int32_t)pv.first);
  }
  std::sort(t->clear_host.begin(), t->clear_host.end());
  t->clear_host.erase(std::unique(t->clear_host.begin(), t->clear_host.end()), t->clear_host.end());
  const int nz = (int)t->clear_host.size();
  if (nz > t->clear_cap) {
    hipFree(t->clear_rows);
    t->clear_rows = nullptr;
    t->clear_cap = 0;
    const int cap = std::max(nz, 1024);
    STRY(ctx, hipMalloc((void **)&t->clear_rows, (size_t)cap * 4));
    t->clear_cap = cap;
  }
  if (nz)
    STRY(ctx, hipMemcpyAsync(t->clear_rows, t->clear_host.data(), (size_t)nz * 4,
                             hipMemcpyHostToDevice, ctx->stream));
  t->hts->all_done = 0;
  STRY(ctx, hipMemcpyAsync(t->ts, t->hts, sizeof(TileState), hipMemcpyHostToDevice, ctx->stream));
  PprArgs a = tile_args(ctx, g, t, p, false);
  STRY(ctx, ppr_launch_admit(ctx->stream, t->T, a, mask, mask, t->clear_rows, nz));
  return GOPS_OK;
}

gops_status sync_state(gops_context *ctx, Tile *t) {
  STRY(ctx, hipMemcpyAsync(t->hts, t->ts, sizeof(TileState), hipMemcpyDeviceToHost, ctx->stream));
  STRY(ctx, hipStreamSynchronize(ctx->stream));
  return GOPS_OK;
}

// Workgroups of a tile sweep (grid-stride over the descriptors): 8 waves of
// 64 lanes on every SIMD of 256 CUs.
int tile_grid(const gops_graph *g) {
  (void)g;
  return kMaxSweepBlocks;
}

// Time per query-sweep of tile width T on this graph: T single-seed trial
// queries (tolerance 0, four sweeps in the context's tile schedule), timed
// with events.  With a leading f64 phase the first sweep is an f64 anchor and
// the estimate weights the two sweep kinds equally, as such a schedule spends
// about half of its sweeps in each; otherwise all four are f32 delta sweeps.
// Infinity if the tile does not fit in device memory.
gops_status trial_tile(gops_context *ctx, gops_graph *g, PprGraphCache *c, int T, double *cost) {
  *cost = std::numeric_limits<double>::infinity();
  TilePlan *plan = nullptr;
  gops_status st = tile_plan(ctx, g, c, T, &plan);
  if (st != GOPS_OK) return st;
  Tile *t = nullptr;
  st = alloc_tile(ctx, g, T, tile_grid(g), &t);
  if (st == GOPS_OUT_OF_MEMORY) {
    ctx->last_error.clear();
    (void)hipGetLastError();  // the failed allocation's error is not a later launch's
    return GOPS_OK;
  }
  if (st != GOPS_OK) return st;
  const uint64_t n = g->dev.n;
  std::memset(t->hts, 0, sizeof(TileState));
  uint64_t mask = 0;
  for (int s = 0; s < T; s++) {
    Query q;
    q.cfg = gops_pagerank_config{0.85, 4, 0.0};
    const uint32_t r = (uint32_t)(((uint64_t)s * 2654435761ull + 12345) % n);
    q.pers.push_back({r, 1.0});
    q.init_dsum = c->dangling[r] ? 1.0 : 0.0;
    q.init_mass = 1.0;
    t->slot[s] = q;
    t->used[s] = true;
    set_slot_params(t, s, q, ctx->tile_f32_below, 64);
    mask |= 1ull << s;
  }
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  hipError_t e = hipSuccess;
  for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
  if (e == hipSuccess) {
    st = admit(ctx, g, t, plan, mask);
    if (st == GOPS_OK) {
      PprArgs a = tile_args(ctx, g, t, plan, false);
      e = hipEventRecord(ev[0], ctx->stream);
      if (e == hipSuccess) e = ppr_launch_sweep(ctx->stream, T, a, t->grid);  // admission sweep
      if (e == hipSuccess) e = hipEventRecord(ev[1], ctx->stream);
      for (int k = 0; k < 3 && e == hipSuccess; k++) e = ppr_launch_sweep(ctx->stream, T, a, t->grid);
      if (e == hipSuccess) e = hipEventRecord(ev[2], ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
      float ms_anchor = 0.f, ms_delta = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms_anchor, ev[0], ev[1]);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms_delta, ev[1], ev[2]);
      if (e == hipSuccess)
        *cost = (ctx->tile_f32_below > 0.0 ? 0.5 * ((double)ms_anchor + (double)ms_delta / 3.0)
                                           : ((double)ms_anchor + (double)ms_delta) / 4.0) / T;
    }
  }
  for (hipEvent_t x : ev)
    if (x) hipEventDestroy(x);
  free_tile(t);
  if (st != GOPS_OK) return st;
  if (e != hipSuccess) return sfail(ctx, GOPS_DEVICE_ERROR, std::string("tile trial: ") + hipGetErrorString(e));
  return GOPS_OK;
}

// Tile width of a batch of nq queries.  A fixed width (GOPS_OPT_PPR_TILE)
// wins; otherwise the autotuned width of this graph (the least time per
// query-sweep over 8/16/32/64, measured on first use and kept with the
// graph), narrowed to the smallest width that holds nq queries.
gops_status pick_tile(gops_context *ctx, gops_graph *g, PprGraphCache *c, int nq, int *T) {
  if (ctx->ppr_tile > 0) {
    *T = ctx->ppr_tile;
    return GOPS_OK;
  }
  if (!c->tuned_T) {
    double best = std::numeric_limits<double>::infinity();
    int best_T = 0;
    for (int w : {8, 16, 32, 64}) {
      double cost;
      gops_status st = trial_tile(ctx, g, c, w, &cost);
      if (st != GOPS_OK) return st;
      c->tile_cost[plan_slot(w)] = cost;
      if (cost < best) {
        best = cost;
        best_T = w;
      }
    }
    if (!best_T) return sfail(ctx, GOPS_OUT_OF_MEMORY, "no tile width fits in device memory");
    c->tuned_T = best_T;
  }
  int w = 8;
  while (w < nq && w < c->tuned_T) w <<= 1;
  *T = w;
  return GOPS_OK;
}

}  // namespace

namespace gops {
void free_ppr_cache(gops_graph *g) {
  PprGraphCache *c = g->ppr_cache;
  if (!c) return;
  for (TilePlan &p : c->plans) hipFree(p.desc);
  delete c;
  g->ppr_cache = nullptr;
}
}  // namespace gops

extern "C" {

gops_status gops_ppr_batch_run(gops_context *ctx, gops_graph *g, uint32_t nq,
                               const gops_pagerank_config *cfgs, const uint64_t *seed_off,
                               const uint32_t *seeds, const double *seed_w, double *scores,
                               gops_run_info *infos) {
  if (!ctx || !g) return GOPS_INVALID_PARAMETER;
  if (nq && (!cfgs || !seed_off)) return sfail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return sfail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  ctx->last_error.clear();
  if (ctx->ppr_tile < 0)  // plain path: one query per device run
    return ppr_batch_looped(ctx, g, nq, cfgs, seed_off, seeds, seed_w, scores, infos);
  const uint64_t n = g->dev.n;
  if (nq == 0) return GOPS_OK;
  if (n == 0) {
    for (uint32_t q = 0; q < nq; q++)
      if (infos) infos[q] = gops_run_info{0, 0.0, 1, 0};
    return GOPS_OK;
  }
  PprGraphCache *c = nullptr;
  gops_status st = graph_cache(ctx, g, &c);
  if (st != GOPS_OK) return st;
  int T = 0;
  st = pick_tile(ctx, g, c, (int)std::min<uint32_t>(nq, 1u << 20), &T);
  if (st != GOPS_OK) return st;
  TilePlan *plan = nullptr;
  st = tile_plan(ctx, g, c, T, &plan);
  if (st != GOPS_OK) return st;
  Tile *t = nullptr;
  st = alloc_tile(ctx, g, T, tile_grid(g), &t);
  if (st != GOPS_OK) return st;
  double *dout = nullptr;  // one query's scores in the caller's ids
  if (scores && hipMalloc((void **)&dout, n * sizeof(double)) != hipSuccess) {
    free_tile(t);
    return sfail(ctx, GOPS_OUT_OF_MEMORY, "batch output buffer: out of device memory");
  }
  // Continuous refill by groups of kLaneQ slots (the queries one lane
  // gathers for): a group takes the next queries once all of its queries
  // have stopped, so a lane's queries start together and mostly sweep in the
  // same precision.  The sweeps run until the tile is done or, while queries
  // wait, until refill_groups groups are free; then the finished queries are
  // collected and the free groups refilled, while the others keep their state.
  std::vector<uint32_t> dense_q;  // queries that fall back to the dense uniform case
  std::vector<uint32_t> slot_q(T, 0);
  std::memset(t->hts, 0, sizeof(TileState));
  t->hts->all_done = 1;
  for (int s = 0; s < T; s++) t->used[s] = false;
  uint32_t next = 0;
  const int refill_groups = std::max(1, T / kLaneQ / GOPS_TUNE_TILE_REFILL);
  const uint64_t kAhead = GOPS_TUNE_TILE_AHEAD;  // sweeps queued beyond the last one seen to finish
  while (st == GOPS_OK) {
    // 1. admission into free slots (the host copy of the tile state is current)
    uint64_t mask = 0;
    for (int s = 0; s < T && next < nq && st == GOPS_OK; s++) {
      if (t->used[s]) continue;
      const int g0 = s - s % kLaneQ;  // only groups with no query left take new ones
      bool group_free = true;
      for (int i = g0; i < g0 + kLaneQ; i++) group_free = group_free && (!t->used[i] || ((mask >> i) & 1));
      if (!group_free) continue;
      while (next < nq) {
        const uint32_t q = next++;
        Query qq;
        st = make_query(ctx, g, c, &cfgs[q], seeds + seed_off[q], seed_off[q + 1] - seed_off[q],
                        seed_w ? seed_w + seed_off[q] : nullptr, qq);
        if (st != GOPS_OK) break;
        if (qq.dense_fallback || cfgs[q].max_iterations == 0) {
          dense_q.push_back(q);
          continue;
        }
        t->slot[s] = qq;
        t->used[s] = true;
        slot_q[s] = q;
        set_slot_params(t, s, qq, ctx->tile_f32_below, (uint64_t)std::max(2, ctx->ppr_anchor_period));
        mask |= 1ull << s;
        break;
      }
    }
    if (st != GOPS_OK) break;
    uint64_t cap = 0;  // no query runs more sweeps than its cap from now on
    for (int s = 0; s < T; s++)
      if (t->used[s]) cap = std::max<uint64_t>(cap, t->slot[s].cfg.max_iterations);
    if (cap == 0) break;  // no query left
    if (mask) {
      st = admit(ctx, g, t, plan, mask);
      if (st != GOPS_OK) break;
    }
    // (after admission: it may move the seed arrays)
    const PprArgs a = tile_args(ctx, g, t, plan, true);
    // 2. sweeps, a few queued ahead of the last one seen to finish
    const uint64_t base = t->hts->sweep;
    const bool waiting = next < nq;
    __atomic_store_n(ctx->mirror_host, base << 8, __ATOMIC_RELEASE);
    uint64_t j = 0, seen = 0;
    while (j < cap && st == GOPS_OK) {
      if (j < cap && j < seen + kAhead) {
        uint32_t queued = 0;
        const hipError_t e = queue_tile_sweeps(
            ctx, t, a, (uint32_t)std::min<uint64_t>(cap - j, seen + kAhead - j), &queued);
        j += queued;
        if (e != hipSuccess) {
          st = sfail(ctx, GOPS_DEVICE_ERROR, hipGetErrorString(e));
          break;
        }
      }
      if (j >= cap || st != GOPS_OK) break;
      bool stop = false;
      for (uint64_t spin = 0;; spin++) {
        const uint64_t w = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
        if (w & 1) {
          stop = true;
          break;
        }
        if ((w >> 8) - base > seen) {
          seen = (w >> 8) - base;
          const int free_groups = (int)((w >> 1) & 127);
          stop = waiting && free_groups >= refill_groups;
          break;
        }
        if ((spin & 1023) == 1023) {
          const hipError_t qe = hipStreamQuery(ctx->stream);
          if (qe == hipSuccess) {
            const uint64_t w2 = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
            stop = (w2 & 1) != 0;
            if (!stop && (w2 >> 8) - base <= seen)
              st = sfail(ctx, GOPS_DEVICE_ERROR, "tile sweep chain stalled without progress");
            seen = (w2 >> 8) - base;
            break;
          }
          if (qe != hipErrorNotReady) {
            st = sfail(ctx, GOPS_DEVICE_ERROR, hipGetErrorString(qe));
            break;
          }
        }
        __builtin_ia32_pause();
      }
      if (stop) break;
    }
    if (st != GOPS_OK) break;
    st = sync_state(ctx, t);  // waits for the queued sweeps
    if (st != GOPS_OK) break;
    // 3. collect the queries that have stopped
    for (int s = 0; s < T && st == GOPS_OK; s++) {
      if (!t->used[s] || t->hts->live[s]) continue;
      const uint32_t q = slot_q[s];
      if (infos) {
        infos[q].iterations = t->hts->iter[s];
        infos[q].diff_l1 = t->hts->diff[s];
        infos[q].converged = t->hts->converged[s];
        infos[q].reserved = 0;
      }
      if (scores) {
        hipError_t e = ppr_launch_gather(ctx->stream, T, a, g->dev.rank, (int)n, s, dout);
        if (e == hipSuccess)
          e = staged_download(ctx, scores + (size_t)q * n, dout, n * sizeof(double), ctx->stream);
        if (e != hipSuccess) st = sfail(ctx, GOPS_DEVICE_ERROR, hipGetErrorString(e));
      }
      t->used[s] = false;
    }
  }
  free_tile(t);
  hipFree(dout);
  if (st != GOPS_OK) return st;
  // queries whose personalization is not sparse-positive: the dense path
  for (uint32_t q : dense_q) {
    std::vector<double> p(n, 0.0);
    for (uint64_t j = seed_off[q]; j < seed_off[q + 1]; j++) p[seeds[j]] += seed_w ? seed_w[j] : 1.0;
    gops_run_info info{};
    st = run_single(ctx, g, &cfgs[q], kPersonalized, p.data(), n,
                    scores ? scores + (size_t)q * n : nullptr, &info);
    if (st != GOPS_OK) return st;
    if (infos) infos[q] = info;
  }
  return GOPS_OK;
}

gops_status gops_ppr_tile_info(gops_context *ctx, gops_graph *g, uint32_t nq, uint32_t *tile,
                               double *ms_per_query_sweep) {
  if (!ctx || !g) return GOPS_INVALID_PARAMETER;
  if (g->ctx != ctx) return sfail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  if (g->dev.n == 0) return sfail(ctx, GOPS_INVALID_PARAMETER, "empty graph");
  ctx->last_error.clear();
  PprGraphCache *c = nullptr;
  gops_status st = graph_cache(ctx, g, &c);
  if (st != GOPS_OK) return st;
  // the trial runs whatever the options say, so the costs are always there
  const int saved = ctx->ppr_tile;
  ctx->ppr_tile = 0;
  int T = 0;
  st = pick_tile(ctx, g, c, (int)std::min<uint32_t>(nq, 1u << 20), &T);
  ctx->ppr_tile = saved;
  if (st != GOPS_OK) return st;
  if (saved > 0) T = saved;
  if (saved < 0) T = 0;
  if (tile) *tile = (uint32_t)T;
  if (ms_per_query_sweep)
    for (int i = 0; i < 4; i++) ms_per_query_sweep[i] = c->tile_cost[i];
  return GOPS_OK;
}

// ---------------------------------------------------------------------------
gops_status gops_server_create(gops_context *ctx, gops_graph *g, const gops_server_config *cfg,
                               gops_server **out) {
  if (!ctx || !g || !out) return sfail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  *out = nullptr;
  if (g->ctx != ctx) return sfail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  if (g->dev.n == 0) return sfail(ctx, GOPS_INVALID_PARAMETER, "empty graph");
  PprGraphCache *c = nullptr;
  gops_status st = graph_cache(ctx, g, &c);
  if (st != GOPS_OK) return st;
  int T = cfg ? (int)cfg->tile : 0;
  if (T == 0) {
    const int saved = ctx->ppr_tile;
    ctx->ppr_tile = 0;
    st = pick_tile(ctx, g, c, 64, &T);
    ctx->ppr_tile = saved;
    if (st != GOPS_OK) return st;
  }
  if (!(T == 8 || T == 16 || T == 32 || T == 64))
    return sfail(ctx, GOPS_INVALID_PARAMETER, "tile must be 8, 16, 32 or 64");
  // admission by HBM: as many tiles as the budget holds
  size_t free_b = 0, total_b = 0;
  STRY(ctx, hipMemGetInfo(&free_b, &total_b));
  double budget = (cfg && cfg->hbm_budget_bytes > 0) ? cfg->hbm_budget_bytes : 0.9 * (double)free_b;
  const size_t per_tile = tile_bytes(g->dev.n, T);
  uint32_t ntiles = (uint32_t)std::floor(budget / (double)per_tile);
  if (cfg && cfg->max_slots) ntiles = std::min<uint32_t>(ntiles, (cfg->max_slots + T - 1) / T);
  if (ntiles == 0) return sfail(ctx, GOPS_OUT_OF_MEMORY, "HBM budget holds no query tile");
  TilePlan *plan = nullptr;
  st = tile_plan(ctx, g, c, T, &plan);
  if (st != GOPS_OK) return st;
  auto *s = new (std::nothrow) gops_server;
  if (!s) return sfail(ctx, GOPS_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->g = g;
  s->T = T;
  s->topk = cfg ? cfg->topk : 0;
  s->t0 = std::chrono::steady_clock::now();
  for (uint32_t i = 0; i < ntiles; i++) {
    Tile *t = nullptr;
    st = alloc_tile(ctx, g, T, tile_grid(g), &t);
    if (st != GOPS_OK) {
      if (i == 0) {
        delete s;
        return st;
      }
      ctx->last_error.clear();
      (void)hipGetLastError();  // the failed allocation's error is not a later launch's
      break;  // fewer tiles than the budget suggested
    }
    s->tiles.push_back(t);
  }
  g->refs.fetch_add(1);  // the server keeps its graph (and so the context) until destroyed
  s->holds_graph = true;
  *out = s;
  return GOPS_OK;
}

void gops_server_destroy(gops_server *s) {
  if (!s) return;
  hipStreamSynchronize(s->ctx->stream);
  for (Tile *t : s->tiles) free_tile(t);
  for (auto &d : s->finished) hipFree(d.scores);
  hipFree(s->gather);
  hipFree(s->top_ids);
  hipFree(s->top_vals);
  hipFree(s->top_cnt);
  gops_graph *g = s->holds_graph ? s->g : nullptr;
  delete s;
  release_graph(g);
}

gops_status gops_server_submit(gops_server *s, const gops_pagerank_config *cfg,
                               const uint32_t *seeds, uint64_t ns, const double *w,
                               uint64_t *ticket) {
  if (!s || !cfg || (ns && !seeds)) return GOPS_INVALID_PARAMETER;
  PprGraphCache *c = nullptr;
  gops_status st = graph_cache(s->ctx, s->g, &c);
  if (st != GOPS_OK) return st;
  Query q;
  st = make_query(s->ctx, s->g, c, cfg, seeds, ns, w, q);
  if (st != GOPS_OK) return st;
  if (q.dense_fallback || cfg->max_iterations == 0)
    return sfail(s->ctx, GOPS_NOT_SUPPORTED,
                 "serving needs a personalization with positive sum and max_iterations > 0");
  q.ticket = s->next_ticket++;
  q.submit_ms = s->now_ms();
  s->queue.push_back(std::move(q));
  if (ticket) *ticket = s->queue.back().ticket;
  return GOPS_OK;
}

gops_status gops_server_step(gops_server *s, uint32_t sweeps, uint32_t *active) {
  if (!s) return GOPS_INVALID_PARAMETER;
  gops_context *ctx = s->ctx;
  gops_graph *g = s->g;
  PprGraphCache *c = nullptr;
  gops_status st = graph_cache(ctx, g, &c);
  if (st != GOPS_OK) return st;
  TilePlan *plan = nullptr;
  st = tile_plan(ctx, g, c, s->T, &plan);
  if (st != GOPS_OK) return st;
  const int n = (int)g->dev.n;
  // 1. admission (host state is current: no kernel is queued).  With the
  // default all-f32 schedule every free slot takes the queue head.  With a
  // leading f64 phase (anchored admissions), by groups of kLaneQ slots, the
  // queries one lane gathers for: a group takes new queries once none of its
  // own is left, so a lane's queries start together and sweep in the same
  // precision.  The first queued query opens the group; its other slots take
  // the next queued queries of the same damping (which sets the iteration
  // count most) found within kMatchWindow entries, else the next in order.
  // The queue head is always admitted first.
  constexpr size_t kMatchWindow = 64;
  const bool by_group = ctx->tile_f32_below > 0.0;
  for (Tile *t : s->tiles) {
    if (s->queue.empty()) break;
    uint64_t mask = 0;
    for (int q = 0; q < t->T && !by_group && !s->queue.empty(); q++) {
      if (t->used[q]) continue;
      t->slot[q] = std::move(s->queue.front());
      s->queue.pop_front();
      t->used[q] = true;
      set_slot_params(t, q, t->slot[q], ctx->tile_f32_below, (uint64_t)std::max(2, ctx->ppr_anchor_period));
      mask |= 1ull << q;
    }
    for (int g0 = 0; g0 < t->T && by_group && !s->queue.empty(); g0 += kLaneQ) {
      bool group_free = true;
      for (int q = g0; q < g0 + kLaneQ; q++) group_free = group_free && !t->used[q];
      if (!group_free) continue;
      const double damping = s->queue.front().cfg.damping;
      for (int q = g0; q < g0 + kLaneQ && !s->queue.empty(); q++) {
        size_t pick = 0;
        const size_t window = std::min(s->queue.size(), kMatchWindow);
        for (size_t i = 0; i < window; i++) {
          if (s->queue[i].cfg.damping == damping) {
            pick = i;
            break;
          }
        }
        t->slot[q] = std::move(s->queue[pick]);
        s->queue.erase(s->queue.begin() + (std::ptrdiff_t)pick);
        t->used[q] = true;
        set_slot_params(t, q, t->slot[q], ctx->tile_f32_below, (uint64_t)std::max(2, ctx->ppr_anchor_period));
        mask |= 1ull << q;
      }
    }
    if (mask) {
      st = admit(ctx, g, t, plan, mask);
      if (st != GOPS_OK) return st;
      st = sync_state(ctx, t);
      if (st != GOPS_OK) return st;
    }
  }
  // 2. sweeps of every tile with live queries, queued back to back
  for (Tile *t : s->tiles) {
    if (t->hts->all_done) continue;
    PprArgs a = tile_args(ctx, g, t, plan, false);
    uint32_t queued = 0;
    STRY(ctx, queue_tile_sweeps(ctx, t, a, sweeps, &queued));
  }
  STRY(ctx, hipStreamSynchronize(ctx->stream));
  // 3. harvest finished queries.  With topk: a tile's finished queries are
  // gathered in one pass over its rows and their top_k selected together
  // (every selection kernel launched once for all of them); all of this
  // step's pairs are copied back once at the end.
  uint32_t live_total = 0;
  const size_t first_new = s->finished.size();
  uint32_t n_top = 0;  // results selected in this harvest
  auto ensure_top = [&](uint32_t need) -> gops_status {
    if (need <= s->top_cap) return GOPS_OK;
    const uint32_t cap = std::max<uint32_t>(std::max<uint32_t>(64, 2 * s->top_cap), need);
    const size_t k = s->topk;
    uint32_t *ids = nullptr, *cnt = nullptr;
    double *vals = nullptr;
    STRY(ctx, hipMalloc((void **)&ids, (size_t)cap * k * sizeof(uint32_t)));
    STRY(ctx, hipMalloc((void **)&vals, (size_t)cap * k * sizeof(double)));
    STRY(ctx, hipMalloc((void **)&cnt, (size_t)cap * sizeof(uint32_t)));
    if (n_top) {  // this harvest's earlier selections
      STRY(ctx, hipMemcpyAsync(ids, s->top_ids, (size_t)n_top * k * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, ctx->stream));
      STRY(ctx, hipMemcpyAsync(vals, s->top_vals, (size_t)n_top * k * sizeof(double),
                               hipMemcpyDeviceToDevice, ctx->stream));
      STRY(ctx, hipMemcpyAsync(cnt, s->top_cnt, (size_t)n_top * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    STRY(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(s->top_ids);
    hipFree(s->top_vals);
    hipFree(s->top_cnt);
    s->top_ids = ids;
    s->top_vals = vals;
    s->top_cnt = cnt;
    s->top_cap = cap;
    return GOPS_OK;
  };
  const bool batched_topk = s->topk > 0 && s->topk <= (uint32_t)kNoAllocTopK;
  for (Tile *t : s->tiles) {
    bool any = false;
    for (int q = 0; q < t->T; q++) any = any || t->used[q];
    if (!any) continue;  // an idle tile's state has not changed
    st = sync_state(ctx, t);
    if (st != GOPS_OK) return st;
    PprArgs a = tile_args(ctx, g, t, plan, false);
    uint64_t done_mask = 0;
    for (int q = 0; q < t->T; q++) {
      if (!t->used[q]) continue;
      if (t->hts->live[q]) {
        live_total++;
        continue;
      }
      done_mask |= 1ull << q;
    }
    if (!done_mask) continue;
    const uint32_t F = (uint32_t)__builtin_popcountll(done_mask);
    if (batched_topk) {
      if (s->gather_cap < F && s->gather_cap < (uint32_t)t->T) {
        // room for a tile's worth of columns, or as many as fit (at least one)
        hipFree(s->gather);
        s->gather = nullptr;
        s->gather_cap = 0;
        for (uint32_t c = (uint32_t)t->T; c >= 1 && !s->gather; c /= 2) {
          if (hipMalloc((void **)&s->gather, (size_t)c * n * sizeof(double)) == hipSuccess)
            s->gather_cap = c;
          else
            (void)hipGetLastError();  // a failed allocation is not a later launch's error
        }
        if (!s->gather) return sfail(ctx, GOPS_OUT_OF_MEMORY, "harvest buffer: out of device memory");
      }
      st = ensure_top(n_top + F);
      if (st != GOPS_OK) return st;
      // groups of at most gather_cap queries, in slot order
      uint64_t rest = done_mask;
      uint32_t at = n_top;
      while (rest) {
        uint64_t grp = 0;
        for (uint32_t c = 0; c < s->gather_cap && rest; c++) {
          grp |= rest & (~rest + 1);
          rest &= rest - 1;
        }
        const uint32_t fg = (uint32_t)__builtin_popcountll(grp);
        STRY(ctx, ppr_launch_gather_mask(ctx->stream, t->T, a, g->dev.rank, n, grp, s->gather));
        const size_t off = (size_t)at * s->topk;
        STRY(ctx, launch_top_k_batch(ctx, s->gather, n, fg, s->topk, s->top_ids + off,
                                     s->top_vals + off, s->top_cnt + at));
        at += fg;
      }
    }
    for (int q = 0; q < t->T; q++) {
      if (!((done_mask >> q) & 1ull)) continue;
      gops_server::Done d;
      d.q = std::move(t->slot[q]);
      d.info = gops_run_info{t->hts->iter[q], t->hts->diff[q], t->hts->converged[q], 0};
      d.finish_ms = s->now_ms();
      d.scores = nullptr;
      if (batched_topk) {
        n_top++;  // selected above, in slot order
      } else if (s->topk) {
        // k beyond the one-workgroup sort: one selection per query, the
        // gather buffer reused in stream order
        if (!s->gather) {
          STRY(ctx, hipMalloc((void **)&s->gather, (size_t)n * sizeof(double)));
          s->gather_cap = 1;
        }
        st = ensure_top(n_top + 1);
        if (st != GOPS_OK) return st;
        STRY(ctx, ppr_launch_gather(ctx->stream, t->T, a, g->dev.rank, n, q, s->gather));
        const size_t off = (size_t)n_top * s->topk;
        STRY(ctx, launch_top_k(ctx, s->gather, n, s->topk, s->top_ids + off, s->top_vals + off,
                               s->top_cnt + n_top));
        n_top++;
      } else {
        STRY(ctx, hipMalloc((void **)&d.scores, (size_t)n * sizeof(double)));
        STRY(ctx, ppr_launch_gather(ctx->stream, t->T, a, g->dev.rank, n, q, d.scores));
      }
      t->used[q] = false;
      s->finished.push_back(std::move(d));
      s->n_finished++;
    }
  }
  if (n_top) {
    const size_t k = s->topk;
    std::vector<uint32_t> ids((size_t)n_top * k), cnt(n_top);
    std::vector<double> vals((size_t)n_top * k);
    STRY(ctx, hipMemcpyAsync(ids.data(), s->top_ids, ids.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, ctx->stream));
    STRY(ctx, hipMemcpyAsync(vals.data(), s->top_vals, vals.size() * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
    STRY(ctx, hipMemcpyAsync(cnt.data(), s->top_cnt, cnt.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, ctx->stream));
    STRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < n_top; i++) {
      gops_server::Done &d = s->finished[first_new + i];
      d.top_ids.assign(ids.begin() + (size_t)i * k, ids.begin() + (size_t)i * k + cnt[i]);
      d.top_vals.assign(vals.begin() + (size_t)i * k, vals.begin() + (size_t)i * k + cnt[i]);
    }
  }
  STRY(ctx, hipStreamSynchronize(ctx->stream));
  if (active) *active = live_total;
  return GOPS_OK;
}

gops_status gops_server_poll(gops_server *s, int *ready, gops_server_result *out, double *scores,
                             uint32_t *topk_ids, double *topk_scores) {
  if (!s || !ready) return GOPS_INVALID_PARAMETER;
  *ready = 0;
  if (s->topk && scores)
    return sfail(s->ctx, GOPS_INVALID_PARAMETER, "a top-k server returns no score vectors");
  if (s->finished.empty()) return GOPS_OK;
  gops_server::Done d = std::move(s->finished.front());
  s->finished.pop_front();
  const size_t n = s->g->dev.n;
  gops_status st = GOPS_OK;
  if (scores && d.scores) {
    hipError_t e = staged_download(s->ctx, scores, d.scores, n * sizeof(double), s->ctx->stream);
    if (e != hipSuccess) st = sfail(s->ctx, GOPS_DEVICE_ERROR, hipGetErrorString(e));
  }
  hipFree(d.scores);
  if (st != GOPS_OK) return st;
  if (s->topk) {
    // top_k semantics of topk.rs:17-42, from the device selection
    for (size_t i = 0; i < s->topk; i++) {
      const bool have = i < d.top_ids.size();
      if (topk_ids) topk_ids[i] = have ? d.top_ids[i] : UINT32_MAX;
      if (topk_scores) topk_scores[i] = have ? d.top_vals[i] : 0.0;
    }
  }
  if (out) {
    out->ticket = d.q.ticket;
    out->info = d.info;
    out->submit_ms = d.q.submit_ms;
    out->finish_ms = d.finish_ms;
  }
  *ready = 1;
  return GOPS_OK;
}

gops_status gops_server_stats(const gops_server *s, uint64_t *queued, uint64_t *active,
                              uint64_t *finished, uint32_t *slots) {
  if (!s) return GOPS_INVALID_PARAMETER;
  uint64_t act = 0;
  for (Tile *t : s->tiles)
    for (int q = 0; q < t->T; q++) act += t->used[q];
  if (queued) *queued = s->queue.size();
  if (active) *active = act;
  if (finished) *finished = s->n_finished;
  if (slots) *slots = (uint32_t)(s->tiles.size() * (size_t)s->T);
  return GOPS_OK;
}

}  // extern "C"
