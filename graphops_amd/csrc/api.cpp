// api.cpp -- extern "C" entry points of libgraphops_amd (see
// include/graphops_amd.h for the reference functions each one replaces).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>

#include "internal.h"

// the Python mirror (ctypes) and bindings in other languages lay it out field by field
static_assert(sizeof(gops_run_stats) == 88, "gops_run_stats layout");

using namespace gops;

namespace {

gops_status fail(gops_context *ctx, gops_status st, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return st;
}

gops_status invalid(gops_context *ctx, const std::string &msg) {
  // Display text of Error::InvalidParameter, lib.rs:117-118
  return fail(ctx, GOPS_INVALID_PARAMETER, "invalid parameter: " + msg);
}

gops_status hip_fail(gops_context *ctx, hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(ctx, GOPS_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
  return fail(ctx, GOPS_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                 \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}

template <typename T>
hipError_t upload(gops_context *ctx, T **p, const T *src, size_t count) {
  hipError_t e = dalloc(p, count);
  if (e != hipSuccess) return e;
  if (count == 0) return hipSuccess;
  return staged_upload(ctx, *p, src, count * sizeof(T), ctx->stream);
}

// PageRankConfig::validate, pagerank.rs:63-85 (messages verbatim)
gops_status validate_config(gops_context *ctx, const gops_pagerank_config *cfg) {
  if (!std::isfinite(cfg->damping)) return invalid(ctx, "damping must be finite");
  if (!(cfg->damping >= 0.0 && cfg->damping <= 1.0)) return invalid(ctx, "damping must be in [0,1]");
  if (cfg->max_iterations == 0) return invalid(ctx, "max_iterations must be > 0");
  if (!std::isfinite(cfg->tolerance) || cfg->tolerance <= 0.0)
    return invalid(ctx, "tolerance must be finite and > 0");
  return GOPS_OK;
}

void free_run(RunBuffers &rb) {
  for (hipGraphExec_t &e : rb.chain)
    if (e) (void)hipGraphExecDestroy(e);
  hipFree(rb.xbar);
  hipFree(rb.yd[0]);
  hipFree(rb.yd[1]);
  hipFree(rb.y64[0]);
  hipFree(rb.y64[1]);
  hipFree(rb.q32);
  hipFree(rb.q64);
  hipFree(rb.out);
  hipFree(rb.partials);
  hipFree(rb.hub_partial);
  hipFree(rb.state);
  hipFree(rb.sweep_log);
  rb = RunBuffers{};
}

void free_power(PowerBuffers &pb) {
  for (double *v : pb.v) hipFree(v);
  hipFree(pb.partials);
  hipFree(pb.hub_partial);
  hipFree(pb.state);
  pb = PowerBuffers{};
}

void free_dev(DevGraph &dg) {
  hipFree(dg.row_ptr);
  hipFree(dg.col);
  hipFree(dg.w);
  hipFree(dg.inv_deg32);
  hipFree(dg.inv_deg64);
  hipFree(dg.inv_ws32);
  hipFree(dg.inv_ws64);
  hipFree(dg.rank);
  hipFree(dg.order);
  hipFree(dg.desc);
  hipFree(dg.hub_row);
  hipFree(dg.hub_nchunks);
  hipFree(dg.hub_first);
  dg = DevGraph{};
}

gops_status alloc_run(gops_context *ctx, gops_graph *g) {
  RunBuffers &rb = g->run;
  if (rb.state) return GOPS_OK;
  const size_t n = g->dev.n;
  rb.n = n;
  HIP_TRY(ctx, dalloc(&rb.xbar, n));
  HIP_TRY(ctx, dalloc(&rb.yd[0], n));
  HIP_TRY(ctx, dalloc(&rb.yd[1], n));
  HIP_TRY(ctx, dalloc(&rb.y64[0], n));
  HIP_TRY(ctx, dalloc(&rb.y64[1], n));
  HIP_TRY(ctx, dalloc(&rb.out, n));
  // sweep grid <= kMaxSweepBlocks, elementwise grids <= 2048 workgroups
  rb.partial_slots = std::max<size_t>(kMaxSweepBlocks, 2048);
  HIP_TRY(ctx, dalloc(&rb.partials, rb.partial_slots * kMaxReduce));
  HIP_TRY(ctx, dalloc(&rb.hub_partial, std::max(g->dev.num_hub_chunks, 1)));
  HIP_TRY(ctx, dalloc(&rb.state, 1));
  HIP_TRY(ctx, dalloc(&rb.sweep_log, kSweepLogCap));
  HIP_TRY(ctx, hipMemsetAsync(rb.yd[0], 0, n * sizeof(float), ctx->stream));
  HIP_TRY(ctx, hipMemsetAsync(rb.yd[1], 0, n * sizeof(float), ctx->stream));
  return GOPS_OK;
}

// the block plan of the sweep kernels
gops_status upload_plan(gops_context *ctx, gops_graph *g, const Plan &p) {
  HIP_TRY(ctx, upload_plan_to(ctx->stream, g->dev, p));
  return GOPS_OK;
}

// the arrays build_csrt made on the host
gops_status upload_graph(gops_context *ctx, gops_graph *g, const HostCSRT &h, int weight_storage) {
  DevGraph &dg = g->dev;
  dg.n = h.n;
  dg.m = h.m;
  dg.weight_storage = h.has_weights ? weight_storage : GOPS_WEIGHT_NONE;
  HIP_TRY(ctx, upload(ctx, &dg.row_ptr, h.row_ptr.data(), h.row_ptr.size()));
  HIP_TRY(ctx, upload(ctx, &dg.col, h.col.data(), h.col.size()));
  HIP_TRY(ctx, upload(ctx, &dg.rank, h.rank.data(), h.rank.size()));
  HIP_TRY(ctx, upload(ctx, &dg.order, h.order.data(), h.order.size()));

  // per-node 1/out-degree (and 1/out_wsum): f64 for anchor sweeps, f32 for
  // delta sweeps.  0 marks a dangling node; a non-dangling value that
  // underflows in f32 is kept at the smallest subnormal so it stays
  // non-dangling.
  auto to32 = [](const std::vector<double> &v) {
    std::vector<float> f(v.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)v.size(); i++) {
      float x = (float)v[i];
      if (v[i] != 0.0 && x == 0.0f) x = std::numeric_limits<float>::denorm_min();
      f[i] = x;
    }
    return f;
  };
  {
    std::vector<float> f = to32(h.inv_deg);
    HIP_TRY(ctx, upload(ctx, &dg.inv_deg64, h.inv_deg.data(), h.inv_deg.size()));
    HIP_TRY(ctx, upload(ctx, &dg.inv_deg32, f.data(), f.size()));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (h.has_weights) {
    std::vector<float> f = to32(h.inv_ws);
    HIP_TRY(ctx, upload(ctx, &dg.inv_ws64, h.inv_ws.data(), h.inv_ws.size()));
    HIP_TRY(ctx, upload(ctx, &dg.inv_ws32, f.data(), f.size()));
    const size_t m = h.m;
    if (dg.weight_storage == GOPS_WEIGHT_F64) {
      HIP_TRY(ctx, upload(ctx, (double **)&dg.w, h.w.data(), m));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    } else if (dg.weight_storage == GOPS_WEIGHT_F32) {
      std::vector<float> w32(m);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)m; i++) w32[i] = (float)h.w[i];
      HIP_TRY(ctx, upload(ctx, (float **)&dg.w, w32.data(), m));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    } else {
      std::vector<uint16_t> w16(m);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < (int64_t)m; i++) w16[i] = f32_to_bf16((float)h.w[i]);
      HIP_TRY(ctx, upload(ctx, (uint16_t **)&dg.w, w16.data(), m));
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GOPS_OK;
}

gops_status ensure_events(gops_context *ctx, size_t count) {
  while (ctx->events.size() < count) {
    hipEvent_t ev;
    HIP_TRY(ctx, hipEventCreate(&ev));
    ctx->events.push_back(ev);
  }
  return GOPS_OK;
}

}  // namespace

namespace gops {

// A graph lives while its caller's handle or a server on it does; it keeps
// its context alive until then.
void release_graph(gops_graph *g) {
  if (!g || g->refs.fetch_sub(1) > 1) return;
  gops_context *ctx = g->ctx;
  if (ctx) {
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
  }
  free_ppr_cache(g);
  free_run(g->run);
  free_power(g->pow);
  free_dev(g->dev);
  free_dev(g->fwd);
  free_blocked(g->pb);
  free_blocked(g->pb_fwd);
  delete g;
  gops_context_destroy(ctx);  // the graph's reference
}

// Shared driver of pagerank_run / pagerank_weighted_run /
// personalized_pagerank_run: host-side setup, the device sweep chain, the
// result in the caller's ids.
gops_status run_single(gops_context *ctx, gops_graph *g, const gops_pagerank_config *cfg,
                       RunKind kind, const double *pers, uint64_t pers_len, double *scores,
                       gops_run_info *info) {
  if (!ctx || !g || !cfg) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  ctx->last_error.clear();
  const uint64_t n = g->dev.n;
  gops_run_info local{};
  if (!info) info = &local;
  ctx->stats = gops_run_stats{};
  ctx->stats.n = n;
  ctx->stats.nnz = g->dev.m;
  if (kind == kWeighted && !g->weighted)
    return fail(ctx, GOPS_INVALID_PARAMETER, "graph was created without edge weights");
  if (kind == kPersonalized && pers_len != n) {
    if (pers_len > n || !pers)
      return fail(ctx, GOPS_DIMENSION_MISMATCH,
                  "dimension mismatch: " + std::to_string(pers_len) + " vs " + std::to_string(n));
    return fail(ctx, GOPS_DIMENSION_MISMATCH,
                "dimension mismatch: " + std::to_string(pers_len) + " vs " + std::to_string(n));
  }
  if (n == 0) {  // pagerank.rs:124-131, ppr.rs:97-103
    info->iterations = 0;
    info->diff_l1 = 0.0;
    info->converged = 1;
    return GOPS_OK;
  }
  gops_status st = alloc_run(ctx, g);
  if (st != GOPS_OK) return st;
  RunBuffers &rb = g->run;
  SweepSetup setup{kind == kWeighted ? 1 : 0, kind == kPersonalized ? 1 : 0, 0, 0};
  // f32 delta sweeps in the blocked layout (built once per graph): when chosen,
  // or by default from 2^20 nodes, where the pull's source vector (4 MB of
  // f32 and more) outgrows the L2 and the blocked passes measured faster.  By
  // default, a run whose layout does not fit in free device memory keeps the
  // pull and says so in its stats; a later run tries again.
  const bool auto_blocked = ctx->delta_sweep == GOPS_DELTA_SWEEP_AUTO && n >= kPbAutoNodes;
  if (ctx->precision == GOPS_PRECISION_MIXED &&
      (ctx->delta_sweep == GOPS_DELTA_SWEEP_BLOCKED || auto_blocked)) {
    std::string err;
    st = build_blocked(ctx, g, err);
    if (st == GOPS_OUT_OF_MEMORY && ctx->delta_sweep == GOPS_DELTA_SWEEP_AUTO) {
      ctx->stats.blocked_fallback = 1;
      (void)hipGetLastError();  // the failed allocation's error is not this run's
      ctx->last_error.clear();
    } else if (st != GOPS_OK) {
      return fail(ctx, st, err);
    } else {
      setup.blocked = 1;
      setup.anchors_blocked = ctx->anchor_sweep != GOPS_ANCHOR_SWEEP_PULL;
      ctx->stats.blocked_bins = g->pb.num_bins;
      ctx->stats.blocked_split_bins = g->pb.num_split_bins;
    }
  }

  if (kind == kPersonalized) {
    // ppr.rs:104-109: p / sum(p), or uniform when the sum is not > 0
    double psum = 0.0;
    for (uint64_t i = 0; i < n; i++) psum += pers[i];
    std::vector<double> q64(n);
    std::vector<float> q32(n);
    const uint32_t *order = g->order_host.data();
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)n; r++) {
      const double v = psum > 0.0 ? pers[order[r]] / psum : 1.0 / (double)n;
      q64[r] = v;
      q32[r] = (float)v;
    }
    if (!rb.q64) {
      HIP_TRY(ctx, dalloc(&rb.q64, n));
      HIP_TRY(ctx, dalloc(&rb.q32, n));
    }
    HIP_TRY(ctx, hipMemcpyAsync(rb.q64, q64.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(rb.q32, q32.data(), n * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }

  RunState hs{};
  hs.damping = cfg->damping;
  hs.tolerance = cfg->tolerance;
  hs.n_f64 = (double)n;
  hs.inv_n = 1.0 / (double)n;
  hs.max_iter = cfg->max_iterations;
  hs.mirror = ctx->mirror_dev;
  hs.all_f64 = ctx->precision == GOPS_PRECISION_FP64;
  hs.mode = kModeAnchor;  // the first sweep gathers x0 in f64
  // Mixed-precision schedule (see pagerank.hip).  Global PageRank: f32 delta
  // sweeps from the second sweep, re-anchored in f64 every anchor_drop of
  // residual.  Personalized PageRank: f64 sweeps while the residual is above
  // f32_below * tolerance, then an f32 delta chain re-anchored only every
  // anchor_period sweeps (its localized residual decays faster than the f32
  // errors an anchor would feed back from the iterate).
  const bool ppr = kind == kPersonalized;
  hs.f32_below = ppr ? ctx->ppr_f32_below : ctx->f32_below;
  hs.anchor_drop = ppr ? 0.0 : ctx->anchor_drop;
  hs.anchor_period = (uint64_t)std::max(2, ppr ? ctx->ppr_anchor_period : ctx->anchor_period);
  hs.anchor_min = ppr ? ctx->ppr_anchor_min : 0.0;
  hs.f64_phase = hs.all_f64 || hs.f32_below > 0.0;
  hs.x_scale = 1.0;
  // global runs with every sweep blocked keep the isolated rows' iterate in
  // the state (the sweeps' SweepArgs::iso under the same condition)
  hs.x_iso = hs.inv_n;
  hs.n_iso = (setup.blocked && setup.anchors_blocked && !ppr && g->has_pb) ? (double)g->pb.n_iso : 0.0;
  hs.sweep_log = rb.sweep_log;
  hs.sweep_log_cap = kSweepLogCap;
  __atomic_store_n(ctx->mirror_host, (uint64_t)0, __ATOMIC_RELEASE);
  // from pinned slot 1, in stream order before the first kernel; slot 0
  // takes the state back at the end
  ctx->host_state[1] = hs;
  HIP_TRY(ctx, hipMemcpyAsync(rb.state, &ctx->host_state[1], sizeof(RunState),
                              hipMemcpyHostToDevice, ctx->stream));

  const uint64_t max_it = cfg->max_iterations;
  const bool timing = ctx->timing != 0;
  const uint64_t timed_cap = timing ? std::min<uint64_t>(max_it, kSweepLogCap) : 0;
  st = ensure_events(ctx, 2 + 2 * (size_t)timed_cap);
  if (st != GOPS_OK) return st;
  hipEvent_t ev_begin = ctx->events[0], ev_end = ctx->events[1];

  // The sweeps go kernel by kernel (the default) or as replays of a captured
  // chain of kChainSweeps sweeps (GOPS_LAUNCH_GRAPH: one launch per chain
  // instead of one per kernel).  Every launch of a chain has identical
  // arguments (the device picks modes and buffers), so one capture per graph
  // and kind of run serves every later run.  Measured per run, queued against
  // replayed: 0.37 / 0.39 ms (ER 2000, 19 sweeps), 2.75 / 2.75 ms (R-MAT 20),
  // 27.5 / 27.3 ms (R-MAT 24): ROCm dispatches a graph's kernels at about the
  // cost of separate launches, and the replays past convergence cost more.
  const bool use_graph = ctx->launch == GOPS_LAUNCH_GRAPH && !timing;
  hipGraphExec_t chain = nullptr;
  if (use_graph) {
    hipGraphExec_t &slot = rb.chain[setup.weighted | (setup.personalized << 1) | (setup.blocked << 2)];
    if (!slot) {
      hipGraph_t graph = nullptr;
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the state upload first, uncaptured
      HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      for (int k = 0; k < kChainSweeps && ce == hipSuccess; k++) ce = launch_sweep(ctx, g, setup);
      const hipError_t ee = hipStreamEndCapture(ctx->stream, &graph);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(ctx, ce != hipSuccess ? ce : ee, "capturing the sweep chain");
      }
      const hipError_t ie = hipGraphInstantiate(&slot, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        slot = nullptr;
        return hip_fail(ctx, ie, "hipGraphInstantiate");
      }
    }
    chain = slot;
  }
  ctx->stats.graph_replay = use_graph ? 1 : 0;

  HIP_TRY(ctx, hipEventRecord(ev_begin, ctx->stream));
  HIP_TRY(ctx, launch_init(ctx, g, setup));

  // The device decides when to stop (RunState::done; later sweeps return at
  // once) and which kind each sweep is.  The host keeps a few identical
  // launches queued beyond the last sweep it has seen finish, reading the
  // progress word the last workgroup of every sweep writes to mapped host
  // memory, so no sweep waits on the host and at most kAhead launches are
  // wasted after convergence.
  // A chain replay may run up to kChainSweeps - 1 sweeps past the iteration
  // cap; those return at once (the cap sets the done flag).
  //
  // With every sweep blocked, a sweep's kernels are of one kind: the f32
  // delta kernels or their f64 anchor forms, each returning at once when the
  // sweep is of the other kind.  The progress word also carries the mode of
  // the next sweep, and the mode after it follows from that one (an anchor
  // comes right after a fold sweep and, past a leading f64 phase, is followed
  // by a delta sweep), so once sweep j - 2 has finished the host knows the
  // kind of sweep j and launches only its kernels: one sweep runs while the
  // next is queued, and no empty launches of the other kind remain (about
  // 14 us per sweep at R-MAT scale 24).
  const bool by_kind =
      !use_graph && setup.blocked && setup.anchors_blocked && g->has_pb && !hs.f64_phase;
  const uint64_t step = use_graph ? kChainSweeps : 1;
  const uint64_t kAhead = use_graph ? kChainAhead * step : (by_kind ? 2 : 4);
  uint64_t j = 0, seen = 0;
  int mode_seen = hs.mode;  // the mode of sweep `seen`, the first not yet finished
  while (j < max_it) {
    for (; j < max_it && j < seen + kAhead; j += step) {
      if (use_graph) {
        HIP_TRY(ctx, hipGraphLaunch(chain, ctx->stream));
        continue;
      }
      int kinds = 3;
      if (by_kind) {
        const bool anchor = j == seen ? mode_seen == kModeAnchor : mode_seen == kModeDeltaFold;
        kinds = anchor ? 2 : 1;
      }
      if (j < timed_cap) HIP_TRY(ctx, hipEventRecord(ctx->events[2 + 2 * j], ctx->stream));
      HIP_TRY(ctx, launch_sweep(ctx, g, setup, kinds));
      if (j < timed_cap) HIP_TRY(ctx, hipEventRecord(ctx->events[3 + 2 * j], ctx->stream));
    }
    if (j >= max_it) break;
    bool done = false;
    for (uint64_t spin = 0;; spin++) {
      const uint64_t w = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
      if (w & 1) {
        done = true;
        break;
      }
      if ((w >> 4) > seen) {
        seen = w >> 4;
        mode_seen = (int)((w >> 2) & 3);
        break;
      }
      if ((spin & 1023) == 1023) {
        const hipError_t q = hipStreamQuery(ctx->stream);
        if (q == hipSuccess) {  // queue drained: whatever the word says is final
          const uint64_t w2 = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
          done = (w2 & 1) != 0;
          if (!done && (w2 >> 4) <= seen)
            return fail(ctx, GOPS_DEVICE_ERROR, "sweep chain stalled without progress");
          seen = w2 >> 4;
          mode_seen = (int)((w2 >> 2) & 3);
          break;
        }
        if (q != hipErrorNotReady) return hip_fail(ctx, q, "hipStreamQuery");
      }
      __builtin_ia32_pause();
    }
    if (done) break;
  }
  HIP_TRY(ctx, launch_finalize(ctx, g, ctx->normalize));
  rb.has_result = true;
  HIP_TRY(ctx, hipEventRecord(ev_end, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(&ctx->host_state[0], rb.state, sizeof(RunState),
                              hipMemcpyDeviceToHost, ctx->stream));
  // the mode log of as many sweeps as the run may take, with the state
  const uint64_t log_n = std::min<uint64_t>(max_it, kSweepLogCap);
  if (log_n)
    HIP_TRY(ctx, hipMemcpyAsync(ctx->host_log, rb.sweep_log, log_n, hipMemcpyDeviceToHost,
                                ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  RunState last = ctx->host_state[0];
  if (last.kind_miss) {
    rb.has_result = false;
    return fail(ctx, GOPS_DEVICE_ERROR,
                "sweep " + std::to_string(last.iter + 1) + " was launched for the other kind");
  }
  if (max_it == 0) {  // unchecked max_iterations == 0 (pagerank.rs:139-141, 182-186)
    info->iterations = 0;
    info->diff_l1 = std::numeric_limits<double>::infinity();
    info->converged = 0;
  } else {
    info->iterations = last.iter;
    info->diff_l1 = last.diff;
    info->converged = last.converged;
  }
  float ms = 0.0f;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ev_begin, ev_end));
  ctx->stats.device_ms = ms;
  const uint64_t logged = std::min<uint64_t>(last.iter, log_n);
  for (uint64_t k = 0; k < logged; k++) {
    const bool f64 = ctx->host_log[k] == kModeAnchor;
    if (f64) ctx->stats.sweeps_fp64++;
    else ctx->stats.sweeps_fp32++;
    if (setup.blocked && (!f64 || setup.anchors_blocked)) ctx->stats.sweeps_blocked++;
    if (k < timed_cap) {
      float t = 0.0f;
      HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->events[2 + 2 * k], ctx->events[3 + 2 * k]));
      if (f64) ctx->stats.fp64_sweep_ms += t;
      else ctx->stats.fp32_sweep_ms += t;
    }
  }
  if (scores) {
    HIP_TRY(ctx, staged_download(ctx, scores, rb.out, n * sizeof(double), ctx->stream));
  }
  return GOPS_OK;
}

}  // namespace gops

// ---------------------------------------------------------------------------
extern "C" {

gops_status gops_context_create(int device, gops_context **out) {
  if (!out) return GOPS_INVALID_PARAMETER;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return GOPS_DEVICE_ERROR;
  if (device < 0 || device >= count) return GOPS_INVALID_PARAMETER;
  if (hipSetDevice(device) != hipSuccess) return GOPS_DEVICE_ERROR;
  auto *ctx = new (std::nothrow) gops_context;
  if (!ctx) return GOPS_OUT_OF_MEMORY;
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return GOPS_DEVICE_ERROR;
  }
  if (hipHostMalloc((void **)&ctx->host_state, 2 * sizeof(RunState), hipHostMallocDefault) !=
      hipSuccess) {
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return GOPS_OUT_OF_MEMORY;
  }
  if (hipHostMalloc((void **)&ctx->host_log, kSweepLogCap, hipHostMallocDefault) != hipSuccess) {
    hipHostFree(ctx->host_state);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return GOPS_OUT_OF_MEMORY;
  }
  if (hipHostMalloc((void **)&ctx->mirror_host, 64, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&ctx->mirror_dev, ctx->mirror_host, 0) != hipSuccess) {
    hipHostFree(ctx->host_log);
    hipHostFree(ctx->host_state);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return GOPS_OUT_OF_MEMORY;
  }
  *ctx->mirror_host = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  // set-up of the allocation-free top_k: its scratch array's address on this
  // device (looked up once per device; a failure here is retried by the call)
  if (gops::noalloc_scratch(device, &ctx->noalloc_scratch) != hipSuccess) {
    ctx->noalloc_scratch = nullptr;
    (void)hipGetLastError();
  }
  *out = ctx;
  return GOPS_OK;
}

void gops_context_destroy(gops_context *ctx) {
  if (!ctx) return;
  // graphs made on the context keep it until they are destroyed too (Python
  // finalizers, for one, may run in any order)
  if (ctx->refs.fetch_sub(1) > 1) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (hipEvent_t e : ctx->events) hipEventDestroy(e);
  free_topk(ctx->topk);
  free_topk_batch(ctx->topk_batch);
  hipFree(ctx->staging);
  free_pinned_chunks(ctx);
  hipHostFree(ctx->host_state);
  hipHostFree(ctx->host_log);
  hipHostFree(ctx->mirror_host);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *gops_last_error(const gops_context *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

gops_status gops_context_set_option(gops_context *ctx, int option, double value) {
  if (!ctx) return GOPS_INVALID_PARAMETER;
  switch (option) {
    case GOPS_OPT_PRECISION:
      if (value != GOPS_PRECISION_MIXED && value != GOPS_PRECISION_FP64)
        return fail(ctx, GOPS_INVALID_PARAMETER, "unknown precision");
      ctx->precision = (int)value;
      return GOPS_OK;
    case GOPS_OPT_ANCHOR_PERIOD:
      if (!(value >= 1 && value <= 1e6)) return fail(ctx, GOPS_INVALID_PARAMETER, "anchor period");
      ctx->anchor_period = (int)value;
      return GOPS_OK;
    case GOPS_OPT_NORMALIZE:
      ctx->normalize = value != 0;
      return GOPS_OK;
    case GOPS_OPT_TIMING:
      ctx->timing = value != 0;
      return GOPS_OK;
    case GOPS_OPT_ANCHOR_DROP:
      if (!(value > 0.0 && value < 1.0)) return fail(ctx, GOPS_INVALID_PARAMETER, "anchor drop");
      ctx->anchor_drop = value;
      return GOPS_OK;
    case GOPS_OPT_F32_BELOW:
    case GOPS_OPT_PPR_F32_BELOW:
    case GOPS_OPT_TILE_F32_BELOW:
      if (!(value >= 0.0)) return fail(ctx, GOPS_INVALID_PARAMETER, "f32 threshold must be >= 0");
      (option == GOPS_OPT_F32_BELOW       ? ctx->f32_below
       : option == GOPS_OPT_PPR_F32_BELOW ? ctx->ppr_f32_below
                                          : ctx->tile_f32_below) = value;
      return GOPS_OK;
    case GOPS_OPT_PPR_ANCHOR_PERIOD:
      if (!(value >= 2 && value <= 1e6)) return fail(ctx, GOPS_INVALID_PARAMETER, "anchor period");
      ctx->ppr_anchor_period = (int)value;
      return GOPS_OK;
    case GOPS_OPT_PPR_TILE:
      if (!(value == -1 || value == 0 || value == 8 || value == 16 || value == 32 || value == 64))
        return fail(ctx, GOPS_INVALID_PARAMETER, "ppr tile must be -1, 0, 8, 16, 32 or 64");
      ctx->ppr_tile = (int)value;
      return GOPS_OK;
    case GOPS_OPT_GRAPH_BUILD:
      if (value != GOPS_BUILD_DEVICE && value != GOPS_BUILD_HOST)
        return fail(ctx, GOPS_INVALID_PARAMETER, "graph build must be GOPS_BUILD_DEVICE or _HOST");
      ctx->graph_build = (int)value;
      return GOPS_OK;
    case GOPS_OPT_LAUNCH:
      if (value != GOPS_LAUNCH_AUTO && value != GOPS_LAUNCH_QUEUED && value != GOPS_LAUNCH_GRAPH)
        return fail(ctx, GOPS_INVALID_PARAMETER,
                    "launch must be GOPS_LAUNCH_AUTO, _QUEUED or _GRAPH");
      ctx->launch = (int)value;
      return GOPS_OK;
    case GOPS_OPT_ANCHOR_SWEEP:
      if (value != GOPS_ANCHOR_SWEEP_AUTO && value != GOPS_ANCHOR_SWEEP_PULL)
        return fail(ctx, GOPS_INVALID_PARAMETER,
                    "anchor sweep must be GOPS_ANCHOR_SWEEP_AUTO or _PULL");
      ctx->anchor_sweep = (int)value;
      return GOPS_OK;
    case GOPS_OPT_DELTA_SWEEP:
      if (value != GOPS_DELTA_SWEEP_AUTO && value != GOPS_DELTA_SWEEP_PULL &&
          value != GOPS_DELTA_SWEEP_BLOCKED)
        return fail(ctx, GOPS_INVALID_PARAMETER,
                    "delta sweep must be GOPS_DELTA_SWEEP_AUTO, _PULL or _BLOCKED");
      ctx->delta_sweep = (int)value;
      return GOPS_OK;
    case GOPS_OPT_BLOCKED_BIN_RUNS:
      if (!(value == 0 || (value >= 4096 && value <= 1 << 24)))
        return fail(ctx, GOPS_INVALID_PARAMETER, "blocked bin runs must be 0 or in [4096, 2^24]");
      ctx->pb_bin_runs = (int)value;
      return GOPS_OK;
    case GOPS_OPT_CENTRALITY_SWEEP:
      if (value != GOPS_CENTRALITY_SWEEP_AUTO && value != GOPS_CENTRALITY_SWEEP_PULL &&
          value != GOPS_CENTRALITY_SWEEP_BLOCKED)
        return fail(ctx, GOPS_INVALID_PARAMETER,
                    "centrality sweep must be GOPS_CENTRALITY_SWEEP_AUTO, _PULL or _BLOCKED");
      ctx->centrality_sweep = (int)value;
      return GOPS_OK;
    default:
      return fail(ctx, GOPS_INVALID_PARAMETER, "unknown option");
  }
}

// Replace the graph's row plan (the pull sweeps' work descriptors) by p.
static gops_status replace_plan(gops_context *ctx, gops_graph *g, const Plan &p) {
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  DevGraph &dg = g->dev;
  for (void *q : {(void *)dg.desc, (void *)dg.hub_row, (void *)dg.hub_nchunks, (void *)dg.hub_first})
    (void)hipFree(q);
  dg.desc = nullptr;
  dg.hub_row = dg.hub_nchunks = dg.hub_first = nullptr;
  HIP_TRY(ctx, upload_plan_to(ctx->stream, dg, p));
  RunBuffers &rb = g->run;
  if (rb.state) {  // the run buffers' hub partials follow the plan's chunk count
    (void)hipFree(rb.hub_partial);
    rb.hub_partial = nullptr;
    HIP_TRY(ctx, dalloc(&rb.hub_partial, std::max(dg.num_hub_chunks, 1)));
  }
  return GOPS_OK;
}

static gops_status download_rows(gops_context *ctx, gops_graph *g, HostCSRT &h) {
  h.n = g->dev.n;
  h.m = g->dev.m;
  h.row_ptr.resize(h.n + 1);
  HIP_TRY(ctx, staged_download(ctx, h.row_ptr.data(), g->dev.row_ptr,
                               (h.n + 1) * sizeof(int32_t), ctx->stream));
  return GOPS_OK;
}

gops_status gops_graph_set_rows(gops_context *ctx, gops_graph *g, const gops_row_plan *plan) {
  if (!ctx || !g || !plan) return GOPS_INVALID_PARAMETER;
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  ctx->last_error.clear();
  if (!(plan->row_min >= 64 && plan->row_min <= kStageNnz))
    return fail(ctx, GOPS_INVALID_PARAMETER, "row_min must be in [64, 4096]");
  if (!(plan->hub_chunk >= plan->row_min && plan->hub_chunk <= (1 << 26)))
    return fail(ctx, GOPS_INVALID_PARAMETER, "hub_chunk must be in [row_min, 2^26]");
  if (!(plan->row_iters >= 1 && plan->row_iters <= 64))
    return fail(ctx, GOPS_INVALID_PARAMETER, "row_iters must be in [1, 64]");
  const PlanParams pp{plan->row_min, plan->hub_chunk, plan->row_iters};
  if (g->dev.n) {
    HostCSRT h;
    gops_status st = download_rows(ctx, g, h);
    if (st != GOPS_OK) return st;
    Plan p;
    plan_blocks(h, p, pp);
    st = replace_plan(ctx, g, p);
    if (st != GOPS_OK) return st;
  }
  g->plan_params = pp;
  return GOPS_OK;
}

gops_status gops_graph_tune_rows(gops_context *ctx, gops_graph *g, gops_row_plan *out) {
  if (!ctx || !g) return GOPS_INVALID_PARAMETER;
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  ctx->last_error.clear();
  const uint64_t n = g->dev.n;
  gops_row_plan res{g->plan_params.row_min, g->plan_params.hub_chunk, g->plan_params.row_iters,
                    0, 0.0, 0.0};
  if (n == 0 || g->dev.m == 0) {
    if (out) *out = res;
    return GOPS_OK;
  }
  HostCSRT h;
  gops_status st = download_rows(ctx, g, h);
  if (st != GOPS_OK) return st;
  // timed runs: f32 pull sweeps with every kernel timed, the context's
  // options and last statistics restored afterwards
  const int saved_prec = ctx->precision, saved_ds = ctx->delta_sweep, saved_launch = ctx->launch;
  const bool saved_timing = ctx->timing;
  const gops_run_stats saved_stats = ctx->stats;
  ctx->precision = GOPS_PRECISION_MIXED;
  ctx->delta_sweep = GOPS_DELTA_SWEEP_PULL;
  ctx->launch = GOPS_LAUNCH_QUEUED;
  ctx->timing = true;
  const RunKind kind = g->weighted ? kWeighted : kPlain;
  const gops_pagerank_config cfg{0.85, 10, 0.0};
  auto measure = [&](const PlanParams &pp, double *us) {
    Plan p;
    plan_blocks(h, p, pp);
    st = replace_plan(ctx, g, p);
    double best = std::numeric_limits<double>::infinity();
    for (int rep = 0; rep < 3 && st == GOPS_OK; rep++) {
      gops_run_info info{};
      st = run_single(ctx, g, &cfg, kind, nullptr, 0, nullptr, &info);
      if (st == GOPS_OK && ctx->stats.sweeps_fp32)
        best = std::min(best, ctx->stats.fp32_sweep_ms * 1e3 / (double)ctx->stats.sweeps_fp32);
    }
    *us = best;
    res.candidates++;
  };
  // coordinate search from the defaults: the run cut-off, then the hub
  // chunk, then the rows per lane group; a change must gain 1 %
  PlanParams best_pp{}, pp{};
  double best_us = 0.0;
  measure(best_pp, &best_us);
  res.default_us = best_us;
  const int32_t row_mins[] = {2048, 1024, 512};
  const int32_t hub_chunks[] = {32768, 16384};
  const int32_t row_iters[] = {4, 16};
  for (int32_t v : row_mins) {
    if (st != GOPS_OK) break;
    pp = best_pp;
    pp.row_min = v;
    double us;
    measure(pp, &us);
    if (us < 0.99 * best_us) best_pp = pp, best_us = us;
  }
  for (int32_t v : hub_chunks) {
    if (st != GOPS_OK) break;
    pp = best_pp;
    pp.hub_chunk = v;
    double us;
    measure(pp, &us);
    if (us < 0.99 * best_us) best_pp = pp, best_us = us;
  }
  for (int32_t v : row_iters) {
    if (st != GOPS_OK) break;
    pp = best_pp;
    pp.row_iters = v;
    double us;
    measure(pp, &us);
    if (us < 0.99 * best_us) best_pp = pp, best_us = us;
  }
  ctx->precision = saved_prec;
  ctx->delta_sweep = saved_ds;
  ctx->launch = saved_launch;
  ctx->timing = saved_timing;
  ctx->stats = saved_stats;
  if (st != GOPS_OK) return st;
  Plan p;
  plan_blocks(h, p, best_pp);
  st = replace_plan(ctx, g, p);
  if (st != GOPS_OK) return st;
  g->plan_params = best_pp;
  res.row_min = best_pp.row_min;
  res.hub_chunk = best_pp.hub_chunk;
  res.row_iters = best_pp.row_iters;
  res.tuned_us = best_us;
  if (out) *out = res;
  return GOPS_OK;
}

gops_status gops_last_run_stats(const gops_context *ctx, gops_run_stats *out) {
  if (!ctx || !out) return GOPS_INVALID_PARAMETER;
  *out = ctx->stats;
  return GOPS_OK;
}

gops_status gops_graph_create(gops_context *ctx, uint64_t n, const uint64_t *offsets,
                              const uint32_t *targets, const double *edge_weights,
                              int weight_storage, gops_graph **out) {
  if (!ctx || !out) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  *out = nullptr;
  if (!offsets) return fail(ctx, GOPS_INVALID_PARAMETER, "offsets must not be null");
  if (offsets[n] > 0 && !targets) return fail(ctx, GOPS_INVALID_PARAMETER, "targets must not be null");
  if (weight_storage < GOPS_WEIGHT_NONE || weight_storage > GOPS_WEIGHT_BF16)
    return fail(ctx, GOPS_INVALID_PARAMETER, "unknown weight storage");
  if (weight_storage != GOPS_WEIGHT_NONE && !edge_weights && offsets[n] > 0)
    return fail(ctx, GOPS_INVALID_PARAMETER, "weight storage given without edge weights");
  // a weighted graph without edges still has to be flagged as weighted
  static const double kNoWeights[1] = {0.0};
  const double *ew = nullptr;
  if (weight_storage != GOPS_WEIGHT_NONE) ew = edge_weights ? edge_weights : kNoWeights;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto *g = new (std::nothrow) gops_graph;
  if (!g) return fail(ctx, GOPS_OUT_OF_MEMORY, "host allocation failed");
  g->ctx = ctx;
  g->weighted = weight_storage != GOPS_WEIGHT_NONE;
  std::string err;
  Plan p;
  gops_status st;
  if (ctx->graph_build == GOPS_BUILD_DEVICE) {
    st = build_on_device(ctx, n, offsets, targets, ew, weight_storage, g, p, err);
    if (st != GOPS_OK) {
      delete g;
      return fail(ctx, st, err);
    }
  } else {
    HostCSRT h;
    st = build_csrt(n, offsets, targets, ew, h, err);
    if (st != GOPS_OK) {
      delete g;
      return fail(ctx, st, err);
    }
    plan_blocks(h, p);
    g->first_bad_weight = h.first_bad_weight;
    g->order_host = h.order;
    g->rank_host = h.rank;
    st = upload_graph(ctx, g, h, weight_storage);
  }
  if (st == GOPS_OK) st = upload_plan(ctx, g, p);
  if (st != GOPS_OK) {
    free_dev(g->dev);
    delete g;
    return st;
  }
  ctx->refs.fetch_add(1);  // released with the graph
  *out = g;
  return GOPS_OK;
}

gops_status gops_graph_from_dense(gops_context *ctx, uint64_t n, const double *matrix,
                                  int weight_storage, gops_graph **out) {
  if (!ctx || !out) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (n > 0 && !matrix) return fail(ctx, GOPS_INVALID_PARAMETER, "matrix must not be null");
  if (n >= (1ull << 31)) return fail(ctx, GOPS_NOT_SUPPORTED, "node count must be < 2^31");
  // AdjacencyMatrix::neighbors (graph.rs:128-135): columns with entry > 0.0
  std::vector<uint64_t> off(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    uint64_t c = 0;
    const double *row = matrix + (uint64_t)u * n;
    for (uint64_t v = 0; v < n; v++) c += row[v] > 0.0;
    off[u + 1] = c;
  }
  for (uint64_t u = 0; u < n; u++) off[u + 1] += off[u];
  std::vector<uint32_t> nb(off[n]);
  std::vector<double> ew(weight_storage != GOPS_WEIGHT_NONE ? off[n] : 0);
#pragma omp parallel for schedule(static)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    uint64_t k = off[u];
    const double *row = matrix + (uint64_t)u * n;
    for (uint64_t v = 0; v < n; v++) {
      if (row[v] > 0.0) {
        nb[k] = (uint32_t)v;
        if (!ew.empty()) ew[k] = row[v];  // edge_weight = raw entry, graph.rs:139-141
        k++;
      }
    }
  }
  return gops_graph_create(ctx, n, off.data(), nb.data(), ew.empty() ? nullptr : ew.data(),
                           weight_storage, out);
}

void gops_graph_destroy(gops_graph *g) { gops::release_graph(g); }

gops_status gops_graph_info(const gops_graph *g, uint64_t *n, uint64_t *nnz) {
  if (!g) return GOPS_INVALID_PARAMETER;
  if (n) *n = g->dev.n;
  if (nnz) *nnz = g->dev.m;
  return GOPS_OK;
}

gops_status gops_graph_export(const gops_graph *g, int32_t *row_ptr, int32_t *col, void *weights,
                              uint32_t *order, double *inv_deg, double *inv_ws) {
  if (!g || !g->ctx) return GOPS_INVALID_PARAMETER;
  gops_context *ctx = g->ctx;
  const DevGraph &dg = g->dev;
  const size_t n = dg.n, m = dg.m;
  if (weights && dg.weight_storage == GOPS_WEIGHT_NONE)
    return fail(ctx, GOPS_INVALID_PARAMETER, "graph has no edge weights");
  if (inv_ws && !dg.inv_ws64) return fail(ctx, GOPS_INVALID_PARAMETER, "graph has no edge weights");
  const size_t wb = dg.weight_storage == GOPS_WEIGHT_F64 ? 8 : dg.weight_storage == GOPS_WEIGHT_F32 ? 4 : 2;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  auto down = [&](void *dst, const void *src, size_t bytes) {
    return (dst && bytes) ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
  };
  HIP_TRY(ctx, down(row_ptr, dg.row_ptr, (n + 1) * sizeof(int32_t)));
  HIP_TRY(ctx, down(col, dg.col, m * sizeof(int32_t)));
  HIP_TRY(ctx, down(weights, dg.w, m * wb));
  HIP_TRY(ctx, down(order, dg.order, n * sizeof(uint32_t)));
  HIP_TRY(ctx, down(inv_deg, dg.inv_deg64, n * sizeof(double)));
  HIP_TRY(ctx, down(inv_ws, dg.inv_ws64, n * sizeof(double)));
  HIP_TRY(ctx, hipStreamSynchronize(s));
  return GOPS_OK;
}

gops_status gops_pagerank_run(gops_context *ctx, gops_graph *g, const gops_pagerank_config *cfg,
                              double *scores, gops_run_info *info) {
  return run_single(ctx, g, cfg, kPlain, nullptr, 0, scores, info);
}

gops_status gops_pagerank_checked_run(gops_context *ctx, gops_graph *g,
                                      const gops_pagerank_config *cfg, double *scores,
                                      gops_run_info *info) {
  if (!ctx || !cfg) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  gops_status st = validate_config(ctx, cfg);
  if (st != GOPS_OK) return st;
  return run_single(ctx, g, cfg, kPlain, nullptr, 0, scores, info);
}

gops_status gops_pagerank_weighted_run(gops_context *ctx, gops_graph *g,
                                       const gops_pagerank_config *cfg, double *scores,
                                       gops_run_info *info) {
  return run_single(ctx, g, cfg, kWeighted, nullptr, 0, scores, info);
}

gops_status gops_pagerank_weighted_checked_run(gops_context *ctx, gops_graph *g,
                                               const gops_pagerank_config *cfg, double *scores,
                                               gops_run_info *info) {
  if (!ctx || !cfg || !g) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  gops_status st = validate_config(ctx, cfg);
  if (st != GOPS_OK) return st;
  // pagerank.rs:300-314: first offending neighbour weight decides the message
  if (g->first_bad_weight == 1) return invalid(ctx, "edge weights must be finite");
  if (g->first_bad_weight == 2) return invalid(ctx, "edge weights must be non-negative");
  return run_single(ctx, g, cfg, kWeighted, nullptr, 0, scores, info);
}

gops_status gops_personalized_pagerank_run(gops_context *ctx, gops_graph *g,
                                           const gops_pagerank_config *cfg,
                                           const double *personalization, uint64_t len,
                                           double *scores, gops_run_info *info) {
  return run_single(ctx, g, cfg, kPersonalized, personalization, len, scores, info);
}

gops_status gops_personalized_pagerank_checked_run(gops_context *ctx, gops_graph *g,
                                                   const gops_pagerank_config *cfg,
                                                   const double *p, uint64_t len,
                                                   double *scores, gops_run_info *info) {
  if (!ctx || !cfg || !g) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  gops_status st = validate_config(ctx, cfg);
  if (st != GOPS_OK) return st;
  const uint64_t n = g->dev.n;
  if (len != n)  // ppr.rs:18-24
    return invalid(ctx, "personalization length must equal node_count (len=" +
                            std::to_string(len) + " node_count=" + std::to_string(n) + ")");
  if (len > 0 && !p) return fail(ctx, GOPS_INVALID_PARAMETER, "personalization must not be null");
  for (uint64_t i = 0; i < len; i++) {  // ppr.rs:25-36
    if (!std::isfinite(p[i])) return invalid(ctx, "personalization entries must be finite");
    if (p[i] < 0.0) return invalid(ctx, "personalization entries must be non-negative");
  }
  double sum = 0.0;
  for (uint64_t i = 0; i < len; i++) sum += p[i];
  if (sum <= 0.0) return invalid(ctx, "personalization sum must be > 0");  // ppr.rs:37-42
  return run_single(ctx, g, cfg, kPersonalized, p, len, scores, info);
}

gops_status gops_gather_probe(gops_context *ctx, gops_graph *g, uint32_t reps, double *best_ms) {
  if (!ctx || !g || !best_ms) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  gops_status st = alloc_run(ctx, g);
  if (st != GOPS_OK) return st;
  if ((st = ensure_events(ctx, 2)) != GOPS_OK) return st;
  double best = std::numeric_limits<double>::infinity();
  for (uint32_t r = 0; r <= reps; r++) {  // the first launch warms up
    HIP_TRY(ctx, hipEventRecord(ctx->events[0], ctx->stream));
    HIP_TRY(ctx, launch_gather_probe(ctx, g));
    HIP_TRY(ctx, hipEventRecord(ctx->events[1], ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->events[1]));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
    if (r > 0) best = std::min(best, (double)ms);
  }
  *best_ms = best;
  return GOPS_OK;
}

gops_status gops_top_k(gops_context *ctx, const double *scores, uint64_t n, uint32_t k,
                       uint32_t *ids, double *values, uint32_t *count) {
  if (!ctx || !count) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  *count = 0;
  if (k == 0 || n == 0) return GOPS_OK;  // topk.rs:18-20
  if (!scores) return fail(ctx, GOPS_INVALID_PARAMETER, "scores must not be null");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (ctx->staging_cap < n) {
    hipFree(ctx->staging);
    ctx->staging = nullptr;
    ctx->staging_cap = 0;
    HIP_TRY(ctx, dalloc(&ctx->staging, n));
    ctx->staging_cap = n;
  }
  HIP_TRY(ctx, staged_upload(ctx, ctx->staging, scores, n * sizeof(double), ctx->stream));
  std::string err;
  const gops_status st = device_top_k(ctx, ctx->staging, n, k, ids, values, count, err);
  return st == GOPS_OK ? st : fail(ctx, st, err);
}

gops_status gops_result_top_k(gops_context *ctx, gops_graph *g, uint32_t k, uint32_t *ids,
                              double *values, uint32_t *count) {
  if (!ctx || !g || !count) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  *count = 0;
  if (!g->run.has_result)
    return fail(ctx, GOPS_INVALID_PARAMETER, "no PageRank result on this graph yet");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::string err;
  const gops_status st = device_top_k(ctx, g->run.out, g->dev.n, k, ids, values, count, err);
  return st == GOPS_OK ? st : fail(ctx, st, err);
}

gops_status gops_result_top_k_noalloc(gops_context *ctx, gops_graph *g, uint32_t k, uint32_t *ids,
                                      double *values, uint32_t *count) {
  if (!ctx || !g || !count) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  *count = 0;
  if (!g->run.has_result)
    return fail(ctx, GOPS_INVALID_PARAMETER, "no PageRank result on this graph yet");
  if (k > kNoAllocTopK)
    return fail(ctx, GOPS_INVALID_PARAMETER,
                "top_k without allocation: k must be at most " + std::to_string(kNoAllocTopK) +
                    " (got " + std::to_string(k) + ")");
  const uint64_t n = g->dev.n;
  if (k == 0 || n == 0) return GOPS_OK;  // topk.rs:18-20
  if (n >= (1ull << 32)) return fail(ctx, GOPS_NOT_SUPPORTED, "top_k: at most 2^32 - 1 scores");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!ctx->noalloc_scratch) HIP_TRY(ctx, noalloc_scratch(ctx->device, &ctx->noalloc_scratch));
  std::string err;
  const gops_status st = top_k_noalloc(ctx, static_cast<unsigned char *>(ctx->noalloc_scratch),
                                       g->run.out, n, k, ids, values, count, err);
  return st == GOPS_OK ? st : fail(ctx, st, err);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched personalized PageRank, one query per device run (the plain path).
namespace gops {
gops_status ppr_batch_looped(gops_context *ctx, gops_graph *g, uint32_t nq,
                             const gops_pagerank_config *cfgs, const uint64_t *seed_off,
                             const uint32_t *seeds, const double *seed_w, double *scores,
                             gops_run_info *infos) {
  const uint64_t n = g->dev.n;
  std::vector<double> p(n, 0.0);
  for (uint32_t q = 0; q < nq; q++) {
    for (uint64_t j = seed_off[q]; j < seed_off[q + 1]; j++) {
      if (seeds[j] >= n)
        return fail(ctx, GOPS_INDEX_OUT_OF_BOUNDS, "index out of bounds: " + std::to_string(seeds[j]));
      p[seeds[j]] += seed_w ? seed_w[j] : 1.0;
    }
    gops_run_info info{};
    gops_status st = run_single(ctx, g, &cfgs[q], kPersonalized, p.data(), n,
                                scores ? scores + (size_t)q * n : nullptr, &info);
    if (st != GOPS_OK) return st;
    if (infos) infos[q] = info;
    for (uint64_t j = seed_off[q]; j < seed_off[q + 1]; j++) p[seeds[j]] = 0.0;
  }
  return GOPS_OK;
}
}  // namespace gops
