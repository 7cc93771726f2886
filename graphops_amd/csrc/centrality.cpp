// centrality.cpp -- host side of Katz, eigenvector and HITS centrality
// (katz.rs, eigenvector.rs, centrality.rs:114-166): validation, buffers,
// the device-driven iteration chain and the results in the caller's ids.
// Kernels: centrality.hip.
#include <algorithm>
#include <cmath>
#include <limits>
#include <string>

#include "internal.h"

using namespace gops;

namespace {

gops_status fail(gops_context *ctx, gops_status st, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return st;
}

gops_status invalid(gops_context *ctx, const std::string &msg) {
  return fail(ctx, GOPS_INVALID_PARAMETER, "invalid parameter: " + msg);  // lib.rs:117-118
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

// KatzConfig::validate, katz.rs:57-84 (messages verbatim)
gops_status validate_katz(gops_context *ctx, const gops_katz_config *c) {
  if (!std::isfinite(c->alpha) || c->alpha <= 0.0)
    return invalid(ctx, "alpha must be finite and > 0");
  if (c->alpha >= 1.0)
    return invalid(ctx,
                   "alpha must be < 1 (required for convergence; tighter bound "
                   "alpha < 1/spectral_radius(A) may be needed for large graphs)");
  if (!std::isfinite(c->beta)) return invalid(ctx, "beta must be finite");
  if (c->max_iterations == 0) return invalid(ctx, "max_iterations must be > 0");
  if (!std::isfinite(c->tolerance) || c->tolerance <= 0.0)
    return invalid(ctx, "tolerance must be finite and > 0");
  return GOPS_OK;
}

gops_status check_args(gops_context *ctx, gops_graph *g) {
  if (!ctx || !g) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  if (g->ctx != ctx) return fail(ctx, GOPS_INVALID_PARAMETER, "graph belongs to another context");
  ctx->last_error.clear();
  ctx->stats = gops_run_stats{};
  ctx->stats.n = g->dev.n;
  ctx->stats.nnz = g->dev.m;
  return GOPS_OK;
}

gops_status alloc_power(gops_context *ctx, gops_graph *g) {
  PowerBuffers &pb = g->pow;
  const size_t n = g->dev.n;
  if (!pb.state) {
    pb.n = n;
    for (double *&v : pb.v)
      if (!v) HIP_TRY(ctx, hipMalloc((void **)&v, std::max<size_t>(n, 1) * 8));
    if (!pb.partials)
      HIP_TRY(ctx, hipMalloc((void **)&pb.partials,
                             std::max<size_t>(2 * kMaxSweepBlocks, 2048) * sizeof(double)));
    HIP_TRY(ctx, hipMalloc((void **)&pb.state, sizeof(PowerState)));
  }
  const size_t need = (size_t)std::max({g->dev.num_hub_chunks, g->fwd.num_hub_chunks, 1});
  if (pb.hub_cap < need) {
    (void)hipFree(pb.hub_partial);
    pb.hub_partial = nullptr;
    pb.hub_cap = 0;
    HIP_TRY(ctx, hipMalloc((void **)&pb.hub_partial, need * sizeof(double)));
    pb.hub_cap = need;
  }
  return GOPS_OK;
}

gops_status find_degrees(gops_context *ctx, gops_graph *g);

// Blocked eigenvector / HITS passes: when chosen, or by default from 2^21
// nodes; the out-edge layout (and for HITS the in-edge one) built on first
// use.  By default, layouts that do not fit keep the pulls.
gops_status choose_power_blocked(gops_context *ctx, gops_graph *g, bool hits) {
  PowerBuffers &pb = g->pow;
  pb.power_blocked = false;
  if (!(ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_BLOCKED ||
        (ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_AUTO && g->dev.n >= kPowerAutoNodes)))
    return GOPS_OK;
  std::string err;
  gops_status st = build_blocked_fwd(ctx, g, err);
  if (st == GOPS_OK && hits) st = build_blocked(ctx, g, err);
  if (st == GOPS_OUT_OF_MEMORY && ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_AUTO) {
    ctx->stats.blocked_fallback = 1;
    (void)hipGetLastError();
    return GOPS_OK;
  }
  if (st != GOPS_OK) return fail(ctx, st, err);
  if ((st = find_degrees(ctx, g)) != GOPS_OK) return st;
  pb.power_blocked = true;
  return GOPS_OK;
}

// Rows with in-edges (n_in: rows are sorted by in-degree, descending) and
// the largest in-degree (row 0's), found once per graph from the row
// pointers by a binary search of single-word copies.
gops_status find_degrees(gops_context *ctx, gops_graph *g) {
  PowerBuffers &pb = g->pow;
  if (pb.have_degrees) return GOPS_OK;
  const uint64_t n = g->dev.n;
  auto rp = [&](uint64_t i, int32_t *v) {
    return hipMemcpy(v, g->dev.row_ptr + i, sizeof(int32_t), hipMemcpyDeviceToHost);
  };
  int32_t r0 = 0, r1 = 0;
  HIP_TRY(ctx, rp(0, &r0));
  HIP_TRY(ctx, rp(n > 0 ? 1 : 0, &r1));
  pb.indeg_max = (uint64_t)(r1 - r0);
  // the first row r with row_ptr[r + 1] == row_ptr[r] (n if none)
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    int32_t a = 0, b = 0;
    HIP_TRY(ctx, rp(mid, &a));
    HIP_TRY(ctx, rp(mid + 1, &b));
    if (b > a) lo = mid + 1;
    else hi = mid;
  }
  pb.n_in = lo;
  if (g->has_fwd) {  // the out-edge CSR's rows are sorted by out-degree, descending
    int32_t f0 = 0, f1 = 0;
    HIP_TRY(ctx, hipMemcpy(&f0, g->fwd.row_ptr, sizeof(int32_t), hipMemcpyDeviceToHost));
    HIP_TRY(ctx, hipMemcpy(&f1, g->fwd.row_ptr + (n > 0 ? 1 : 0), sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    pb.outdeg_max = (uint64_t)(f1 - f0);
  }
  pb.have_degrees = g->has_fwd;  // again once the out-edge CSR exists
  return GOPS_OK;
}

// Runs iterations of algo until the device state says done (at most max_it),
// keeping a few iterations queued beyond the last one seen to finish (as the
// PageRank chain does), then returns the final state.
gops_status drive(gops_context *ctx, gops_graph *g, PowerAlgo algo, const PowerState &init,
                  PowerState &last) {
  PowerBuffers &pb = g->pow;
  PowerState hs = init;
  hs.mirror = ctx->mirror_dev;
  __atomic_store_n(ctx->mirror_host, (uint64_t)0, __ATOMIC_RELEASE);
  HIP_TRY(ctx, hipMemcpyAsync(pb.state, &hs, sizeof(hs), hipMemcpyHostToDevice, ctx->stream));
  const uint64_t max_it = init.max_iter;
  const uint64_t kAhead = 3;
  uint64_t j = 0, seen = 0;
  while (j < max_it) {
    for (; j < max_it && j < seen + kAhead; j++)
      HIP_TRY(ctx, launch_power_iteration(ctx, g, algo));
    if (j >= max_it) break;
    bool done = false;
    for (uint64_t spin = 0;; spin++) {
      const uint64_t w = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
      if (w & 1) {
        done = true;
        break;
      }
      if ((w >> 2) > seen) {
        seen = w >> 2;
        break;
      }
      if ((spin & 1023) == 1023) {
        const hipError_t q = hipStreamQuery(ctx->stream);
        if (q == hipSuccess) {
          const uint64_t w2 = __atomic_load_n(ctx->mirror_host, __ATOMIC_ACQUIRE);
          done = (w2 & 1) != 0;
          if (!done && (w2 >> 2) <= seen)
            return fail(ctx, GOPS_DEVICE_ERROR, "iteration chain stalled without progress");
          seen = w2 >> 2;
          break;
        }
        if (q != hipErrorNotReady) return hip_fail(ctx, q, "hipStreamQuery");
      }
      __builtin_ia32_pause();
    }
    if (done) break;
  }
  HIP_TRY(ctx, hipMemcpyAsync(&last, pb.state, sizeof(last), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return GOPS_OK;
}

// v (new ids) -> out (caller ids) on the host, through the spare vector v[3]
gops_status download(gops_context *ctx, gops_graph *g, const double *v, double *out) {
  if (!out) return GOPS_OK;
  double *dst = g->pow.v[3];
  HIP_TRY(ctx, launch_permute(ctx, g, v, dst));
  HIP_TRY(ctx, staged_download(ctx, out, dst, g->dev.n * sizeof(double), ctx->stream));
  return GOPS_OK;
}

gops_status katz_run(gops_context *ctx, gops_graph *g, const gops_katz_config *cfg,
                     double *scores, gops_run_info *info) {
  gops_status st = check_args(ctx, g);
  if (st != GOPS_OK) return st;
  if (!cfg) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  gops_run_info local{};
  if (!info) info = &local;
  const uint64_t n = g->dev.n;
  if (n == 0) {  // katz.rs:138-145
    *info = gops_run_info{0, 0.0, 1, 0};
    return GOPS_OK;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if ((st = alloc_power(ctx, g)) != GOPS_OK) return st;
  PowerBuffers &pb = g->pow;
  // blocked sweeps: when chosen, or by default from 2^21 nodes; by default a layout that does not fit keeps the pull
  pb.katz_blocked = false;
  if (ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_BLOCKED ||
      (ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_AUTO && n >= kPowerAutoNodes)) {
    std::string err;
    st = build_blocked(ctx, g, err);
    if (st == GOPS_OUT_OF_MEMORY && ctx->centrality_sweep == GOPS_CENTRALITY_SWEEP_AUTO) {
      ctx->stats.blocked_fallback = 1;
      (void)hipGetLastError();  // the failed allocation's error is not this run's
    } else if (st != GOPS_OK) {
      return fail(ctx, st, err);
    } else {
      if ((st = find_degrees(ctx, g)) != GOPS_OK) return st;
      pb.katz_blocked = true;
      pb.katz_bound_scale = cfg->alpha * (double)pb.indeg_max;
      ctx->stats.blocked_bins = g->pb.num_bins;
      ctx->stats.blocked_split_bins = g->pb.num_split_bins;
    }
  }
  HIP_TRY(ctx, launch_power_init(ctx, g, kKatz, cfg->beta, cfg->alpha * cfg->beta));
  PowerState init{};
  init.alpha = cfg->alpha;
  init.beta = cfg->beta;
  init.tolerance = cfg->tolerance;
  init.max_iter = cfg->max_iterations;
  if (pb.katz_blocked) init.fix = katz_initial_fix(cfg->alpha, pb.indeg_max, n, cfg->beta);
  PowerState last{};
  if (cfg->max_iterations == 0) {  // the loop never runs: x = beta, diff = +inf
    *info = gops_run_info{0, std::numeric_limits<double>::infinity(), 0, 0};
  } else {
    if ((st = drive(ctx, g, kKatz, init, last)) != GOPS_OK) return st;
    *info = gops_run_info{last.iter, last.diff, last.converged, 0};
  }
  ctx->stats.sweeps_fp64 = info->iterations;
  if (pb.katz_blocked) ctx->stats.sweeps_blocked = info->iterations;
  return download(ctx, g, pb.v[0], scores);
}

}  // namespace

extern "C" {

gops_status gops_katz_centrality_run(gops_context *ctx, gops_graph *g,
                                     const gops_katz_config *cfg, double *scores,
                                     gops_run_info *info) {
  return katz_run(ctx, g, cfg, scores, info);
}

gops_status gops_katz_centrality_checked_run(gops_context *ctx, gops_graph *g,
                                             const gops_katz_config *cfg, double *scores,
                                             gops_run_info *info) {
  if (!ctx || !cfg) return fail(ctx, GOPS_INVALID_PARAMETER, "null argument");
  const gops_status st = validate_katz(ctx, cfg);
  if (st != GOPS_OK) return st;
  return katz_run(ctx, g, cfg, scores, info);
}

gops_status gops_eigenvector_centrality_run(gops_context *ctx, gops_graph *g,
                                            uint64_t max_iterations, double tolerance,
                                            double *scores, gops_run_info *info) {
  gops_status st = check_args(ctx, g);
  if (st != GOPS_OK) return st;
  gops_run_info local{};
  if (!info) info = &local;
  const uint64_t n = g->dev.n;
  if (n == 0) {  // eigenvector.rs:55-62
    *info = gops_run_info{0, 0.0, 1, 0};
    return GOPS_OK;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::string err;
  if ((st = build_forward(ctx, g, err)) != GOPS_OK) return fail(ctx, st, err);
  if ((st = alloc_power(ctx, g)) != GOPS_OK) return st;
  if ((st = choose_power_blocked(ctx, g, false)) != GOPS_OK) return st;
  const double init_v = 1.0 / std::sqrt((double)n);  // eigenvector.rs:65
  HIP_TRY(ctx, launch_power_init(ctx, g, kEigen, init_v, 0.0));
  PowerState init{};
  init.tolerance = tolerance;
  init.max_iter = max_iterations;
  PowerState last{};
  if (max_iterations == 0) {  // eigenvector.rs:64-69: the start vector, diff 0.0
    *info = gops_run_info{0, 0.0, 0, 0};
  } else {
    if ((st = drive(ctx, g, kEigen, init, last)) != GOPS_OK) return st;
    *info = gops_run_info{last.iter, last.diff, last.converged, 0};
  }
  ctx->stats.sweeps_fp64 = info->iterations;
  if (g->pow.power_blocked) ctx->stats.sweeps_blocked = info->iterations;
  return download(ctx, g, g->pow.v[0], scores);
}

gops_status gops_hits(gops_context *ctx, gops_graph *g, uint64_t max_iter, double tol,
                      double *hub, double *authority, gops_run_info *info) {
  gops_status st = check_args(ctx, g);
  if (st != GOPS_OK) return st;
  gops_run_info local{};
  if (!info) info = &local;
  const uint64_t n = g->dev.n;
  if (n == 0) {  // centrality.rs:116-118
    *info = gops_run_info{0, 0.0, 1, 0};
    return GOPS_OK;
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::string err;
  if ((st = build_forward(ctx, g, err)) != GOPS_OK) return fail(ctx, st, err);
  if ((st = alloc_power(ctx, g)) != GOPS_OK) return st;
  if ((st = choose_power_blocked(ctx, g, true)) != GOPS_OK) return st;
  const double init_v = 1.0 / std::sqrt((double)n);  // centrality.rs:130-132
  HIP_TRY(ctx, launch_power_init(ctx, g, kHits, init_v, 0.0));
  PowerState init{};
  init.tolerance = tol;
  init.max_iter = max_iter;
  PowerState last{};
  if (max_iter == 0) {
    *info = gops_run_info{0, 0.0, 0, 0};
  } else {
    if ((st = drive(ctx, g, kHits, init, last)) != GOPS_OK) return st;
    *info = gops_run_info{last.iter, last.diff, last.converged, 0};
  }
  ctx->stats.sweeps_fp64 = 2 * info->iterations;
  if (g->pow.power_blocked) ctx->stats.sweeps_blocked = 2 * info->iterations;
  if ((st = download(ctx, g, g->pow.v[0], hub)) != GOPS_OK) return st;
  return download(ctx, g, g->pow.v[2], authority);
}

}  // extern "C"
