// host.cpp -- host-side graph preprocessing and synthetic inputs.
//
// The reference materialises neighbour lists and out-degree / out-weight sums
// on every call (pagerank.rs:136-137, 226-234; ppr.rs:112, 133).  Here that
// happens once per graph: the adapter's neighbour lists are transposed into a
// CSR of in-edges (pull form), rows are renumbered by in-degree (descending;
// among the rows without in-edges, the ones without out-edges last; ties by
// id) so that rows of similar length sit together and the isolated rows form
// the tail, and the per-row work descriptors of the sweep kernels are
// planned.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.h"

namespace gops {

uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / NaN: keep class, quiet NaN
    uint16_t h = (uint16_t)(u >> 16);
    if (u & 0x007fffffu) h |= 0x40;
    return h;
  }
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;  // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}

static inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

gops_status build_csrt(uint64_t n, const uint64_t *off, const uint32_t *nb, const double *ew,
                       HostCSRT &g, std::string &err) {
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
  int64_t bad = -1;
#pragma omp parallel for schedule(static) reduction(max : bad)
  for (int64_t j = 0; j < (int64_t)m; j++) {
    if (nb[j] >= n) bad = std::max(bad, (int64_t)nb[j]);
  }
  if (bad >= 0) {
    err = "index out of bounds: " + std::to_string(bad);
    return GOPS_INDEX_OUT_OF_BOUNDS;
  }
  g.n = n;
  g.m = m;
  g.has_weights = ew != nullptr;

  // out-degree (pagerank.rs:137: neighbour entries, duplicates included) and
  // out_wsum (pagerank.rs:227-234: sum of max(w, 0) in neighbour order).
  std::vector<double> inv_deg_o(n), inv_ws_o;
  if (ew) inv_ws_o.resize(n);
  int64_t first_bad = INT64_MAX;
#pragma omp parallel for schedule(dynamic, 1024) reduction(min : first_bad)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    uint64_t deg = off[u + 1] - off[u];
    inv_deg_o[u] = deg ? 1.0 / (double)deg : 0.0;
    if (ew) {
      double s = 0.0;
      for (uint64_t j = off[u]; j < off[u + 1]; j++) {
        double w = ew[j];
        if (!std::isfinite(w) || w < 0.0) first_bad = std::min(first_bad, (int64_t)j);
        s += std::fmax(w, 0.0);
      }
      inv_ws_o[u] = (s != 0.0) ? 1.0 / s : 0.0;
    }
  }
  g.first_bad_weight = 0;
  if (first_bad != INT64_MAX) g.first_bad_weight = std::isfinite(ew[first_bad]) ? 2 : 1;

  // in-degree histogram
  std::vector<uint32_t> indeg(n, 0);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)m; j++) __atomic_fetch_add(&indeg[nb[j]], 1u, __ATOMIC_RELAXED);

  // order: in-degree descending, then (in-degree 0) nodes with out-edges
  // before isolated ones, ties by ascending id (counting sort over the key
  // 2 (maxdeg - indeg) + isolated)
  uint32_t maxdeg = 0;
  for (uint64_t v = 0; v < n; v++) maxdeg = std::max(maxdeg, indeg[v]);
  auto key = [&](uint64_t v) {
    return 2 * (size_t)(maxdeg - indeg[v]) + ((indeg[v] == 0 && off[v + 1] == off[v]) ? 1 : 0);
  };
  std::vector<uint64_t> start(2 * (size_t)maxdeg + 3, 0);
  for (uint64_t v = 0; v < n; v++) start[key(v) + 1]++;
  for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
  g.order.assign(n, 0);
  g.rank.assign(n, 0);
  for (uint64_t v = 0; v < n; v++) {
    uint64_t pos = start[key(v)]++;
    g.order[pos] = (uint32_t)v;
    g.rank[v] = (uint32_t)pos;
  }
  start.clear();
  start.shrink_to_fit();

  g.row_ptr.assign(n + 1, 0);
  for (uint64_t r = 0; r < n; r++) g.row_ptr[r + 1] = g.row_ptr[r] + (int32_t)indeg[g.order[r]];

  // scatter in-edges, then sort every row by (new source id, adapter
  // position): duplicate edges u -> v keep the order in which the
  // reference's push loop adds them (pagerank.rs:154-162)
  g.col.assign(m, 0);
  std::vector<int32_t> src_pos(ew ? m : 0);
  std::vector<int32_t> cursor(g.row_ptr.begin(), g.row_ptr.end() - 1);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    const int32_t su = (int32_t)g.rank[u];
    for (uint64_t j = off[u]; j < off[u + 1]; j++) {
      const uint32_t r = g.rank[nb[j]];
      const int32_t pos = __atomic_fetch_add(&cursor[r], 1, __ATOMIC_RELAXED);
      g.col[pos] = su;
      if (ew) src_pos[pos] = (int32_t)j;
    }
  }
  cursor.clear();
  cursor.shrink_to_fit();
  if (ew) g.w.assign(m, 0.0);
#pragma omp parallel
  {
    std::vector<std::pair<int32_t, int32_t>> tmp;
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < (int64_t)n; r++) {
      const int32_t b = g.row_ptr[r], e = g.row_ptr[r + 1];
      if (!ew) {
        if (e - b > 1) std::sort(g.col.begin() + b, g.col.begin() + e);
        continue;
      }
      tmp.resize(e - b);
      for (int32_t k = b; k < e; k++) tmp[k - b] = {g.col[k], src_pos[k]};
      std::sort(tmp.begin(), tmp.end());
      for (int32_t k = b; k < e; k++) {
        g.col[k] = tmp[k - b].first;
        g.w[k] = std::fmax(ew[tmp[k - b].second], 0.0);
      }
    }
  }

  g.inv_deg.assign(n, 0.0);
  if (ew) g.inv_ws.assign(n, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < (int64_t)n; r++) {
    g.inv_deg[r] = inv_deg_o[g.order[r]];
    if (ew) g.inv_ws[r] = inv_ws_o[g.order[r]];
  }
  return GOPS_OK;
}

static inline int lanes_for(int32_t deg) {
  int want = (deg + 3) / 4;
  int g = 1;
  while (g < want && g < 64) g <<= 1;
  return g;
}

void plan_blocks(const HostCSRT &g, Plan &p, const PlanParams &pp) {
  p.desc.clear();
  p.hub_row.clear();
  p.hub_nchunks.clear();
  p.hub_first.clear();
  const int64_t n = (int64_t)g.n;
  const int32_t *rp = g.row_ptr.data();
  int64_t i = 0;
  int32_t chunks = 0;
  while (i < n) {
    const int32_t deg = rp[i + 1] - rp[i];
    if (deg > pp.row_min && deg <= pp.hub_chunk) {
      p.desc.push_back({kRowFlag, (int32_t)i, 0, 0});
      i++;
      continue;
    }
    if (deg > pp.hub_chunk) {
      const int32_t h = (int32_t)p.hub_row.size();
      const int32_t nch = (deg + pp.hub_chunk - 1) / pp.hub_chunk;
      p.hub_row.push_back((int32_t)i);
      p.hub_nchunks.push_back(nch);
      p.hub_first.push_back(chunks);
      for (int32_t c = 0; c < nch; c++) {
        const int32_t b = rp[i] + c * pp.hub_chunk;
        const int32_t e = std::min(b + pp.hub_chunk, rp[i + 1]);
        p.desc.push_back({kHubFlag | c, h, b, e});
      }
      chunks += nch;
      i++;
      continue;
    }
    const int G = lanes_for(deg);
    const int groups = kBlock / G;
    const int64_t max_rows = (int64_t)groups * pp.row_iters;
    const int64_t s = i;
    int64_t nnz = 0;
    while (i < n && i - s < max_rows) {
      const int32_t d = rp[i + 1] - rp[i];
      if (d > pp.row_min || lanes_for(d) != G) break;
      // the run's edges must fit the LDS stage, and with the lanes before
      // the run's first edge (its offset within 64 edges, where the sweep
      // kernels' waves start) one pass of the stage's loads; a first row
      // always fits the stage and may need a second pass
      if (nnz > 0 && (rp[s] & 63) + nnz + d > kStageNnz) break;
      nnz += d;
      i++;
    }
    int lg = 0;
    while ((1 << lg) < G) lg++;
    p.desc.push_back({lg, (int32_t)s, (int32_t)i, 0});
  }
  p.num_hubs = (int)p.hub_row.size();
}

// ---------------------------------------------------------------------------
// Synthetic inputs.  Counter-based RNG: SplitMix64's output function applied
// to (mixed seed + counter * golden gamma), so every draw is a pure function
// of (seed, counter) and generation is reproducible under any thread count.

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rng64(uint64_t seed, uint64_t ctr) {
  return mix64(mix64(seed) + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}
static inline double unit53(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

static double round_weight(double w, int rounding) {
  if (rounding == GOPS_ROUND_F32) return (double)(float)w;
  if (rounding == GOPS_ROUND_BF16) return (double)bf16_to_f32(f32_to_bf16((float)w));
  return w;
}

static double draw_weight(uint64_t seed, uint32_t u, uint32_t v, int kind, double p0, double p1,
                          int rounding) {
  const uint64_t key = ((uint64_t)u << 32) | v;
  const uint64_t ws = seed ^ 0x5bd1e9955bd1e995ull;
  double w;
  if (kind == GOPS_GEN_UNIFORM) {
    w = p0 + (p1 - p0) * unit53(rng64(ws, key));
  } else {  // log-normal by Box-Muller
    const double u1 = ((double)(rng64(ws, key) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = unit53(rng64(ws ^ 0x2545F4914F6CDD1Dull, key));
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    w = std::exp(p0 + p1 * z);
  }
  return round_weight(w, rounding);
}

}  // namespace gops

using namespace gops;

extern "C" gops_status gops_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                                     double c, uint64_t seed, int weight_kind, double p0,
                                     double p1, int rounding, gops_edges **out) {
  if (!out) return GOPS_INVALID_PARAMETER;
  *out = nullptr;
  if (scale == 0 || scale > 30 || edge_factor == 0) return GOPS_INVALID_PARAMETER;
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return GOPS_INVALID_PARAMETER;
  const uint64_t n = 1ull << scale;
  const uint64_t E = (uint64_t)edge_factor << scale;
  if (E >= (1ull << 32)) return GOPS_NOT_SUPPORTED;
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)(a * two32);
  const uint64_t tab = (uint64_t)((a + b) * two32);
  const uint64_t tabc = (uint64_t)((a + b + c) * two32);
  const uint32_t halves = (scale + 1) / 2;
  std::vector<uint32_t> src(E), dst(E);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)E; i++) {
    uint32_t s = 0, d = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; l++) {
      if ((l & 1) == 0) r = rng64(seed, (uint64_t)i * halves + (l >> 1));
      const uint64_t x = (l & 1) ? (r >> 32) : (r & 0xffffffffull);
      uint32_t sb, db;
      if (x < ta) { sb = 0; db = 0; }
      else if (x < tab) { sb = 0; db = 1; }
      else if (x < tabc) { sb = 1; db = 0; }
      else { sb = 1; db = 1; }
      s |= sb << (scale - 1 - l);
      d |= db << (scale - 1 - l);
    }
    src[i] = s;
    dst[i] = d;
  }
  // bucket by source, then sort + unique every bucket; drop self-loops
  std::vector<uint64_t> cnt(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)E; i++)
    if (src[i] != dst[i]) __atomic_fetch_add(&cnt[src[i] + 1], 1ull, __ATOMIC_RELAXED);
  for (uint64_t u = 0; u < n; u++) cnt[u + 1] += cnt[u];
  std::vector<uint32_t> bucket(cnt[n]);
  {
    std::vector<uint64_t> cur(cnt.begin(), cnt.end() - 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)E; i++) {
      if (src[i] == dst[i]) continue;
      const uint64_t pos = __atomic_fetch_add(&cur[src[i]], 1ull, __ATOMIC_RELAXED);
      bucket[pos] = dst[i];
    }
  }
  std::vector<uint32_t>().swap(src);
  std::vector<uint32_t>().swap(dst);
  std::vector<uint64_t> uniq(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    auto bb = bucket.begin() + cnt[u], be = bucket.begin() + cnt[u + 1];
    std::sort(bb, be);
    uniq[u + 1] = (uint64_t)(std::unique(bb, be) - bb);
  }
  auto *e = new gops_edges;
  e->n = n;
  e->off.assign(n + 1, 0);
  for (uint64_t u = 0; u < n; u++) e->off[u + 1] = e->off[u] + uniq[u + 1];
  e->m = e->off[n];
  e->nb.resize(e->m);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; u++)
    std::copy(bucket.begin() + cnt[u], bucket.begin() + cnt[u] + uniq[u + 1],
              e->nb.begin() + e->off[u]);
  if (weight_kind != GOPS_GEN_UNIT) {
    e->w.resize(e->m);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t u = 0; u < (int64_t)n; u++)
      for (uint64_t j = e->off[u]; j < e->off[u + 1]; j++)
        e->w[j] = draw_weight(seed, (uint32_t)u, e->nb[j], weight_kind, p0, p1, rounding);
  }
  *out = e;
  return GOPS_OK;
}

extern "C" gops_status gops_gen_erdos_renyi(uint64_t n, double p, uint64_t seed,
                                            gops_edges **out) {
  if (!out) return GOPS_INVALID_PARAMETER;
  *out = nullptr;
  if (n >= (1ull << 31) || !(p >= 0.0 && p <= 1.0)) return GOPS_INVALID_PARAMETER;
  std::vector<std::vector<uint32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    for (uint64_t v = 0; v < n; v++) {
      if (v == (uint64_t)u) continue;
      if (unit53(rng64(seed, (uint64_t)u * n + v)) < p) rows[u].push_back((uint32_t)v);
    }
  }
  auto *e = new gops_edges;
  e->n = n;
  e->off.assign(n + 1, 0);
  for (uint64_t u = 0; u < n; u++) e->off[u + 1] = e->off[u] + rows[u].size();
  e->m = e->off[n];
  e->nb.resize(e->m);
  for (uint64_t u = 0; u < n; u++) std::copy(rows[u].begin(), rows[u].end(), e->nb.begin() + e->off[u]);
  *out = e;
  return GOPS_OK;
}

extern "C" gops_status gops_gen_seed_sets(const gops_edges *e, uint32_t num_queries,
                                          uint32_t min_seeds, uint32_t max_seeds,
                                          int log_uniform, uint64_t seed, uint64_t *offsets,
                                          uint32_t *seeds, uint64_t capacity) {
  if (!e || !offsets || !seeds || min_seeds == 0 || max_seeds < min_seeds || e->m == 0)
    return GOPS_INVALID_PARAMETER;
  uint64_t ctr = 0, pos = 0;
  offsets[0] = 0;
  std::vector<uint32_t> chosen;
  for (uint32_t q = 0; q < num_queries; q++) {
    uint32_t s;
    if (log_uniform) {
      const double lo = std::log((double)min_seeds), hi = std::log((double)max_seeds + 1.0);
      s = (uint32_t)std::floor(std::exp(lo + (hi - lo) * unit53(rng64(seed, ctr++))));
      s = std::min(std::max(s, min_seeds), max_seeds);
    } else {
      s = min_seeds + (uint32_t)(rng64(seed, ctr++) % (uint64_t)(max_seeds - min_seeds + 1));
    }
    chosen.clear();
    uint64_t attempts = 0;
    while (chosen.size() < s && attempts < 64ull * s) {
      attempts++;
      const uint64_t edge = rng64(seed, ctr++) % e->m;
      const uint32_t u = (uint32_t)(std::upper_bound(e->off.begin(), e->off.end(), edge) -
                                    e->off.begin() - 1);
      if (std::find(chosen.begin(), chosen.end(), u) == chosen.end()) chosen.push_back(u);
    }
    if (pos + chosen.size() > capacity) return GOPS_DIMENSION_MISMATCH;
    std::copy(chosen.begin(), chosen.end(), seeds + pos);
    pos += chosen.size();
    offsets[q + 1] = pos;
  }
  return GOPS_OK;
}

extern "C" gops_status gops_edges_view(const gops_edges *e, uint64_t *n, uint64_t *m,
                                       const uint64_t **offsets, const uint32_t **targets,
                                       const double **weights) {
  if (!e) return GOPS_INVALID_PARAMETER;
  if (n) *n = e->n;
  if (m) *m = e->m;
  if (offsets) *offsets = e->off.data();
  if (targets) *targets = e->nb.data();
  if (weights) *weights = e->w.empty() ? nullptr : e->w.data();
  return GOPS_OK;
}

extern "C" void gops_edges_destroy(gops_edges *e) { delete e; }
