// host_check.cpp -- a CPU-only driver for the host code of host.cpp (the
// synthetic generators, the CSR transpose build, the sweeps' row planner and
// the bf16 conversion), built with AddressSanitizer and UndefinedBehavior-
// Sanitizer by `make -C graphops_amd/csrc sanitize`.  It makes no HIP call.
// Every check that fails prints its name; the exit status is the number of
// failures.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../internal.h"

using namespace gops;

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
      failures++;                                                     \
    }                                                                 \
  } while (0)

// The CSR transpose of (off, nb): row r (new id) lists the new ids of the
// sources of node order[r]'s in-edges, ascending; rows by descending in-degree,
// the isolated ones (no in-, no out-edges) after the other rows without in-edges.
static void check_csrt(uint64_t n, const uint64_t *off, const uint32_t *nb, const HostCSRT &h) {
  const uint64_t m = off[n];
  CHECK(h.n == n && h.m == m);
  CHECK(h.row_ptr.size() == n + 1 && h.row_ptr[0] == 0 && (uint64_t)h.row_ptr[n] == m);
  std::vector<uint64_t> indeg(n, 0);
  for (uint64_t j = 0; j < m; j++) indeg[nb[j]]++;
  for (uint64_t r = 0; r < n; r++) {
    CHECK(h.order[r] < n && h.rank[h.order[r]] == r);
    CHECK((uint64_t)(h.row_ptr[r + 1] - h.row_ptr[r]) == indeg[h.order[r]]);
    if (r) CHECK(h.row_ptr[r + 1] - h.row_ptr[r] <= h.row_ptr[r] - h.row_ptr[r - 1]);
    if (r && indeg[h.order[r]] == 0 && off[h.order[r] + 1] != off[h.order[r]]) {
      const uint64_t p = h.order[r - 1];  // a row with out-edges follows no isolated row
      CHECK(!(indeg[p] == 0 && off[p + 1] == off[p]));
    }
    for (int32_t k = h.row_ptr[r]; k < h.row_ptr[r + 1]; k++) {
      CHECK(h.col[k] >= 0 && (uint64_t)h.col[k] < n);
      if (k > h.row_ptr[r]) CHECK(h.col[k] >= h.col[k - 1]);
    }
    const uint64_t u = h.order[r];
    const uint64_t deg = off[u + 1] - off[u];
    CHECK(h.inv_deg[r] == (deg ? 1.0 / (double)deg : 0.0));
  }
}

// Every row is in exactly one descriptor: a run of rows, a long row, or the
// chunks of a hub row, which cover its edges once, in order.
static void check_plan(const HostCSRT &h, const PlanParams &pp) {
  Plan p;
  plan_blocks(h, p, pp);
  const int64_t n = (int64_t)h.n;
  std::vector<int> seen(n, 0);
  int32_t chunks = 0;
  for (const BlockDesc &d : p.desc) {
    if (d.x & kHubFlag) {
      const int32_t hub = d.y, c = d.x & (kHubFlag - 1);
      CHECK(hub >= 0 && hub < p.num_hubs);
      if (hub < 0 || hub >= p.num_hubs) continue;
      const int32_t r = p.hub_row[hub];
      CHECK(d.z == h.row_ptr[r] + c * pp.hub_chunk);
      CHECK(d.w == std::min(d.z + pp.hub_chunk, h.row_ptr[r + 1]));
      if (c == 0) seen[r]++;
      chunks++;
    } else if (d.x & kRowFlag) {
      const int32_t deg = h.row_ptr[d.y + 1] - h.row_ptr[d.y];
      CHECK(deg > pp.row_min && deg <= pp.hub_chunk);
      seen[d.y]++;
    } else {
      CHECK(d.y < d.z && d.z <= n && d.x >= 0 && d.x <= 6);
      for (int32_t r = d.y; r < d.z; r++) {
        CHECK(h.row_ptr[r + 1] - h.row_ptr[r] <= pp.row_min);
        seen[r]++;
      }
      CHECK(d.z - d.y <= (kBlock >> d.x) * pp.row_iters);
      if (d.z - d.y > 1)
        CHECK((h.row_ptr[d.y] & 63) + (h.row_ptr[d.z] - h.row_ptr[d.y]) <= kStageNnz);
    }
  }
  for (int64_t r = 0; r < n; r++) CHECK(seen[r] == 1);
  int32_t total = 0;
  for (int32_t c : p.hub_nchunks) total += c;
  CHECK(total == chunks && (int)p.hub_row.size() == p.num_hubs);
}

int main() {
  // R-MAT, log-normal weights rounded to bf16
  gops_edges *e = nullptr;
  CHECK(gops_gen_rmat(12, 16, 0.57, 0.19, 0.19, 11, GOPS_GEN_LOGNORMAL, 0.0, 1.0, GOPS_ROUND_BF16,
                      &e) == GOPS_OK);
  uint64_t n = 0, m = 0;
  const uint64_t *off = nullptr;
  const uint32_t *nb = nullptr;
  const double *w = nullptr;
  CHECK(gops_edges_view(e, &n, &m, &off, &nb, &w) == GOPS_OK);
  CHECK(n == 4096 && m > 0 && off[n] == m);
  HostCSRT h;
  std::string err;
  CHECK(build_csrt(n, off, nb, w, h, err) == GOPS_OK);
  check_csrt(n, off, nb, h);
  for (const PlanParams &pp : {PlanParams{}, PlanParams{64, 64, 1}, PlanParams{512, 16384, 4},
                               PlanParams{1024, 32768, 16}})
    check_plan(h, pp);
  // a graph with a hub row of 70 000 in-edges and a long row
  {
    const uint64_t hn = 80000;
    std::vector<uint64_t> hoff(hn + 1, 0);
    std::vector<uint32_t> hnb;
    for (uint64_t u = 0; u < hn; u++) {
      if (u >= 1 && u <= 70000) hnb.push_back(0);
      if (u >= 2 && u <= 5001) hnb.push_back(1);
      hnb.push_back((uint32_t)((u * 7919 + 3) % hn));
      hoff[u + 1] = hnb.size();
    }
    HostCSRT hh;
    CHECK(build_csrt(hn, hoff.data(), hnb.data(), nullptr, hh, err) == GOPS_OK);
    check_csrt(hn, hoff.data(), hnb.data(), hh);
    for (const PlanParams &pp : {PlanParams{}, PlanParams{64, 64, 1}, PlanParams{4096, 16384, 8}})
      check_plan(hh, pp);
  }
  // the adapter's errors: bad offsets, a neighbour out of range
  {
    const uint64_t boff[3] = {0, 2, 1};
    const uint32_t bnb[2] = {1, 0};
    HostCSRT hb;
    CHECK(build_csrt(2, boff, bnb, nullptr, hb, err) == GOPS_INVALID_PARAMETER);
    const uint64_t goff[3] = {0, 1, 2};
    const uint32_t gnb[2] = {1, 5};
    CHECK(build_csrt(2, goff, gnb, nullptr, hb, err) == GOPS_INDEX_OUT_OF_BOUNDS);
    CHECK(err == "index out of bounds: 5");
  }
  // seed sets: sizes in range, distinct nodes per query
  {
    const uint32_t nq = 64, maxs = 32;
    std::vector<uint64_t> so(nq + 1);
    std::vector<uint32_t> seeds(nq * maxs);
    CHECK(gops_gen_seed_sets(e, nq, 1, maxs, 1, 5, so.data(), seeds.data(), seeds.size()) ==
          GOPS_OK);
    for (uint32_t q = 0; q < nq; q++) {
      const uint64_t k = so[q + 1] - so[q];
      CHECK(k >= 1 && k <= maxs);
      for (uint64_t i = so[q]; i < so[q + 1]; i++) {
        CHECK(seeds[i] < n);
        for (uint64_t j = so[q]; j < i; j++) CHECK(seeds[j] != seeds[i]);
      }
    }
  }
  gops_edges_destroy(e);
  // Erdos-Renyi: no self-loops, targets ascending per source
  {
    gops_edges *er = nullptr;
    CHECK(gops_gen_erdos_renyi(2000, 0.005, 42, &er) == GOPS_OK);
    CHECK(gops_edges_view(er, &n, &m, &off, &nb, &w) == GOPS_OK);
    for (uint64_t u = 0; u < n; u++)
      for (uint64_t j = off[u]; j < off[u + 1]; j++) {
        CHECK(nb[j] != u && nb[j] < n);
        if (j > off[u]) CHECK(nb[j] > nb[j - 1]);
      }
    gops_edges_destroy(er);
  }
  // bf16 storage rounding: exact values stay, ties to even, specials kept
  CHECK(f32_to_bf16(1.0f) == 0x3f80);
  CHECK(f32_to_bf16(-2.0f) == 0xc000);
  CHECK(f32_to_bf16(std::numeric_limits<float>::infinity()) == 0x7f80);
  CHECK((f32_to_bf16(std::nanf("")) & 0x7f80) == 0x7f80 && (f32_to_bf16(std::nanf("")) & 0x7f));
  std::printf("host_check: %d failure(s)\n", failures);
  return failures;
}
