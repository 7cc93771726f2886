// server.cpp -- batched personalized PageRank entry point and the serving loop.
#include "internal.h"

using namespace gops;

extern "C" {

gops_status gops_ppr_batch_run(gops_context *ctx, gops_graph *g, uint32_t num_queries,
                               const gops_pagerank_config *cfgs, const uint64_t *seed_offsets,
                               const uint32_t *seeds, const double *seed_weights,
                               double *scores, gops_run_info *infos) {
  if (!ctx || !g || (num_queries && (!cfgs || !seed_offsets))) return GOPS_INVALID_PARAMETER;
  return ppr_batch_looped(ctx, g, num_queries, cfgs, seed_offsets, seeds, seed_weights, scores,
                          infos);
}

gops_status gops_server_create(gops_context *, gops_graph *, const gops_server_config *,
                               gops_server **out) {
  if (out) *out = nullptr;
  return GOPS_NOT_SUPPORTED;
}
void gops_server_destroy(gops_server *) {}
gops_status gops_server_submit(gops_server *, const gops_pagerank_config *, const uint32_t *,
                               uint64_t, const double *, uint64_t *) {
  return GOPS_NOT_SUPPORTED;
}
gops_status gops_server_step(gops_server *, uint32_t, uint32_t *) { return GOPS_NOT_SUPPORTED; }
gops_status gops_server_poll(gops_server *, int *, gops_server_result *, double *, uint32_t *,
                             double *) {
  return GOPS_NOT_SUPPORTED;
}
gops_status gops_server_stats(const gops_server *, uint64_t *, uint64_t *, uint64_t *,
                              uint32_t *) {
  return GOPS_NOT_SUPPORTED;
}

}  // extern "C"
