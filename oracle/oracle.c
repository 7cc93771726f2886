/*
 * oracle.c -- CPU f64 restatement of graphops' PageRank family.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke check and the benchmark's cpu_baseline leg call it to check (and, for
 * the baseline, to time) the HIP product path.  Nothing in graphops_amd/
 * links, loads or calls it; the product fails loudly without its own HIP
 * library instead of falling back here.
 *
 * Reference: crate graphops v0.4.1 (Rust).  Every function cites the lines it
 * restates.  The restatement keeps the reference's operation order so that the
 * f64 results are bit-for-bit what the reference computes:
 *   share     = (d * x_u) / deg_u                      pagerank.rs:157
 *   share_w   = (d * x_u) * (w / ws_u)                 pagerank.rs:259
 *   fill (PR) = (1 - d)/n + (d * dsum)/n               pagerank.rs:150-152
 *   fill (PPR)= (1 - d) * p_i + (d * dsum) * p_i       ppr.rs:126-127
 *   dsum, diff: sequential folds in index order         pagerank.rs:144-149,169-173
 *   stop      : diff < tol (strict), checked after the swap  pagerank.rs:174-179
 * Compile with -O2 -ffp-contract=off (no FMA contraction, no fast-math).
 *
 * Katz, eigenvector and HITS (katz.rs, eigenvector.rs, centrality.rs) are
 * restated in the same way: the reference's loops and operation order,
 * sequential sums in index order.
 *
 * Parity status (pins live in tests/test_oracle.py):
 *   orc_pagerank_run, orc_pagerank_weighted_run, orc_ppr_run: pinned by exact
 *     rational fixed points, exact-rational iteration trajectories on small
 *     graphs, dense linear solves and networkx fixed points.
 *   orc_*_pull_run: pinned bit-for-bit against the push forms above.
 *   orc_katz_run: pinned by exact-rational trajectories, the linear solve
 *     beta (I - alpha A^T)^-1 1 and the reference's unit tests (katz.rs:231-329).
 *   orc_top_k: pinned by the reference's doc example (topk.rs:9-16), by brute
 *     force on tie-free inputs and, at ties, by a closed description of the
 *     entries the heap keeps (tests/test_oracle_topk.py).
 *   orc_eigenvector_run, orc_hits: pinned by 60-digit decimal trajectories of
 *     the same iterations, the dominant eigenvector of A + I (numpy) and the
 *     top singular vectors of A (numpy SVD), and the reference's unit tests
 *     (eigenvector.rs:145-197, centrality.rs:244-273).
 *   The reference's optional `simd` L1 residual (third-party innr 0.4,
 *     pagerank.rs:166-167) is NOT restated: parity unpinned for that feature;
 *     the default scalar residual is what is restated here.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_INVALID 1

static void put_msg(char *msg, size_t cap, const char *text) {
    if (msg && cap) {
        snprintf(msg, cap, "%s", text);
    }
}

/* PageRankConfig::validate, pagerank.rs:63-85.  Returns ORC_INVALID with the
 * reference's message text (the Error::InvalidParameter payload). */
int orc_validate_config(double damping, uint64_t max_iterations, double tolerance,
                        char *msg, size_t cap) {
    if (!isfinite(damping)) {
        put_msg(msg, cap, "damping must be finite");
        return ORC_INVALID;
    }
    if (!(damping >= 0.0 && damping <= 1.0)) {
        put_msg(msg, cap, "damping must be in [0,1]");
        return ORC_INVALID;
    }
    if (max_iterations == 0) {
        put_msg(msg, cap, "max_iterations must be > 0");
        return ORC_INVALID;
    }
    if (!isfinite(tolerance) || tolerance <= 0.0) {
        put_msg(msg, cap, "tolerance must be finite and > 0");
        return ORC_INVALID;
    }
    return ORC_OK;
}

/* Weight checks of pagerank_weighted_checked, pagerank.rs:300-314: every
 * neighbour entry's edge_weight must be finite and >= 0. */
int orc_check_weights(uint64_t m, const double *ew, char *msg, size_t cap) {
    for (uint64_t j = 0; j < m; j++) {
        double w = ew[j];
        if (!isfinite(w)) {
            put_msg(msg, cap, "edge weights must be finite");
            return ORC_INVALID;
        }
        if (w < 0.0) {
            put_msg(msg, cap, "edge weights must be non-negative");
            return ORC_INVALID;
        }
    }
    return ORC_OK;
}

/* Personalization checks of personalized_pagerank_checked, ppr.rs:17-42. */
int orc_check_personalization(uint64_t n, const double *p, uint64_t p_len,
                              char *msg, size_t cap) {
    if (p_len != n) {
        if (msg && cap) {
            snprintf(msg, cap,
                     "personalization length must equal node_count (len=%llu node_count=%llu)",
                     (unsigned long long)p_len, (unsigned long long)n);
        }
        return ORC_INVALID;
    }
    for (uint64_t i = 0; i < p_len; i++) {
        if (!isfinite(p[i])) {
            put_msg(msg, cap, "personalization entries must be finite");
            return ORC_INVALID;
        }
        if (p[i] < 0.0) {
            put_msg(msg, cap, "personalization entries must be non-negative");
            return ORC_INVALID;
        }
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < p_len; i++) sum += p[i];
    if (sum <= 0.0) {
        put_msg(msg, cap, "personalization sum must be > 0");
        return ORC_INVALID;
    }
    return ORC_OK;
}

/* AdjacencyMatrix adapter, graph.rs:122-142: neighbours of u are the columns
 * v with m[u][v] > 0.0 (NaN and negatives excluded), in column order;
 * edge_weight(u, v) is the raw entry.  Fills off[n+1]; when nb/ew are
 * non-NULL also the neighbour ids and their edge weights.  Returns m. */
uint64_t orc_dense_neighbors(uint64_t n, const double *mat, uint64_t *off,
                             uint32_t *nb, double *ew) {
    uint64_t m = 0;
    for (uint64_t u = 0; u < n; u++) {
        off[u] = m;
        const double *row = mat + u * n;
        for (uint64_t v = 0; v < n; v++) {
            if (row[v] > 0.0) {
                if (nb) nb[m] = (uint32_t)v;
                if (ew) ew[m] = row[v];
                m++;
            }
        }
    }
    off[n] = m;
    return m;
}

static void run_empty(double *diff, uint64_t *iters, int *conv) {
    /* n == 0 early return, pagerank.rs:124-131 / ppr.rs:97-103 */
    *iters = 0;
    *diff = 0.0;
    *conv = 1;
}

static double l1_seq(uint64_t n, const double *a, const double *b) {
    /* default (non-simd) residual, pagerank.rs:169-173 */
    double s = 0.0;
    for (uint64_t i = 0; i < n; i++) s += fabs(a[i] - b[i]);
    return s;
}

/* pagerank_run, pagerank.rs:122-187 (push form, neighbour lists in adapter
 * order).  scores[n] receives the returned PageRankRun::scores. */
void orc_pagerank_run(uint64_t n, const uint64_t *off, const uint32_t *nb,
                      double damping, uint64_t max_iterations, double tolerance,
                      double *scores, uint64_t *iters_out, double *diff_out,
                      int *conv_out) {
    if (n == 0) {
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
    double n_f64 = (double)n;
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; i++) cur[i] = 1.0 / n_f64;         /* :133 */
    uint64_t iters = 0;
    double last_diff = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) {             /* :142 */
        iters++;
        double dsum = 0.0;                                         /* :144-149 */
        for (uint64_t i = 0; i < n; i++)
            if (off[i + 1] - off[i] == 0) dsum += cur[i];
        double dangling_contrib = damping * dsum / n_f64;          /* :150 */
        double teleport = (1.0 - damping) / n_f64;                 /* :151 */
        double fill = teleport + dangling_contrib;                 /* :152 */
        for (uint64_t i = 0; i < n; i++) nxt[i] = fill;
        for (uint64_t u = 0; u < n; u++) {                         /* :154-162 */
            uint64_t deg = off[u + 1] - off[u];
            if (deg > 0) {
                double share = damping * cur[u] / (double)deg;
                for (uint64_t j = off[u]; j < off[u + 1]; j++) nxt[nb[j]] += share;
            }
        }
        double diff = l1_seq(n, cur, nxt);                         /* :169-173 */
        last_diff = diff;
        double *t = cur; cur = nxt; nxt = t;                       /* :175 */
        if (diff < tolerance) {                                    /* :176-179 */
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    *iters_out = iters;
    *diff_out = last_diff;
    *conv_out = converged;
}

/* out_wsum, pagerank.rs:227-234: ws_u = sum over neighbour entries of
 * max(edge_weight(u, v), 0) in neighbour order (f64::max returns the non-NaN
 * operand, as fmax does). */
void orc_out_wsum(uint64_t n, const uint64_t *off, const double *ew, double *ws) {
    for (uint64_t u = 0; u < n; u++) {
        double s = 0.0;
        for (uint64_t j = off[u]; j < off[u + 1]; j++) s += fmax(ew[j], 0.0);
        ws[u] = s;
    }
}

/* pagerank_weighted_run, pagerank.rs:210-289 (push form).  ew[j] is
 * edge_weight(u, nb[j]) for the j-th neighbour entry. */
void orc_pagerank_weighted_run(uint64_t n, const uint64_t *off, const uint32_t *nb,
                               const double *ew, double damping,
                               uint64_t max_iterations, double tolerance,
                               double *scores, uint64_t *iters_out,
                               double *diff_out, int *conv_out) {
    if (n == 0) {
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
    double n_f64 = (double)n;
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    double *ws = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; i++) cur[i] = 1.0 / n_f64;         /* :220 */
    orc_out_wsum(n, off, ew, ws);                                  /* :227-234 */
    uint64_t iters = 0;
    double last_diff = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) {             /* :239 */
        iters++;
        double dsum = 0.0;                                         /* :241-246 */
        for (uint64_t i = 0; i < n; i++)
            if (ws[i] == 0.0) dsum += cur[i];
        double dangling_contrib = damping * dsum / n_f64;          /* :248 */
        double teleport = (1.0 - damping) / n_f64;                 /* :249 */
        double fill = teleport + dangling_contrib;
        for (uint64_t i = 0; i < n; i++) nxt[i] = fill;            /* :250 */
        for (uint64_t u = 0; u < n; u++) {                         /* :252-263 */
            double w_s = ws[u];
            if (w_s > 0.0) {
                for (uint64_t j = off[u]; j < off[u + 1]; j++) {
                    double w = fmax(ew[j], 0.0);
                    if (w > 0.0) nxt[nb[j]] += damping * cur[u] * (w / w_s);
                }
            }
        }
        double diff = l1_seq(n, cur, nxt);                         /* :270-274 */
        last_diff = diff;
        double *t = cur; cur = nxt; nxt = t;
        if (diff < tolerance) {                                    /* :277-280 */
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    free(ws);
    *iters_out = iters;
    *diff_out = last_diff;
    *conv_out = converged;
}

/* personalized_pagerank_run, ppr.rs:90-156 (push form). */
void orc_ppr_run(uint64_t n, const uint64_t *off, const uint32_t *nb,
                 const double *personalization, double damping,
                 uint64_t max_iterations, double tolerance, double *scores,
                 uint64_t *iters_out, double *diff_out, int *conv_out) {
    if (n == 0) {
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
    double *p_vec = (double *)malloc(n * sizeof(double));
    double p_sum = 0.0;                                            /* :104 */
    for (uint64_t i = 0; i < n; i++) p_sum += personalization[i];
    if (p_sum > 0.0) {                                             /* :105-109 */
        for (uint64_t i = 0; i < n; i++) p_vec[i] = personalization[i] / p_sum;
    } else {
        for (uint64_t i = 0; i < n; i++) p_vec[i] = 1.0 / (double)n;
    }
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    memcpy(cur, p_vec, n * sizeof(double));                        /* :110 */
    uint64_t iters = 0;
    double last_diff = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) {             /* :117 */
        iters++;
        double dsum = 0.0;                                         /* :119-124 */
        for (uint64_t i = 0; i < n; i++)
            if (off[i + 1] - off[i] == 0) dsum += cur[i];
        for (uint64_t i = 0; i < n; i++)                           /* :125-128 */
            nxt[i] = (1.0 - damping) * p_vec[i] + damping * dsum * p_vec[i];
        for (uint64_t u = 0; u < n; u++) {                         /* :129-137 */
            uint64_t deg = off[u + 1] - off[u];
            if (deg > 0) {
                double share = damping * cur[u] / (double)deg;
                for (uint64_t j = off[u]; j < off[u + 1]; j++) nxt[nb[j]] += share;
            }
        }
        double diff = l1_seq(n, cur, nxt);                         /* :138-142 */
        last_diff = diff;
        double *t = cur; cur = nxt; nxt = t;
        if (diff < tolerance) {                                    /* :145-148 */
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    free(p_vec);
    *iters_out = iters;
    *diff_out = last_diff;
    *conv_out = converged;
}

/*
 * Pull forms over the CSR transpose (in-edges of v sorted by ascending source,
 * duplicates adjacent).  For a fixed v the push loop above adds the shares of
 * u = 0, 1, ... in increasing u, so pulling them in ascending-source order
 * reproduces the same f64 addition sequence bit-for-bit; rows are independent,
 * so OpenMP over rows keeps that property.  dsum and diff stay sequential.
 *   kind 0: unweighted PageRank (pagerank_run)      scale[u] = deg_u (as f64)
 *   kind 1: weighted PageRank (pagerank_weighted_run) scale[u] = ws_u, w[] given
 *   kind 2: personalized PageRank (personalized_pagerank_run), p[] given
 * w[e] must already be max(edge_weight, 0).  threads <= 0 means all.
 */
void orc_pull_run(int kind, uint64_t n, const uint64_t *in_off, const uint32_t *in_src,
                  const double *w, const double *scale, const double *personalization,
                  double damping, uint64_t max_iterations, double tolerance,
                  int threads, double *scores, uint64_t *iters_out,
                  double *diff_out, int *conv_out) {
    if (n == 0) {
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    threads = 1;
#endif
    double n_f64 = (double)n;
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    double *p_vec = NULL;
    if (kind == 2) {
        p_vec = (double *)malloc(n * sizeof(double));
        double p_sum = 0.0;
        for (uint64_t i = 0; i < n; i++) p_sum += personalization[i];
        if (p_sum > 0.0) {
            for (uint64_t i = 0; i < n; i++) p_vec[i] = personalization[i] / p_sum;
        } else {
            for (uint64_t i = 0; i < n; i++) p_vec[i] = 1.0 / n_f64;
        }
        memcpy(cur, p_vec, n * sizeof(double));
    } else {
        for (uint64_t i = 0; i < n; i++) cur[i] = 1.0 / n_f64;
    }
    uint64_t iters = 0;
    double last_diff = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) {
        iters++;
        double dsum = 0.0;
        for (uint64_t i = 0; i < n; i++)
            if (scale[i] == 0.0) dsum += cur[i];
        double fill = (1.0 - damping) / n_f64 + damping * dsum / n_f64;
        double ddsum = damping * dsum;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads)
        for (int64_t vv = 0; vv < (int64_t)n; vv++) {
            uint64_t v = (uint64_t)vv;
            double acc;
            if (kind == 2)
                acc = (1.0 - damping) * p_vec[v] + ddsum * p_vec[v];
            else
                acc = fill;
            for (uint64_t e = in_off[v]; e < in_off[v + 1]; e++) {
                uint32_t u = in_src[e];
                if (kind == 1) {
                    double we = w[e];
                    if (we > 0.0) acc += damping * cur[u] * (we / scale[u]);
                } else {
                    acc += damping * cur[u] / scale[u];
                }
            }
            nxt[v] = acc;
        }
        double diff = l1_seq(n, cur, nxt);
        last_diff = diff;
        double *t = cur; cur = nxt; nxt = t;
        if (diff < tolerance) {
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    free(p_vec);
    *iters_out = iters;
    *diff_out = last_diff;
    *conv_out = converged;
}

/*
 * The same iteration in extended precision (x87 long double, 64-bit
 * significand): the reference's recurrence (pagerank.rs:133-179, 221-281;
 * ppr.rs:104-149) with every quantity -- out-degree or out_wsum, dangling
 * sum, fill, shares, row sums, residual -- carried in long double, so that
 * its rounding (~1e-19 relative) is far below the f64 reference's own.  It
 * stands for the exact iteration at sizes where fractions cannot run: the
 * f64 reference drifts from it by about n * 2^-53 in total mass, so at
 * tolerances near 1e-12 the two can stop at different sweeps.
 *   kind 0/1/2 as orc_pull_run; w[e] = max(edge_weight, 0) per in-edge, or
 *   NULL for kinds 0 and 2.  The out-weight sums are formed here from the
 *   in-edges (in long double: exact for the sums this is used on).
 * Rows run in parallel; dsum and diff are summed over a fixed partition of
 * 256 index blocks, in block order, so results do not depend on the thread
 * count.
 */
#define LD_BLOCKS 256
static long double ld_block_sum(uint64_t n, const long double *a, const long double *b,
                                const long double *scale, int mode, int threads) {
    long double part[LD_BLOCKS];
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int k = 0; k < LD_BLOCKS; k++) {
        uint64_t lo = n * (uint64_t)k / LD_BLOCKS, hi = n * (uint64_t)(k + 1) / LD_BLOCKS;
        long double s = 0.0L;
        for (uint64_t i = lo; i < hi; i++) {
            if (mode == 0) s += fabsl(a[i] - b[i]);        /* residual */
            else if (scale[i] == 0.0L) s += a[i];           /* dangling sum */
        }
        part[k] = s;
    }
    long double s = 0.0L;
    for (int k = 0; k < LD_BLOCKS; k++) s += part[k];
    return s;
}

void orc_pull_run_ld(int kind, uint64_t n, const uint64_t *in_off, const uint32_t *in_src,
                     const double *w, const double *personalization, double damping,
                     uint64_t max_iterations, double tolerance, int threads, double *scores,
                     uint64_t *iters_out, double *diff_out, int *conv_out) {
    if (n == 0) {
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    threads = 1;
#endif
    const long double d = (long double)damping, nl = (long double)n;
    long double *cur = (long double *)malloc(n * sizeof(long double));
    long double *nxt = (long double *)malloc(n * sizeof(long double));
    long double *scale = (long double *)calloc(n, sizeof(long double));
    long double *y = (long double *)malloc(n * sizeof(long double));
    long double *q = NULL;
    /* out-degree / out_wsum of every source (pagerank.rs:137, 227-234) */
    for (uint64_t v = 0; v < n; v++)
        for (uint64_t e = in_off[v]; e < in_off[v + 1]; e++)
            scale[in_src[e]] += (kind == 1) ? (long double)w[e] : 1.0L;
    if (kind == 2) { /* ppr.rs:104-110 */
        q = (long double *)malloc(n * sizeof(long double));
        long double ps = 0.0L;
        for (uint64_t i = 0; i < n; i++) ps += (long double)personalization[i];
        for (uint64_t i = 0; i < n; i++)
            q[i] = ps > 0.0L ? (long double)personalization[i] / ps : 1.0L / nl;
        memcpy(cur, q, n * sizeof(long double));
    } else {
        for (uint64_t i = 0; i < n; i++) cur[i] = 1.0L / nl;
    }
    uint64_t iters = 0;
    long double last = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) {
        iters++;
        const long double dsum = ld_block_sum(n, cur, NULL, scale, 1, threads);
        const long double fill = (1.0L - d) / nl + d * dsum / nl;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t u = 0; u < (int64_t)n; u++)
            y[u] = scale[u] != 0.0L ? d * cur[u] / scale[u] : 0.0L;
#pragma omp parallel for schedule(dynamic, 4096) num_threads(threads)
        for (int64_t vv = 0; vv < (int64_t)n; vv++) {
            const uint64_t v = (uint64_t)vv;
            long double acc = kind == 2 ? (1.0L - d) * q[v] + d * dsum * q[v] : fill;
            for (uint64_t e = in_off[v]; e < in_off[v + 1]; e++) {
                if (kind == 1) {
                    if (w[e] > 0.0) acc += (long double)w[e] * y[in_src[e]];
                } else {
                    acc += y[in_src[e]];
                }
            }
            nxt[v] = acc;
        }
        const long double diff = ld_block_sum(n, cur, nxt, NULL, 0, threads);
        last = diff;
        long double *t = cur; cur = nxt; nxt = t;
        if (diff < (long double)tolerance) {
            converged = 1;
            break;
        }
    }
    for (uint64_t i = 0; i < n; i++) scores[i] = (double)cur[i];
    free(cur);
    free(nxt);
    free(scale);
    free(y);
    free(q);
    *iters_out = iters;
    *diff_out = (double)last;
    *conv_out = converged;
}

/* One push sweep of pagerank_weighted_run / pagerank_run (pagerank.rs:239-263,
 * 142-162) from a given vector, for timing a bounded sample of the CPU path.
 * Returns the sweep's L1 residual; writes the new vector into out[n]. */
double orc_push_sweep(int weighted, uint64_t n, const uint64_t *off, const uint32_t *nb,
                      const double *ew, const double *ws, double damping,
                      const double *cur, double *out) {
    double n_f64 = (double)n;
    double dsum = 0.0;
    if (weighted) {
        for (uint64_t i = 0; i < n; i++)
            if (ws[i] == 0.0) dsum += cur[i];
    } else {
        for (uint64_t i = 0; i < n; i++)
            if (off[i + 1] - off[i] == 0) dsum += cur[i];
    }
    double fill = (1.0 - damping) / n_f64 + damping * dsum / n_f64;
    for (uint64_t i = 0; i < n; i++) out[i] = fill;
    for (uint64_t u = 0; u < n; u++) {
        if (weighted) {
            double w_s = ws[u];
            if (w_s > 0.0) {
                for (uint64_t j = off[u]; j < off[u + 1]; j++) {
                    double w = fmax(ew[j], 0.0);
                    if (w > 0.0) out[nb[j]] += damping * cur[u] * (w / w_s);
                }
            }
        } else {
            uint64_t deg = off[u + 1] - off[u];
            if (deg > 0) {
                double share = damping * cur[u] / (double)deg;
                for (uint64_t j = off[u]; j < off[u + 1]; j++) out[nb[j]] += share;
            }
        }
    }
    return l1_seq(n, cur, out);
}

/* The scatter of pagerank_weighted_run's sweep (pagerank.rs:252-263) for the
 * sources listed in us[0..nu), with the weight of every neighbour entry
 * looked up the way the petgraph adapter's edge_weight does (graph.rs:
 * 205-211): find_edge(u, v) scans u's edges for the first one to v, so every
 * entry costs a scan of u's list up to v's first occurrence.  For timing the
 * reference's per-edge lookup on a sample of sources; adds into out[n]. */
void orc_push_find_edge(uint64_t nu, const uint32_t *us, const uint64_t *off, const uint32_t *nb,
                        const double *ew, const double *ws, double damping, const double *cur,
                        double *out) {
    for (uint64_t i = 0; i < nu; i++) {
        uint64_t u = us[i];
        double w_s = ws[u];
        if (!(w_s > 0.0)) continue;
        for (uint64_t j = off[u]; j < off[u + 1]; j++) {
            uint32_t v = nb[j];
            uint64_t f = off[u];
            while (nb[f] != v) f++; /* find_edge: the first entry to v */
            double w = fmax(ew[f], 0.0);
            if (w > 0.0) out[v] += damping * cur[u] * (w / w_s);
        }
    }
}

/* Transpose neighbour lists (adapter order) into in-edge lists sorted by
 * ascending source; duplicates keep their adapter order.  Sequential. */
void orc_transpose(uint64_t n, const uint64_t *off, const uint32_t *nb, const double *ew,
                   uint64_t *in_off, uint32_t *in_src, double *in_w) {
    uint64_t m = off[n];
    memset(in_off, 0, (n + 1) * sizeof(uint64_t));
    for (uint64_t j = 0; j < m; j++) in_off[nb[j] + 1]++;
    for (uint64_t v = 0; v < n; v++) in_off[v + 1] += in_off[v];
    uint64_t *cur = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    memcpy(cur, in_off, (n + 1) * sizeof(uint64_t));
    for (uint64_t u = 0; u < n; u++) {
        for (uint64_t j = off[u]; j < off[u + 1]; j++) {
            uint64_t pos = cur[nb[j]]++;
            in_src[pos] = (uint32_t)u;
            if (in_w) in_w[pos] = fmax(ew[j], 0.0);
        }
    }
    free(cur);
}

/* KatzConfig::validate, katz.rs:57-84 (messages verbatim). */
int orc_validate_katz(double alpha, double beta, uint64_t max_iterations, double tolerance,
                      char *msg, size_t cap) {
    if (!isfinite(alpha) || alpha <= 0.0) {
        put_msg(msg, cap, "alpha must be finite and > 0");
        return ORC_INVALID;
    }
    if (alpha >= 1.0) {
        put_msg(msg, cap,
                "alpha must be < 1 (required for convergence; tighter bound "
                "alpha < 1/spectral_radius(A) may be needed for large graphs)");
        return ORC_INVALID;
    }
    if (!isfinite(beta)) {
        put_msg(msg, cap, "beta must be finite");
        return ORC_INVALID;
    }
    if (max_iterations == 0) {
        put_msg(msg, cap, "max_iterations must be > 0");
        return ORC_INVALID;
    }
    if (!isfinite(tolerance) || tolerance <= 0.0) {
        put_msg(msg, cap, "tolerance must be finite and > 0");
        return ORC_INVALID;
    }
    return ORC_OK;
}

/* katz_centrality_run, katz.rs:136-201 (push form, neighbour lists in
 * adapter order; ids >= n are skipped, katz.rs:170). */
void orc_katz_run(uint64_t n, const uint64_t *off, const uint32_t *nb, double alpha,
                  double beta, uint64_t max_iterations, double tolerance, double *scores,
                  uint64_t *iters_out, double *diff_out, int *conv_out) {
    if (n == 0) { /* :138-145 */
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; i++) cur[i] = beta; /* :149 */
    uint64_t iters = 0;
    double last_diff = INFINITY;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) { /* :160 */
        iters++;
        for (uint64_t i = 0; i < n; i++) nxt[i] = beta; /* :166 */
        for (uint64_t u = 0; u < n; u++) {              /* :167-174 */
            const double contrib = alpha * cur[u];
            for (uint64_t j = off[u]; j < off[u + 1]; j++)
                if (nb[j] < n) nxt[nb[j]] += contrib;
        }
        const double diff = l1_seq(n, cur, nxt); /* :181-185 */
        last_diff = diff;
        double *t = cur; cur = nxt; nxt = t;      /* :188 */
        if (diff < tolerance) {                   /* :189-192 */
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    *iters_out = iters;
    *diff_out = last_diff;
    *conv_out = converged;
}

/* eigenvector_centrality_run, eigenvector.rs:49-125: x <- (A + I) x with
 * (A x)[v] summed over v's own neighbour list, L2-normalised with the sign
 * that makes the sum non-negative. */
void orc_eigenvector_run(uint64_t n, const uint64_t *off, const uint32_t *nb,
                         uint64_t max_iterations, double tolerance, double *scores,
                         uint64_t *iters_out, double *diff_out, int *conv_out) {
    if (n == 0) { /* :55-62 */
        run_empty(diff_out, iters_out, conv_out);
        return;
    }
    double *cur = scores;
    double *nxt = (double *)malloc(n * sizeof(double));
    const double init = 1.0 / sqrt((double)n); /* :65 */
    for (uint64_t i = 0; i < n; i++) cur[i] = init;
    double diff = 0.0; /* :68 */
    uint64_t iters = 0;
    int converged = 0;
    for (uint64_t it = 0; it < max_iterations; it++) { /* :71 */
        iters = it + 1;
        for (uint64_t v = 0; v < n; v++) { /* :76-84 */
            double sum = cur[v];
            for (uint64_t j = off[v]; j < off[v + 1]; j++)
                if (nb[j] < n) sum += cur[nb[j]];
            nxt[v] = sum;
        }
        double sq = 0.0; /* :87 */
        for (uint64_t i = 0; i < n; i++) sq += nxt[i] * nxt[i];
        const double norm = sqrt(sq);
        if (norm > 0.0) { /* :88-98 */
            double tot = 0.0;
            for (uint64_t i = 0; i < n; i++) tot += nxt[i];
            const double sign = tot >= 0.0 ? 1.0 : -1.0;
            for (uint64_t i = 0; i < n; i++) nxt[i] = nxt[i] * sign / norm;
        }
        diff = l1_seq(n, cur, nxt); /* :101-105 */
        double *t = cur; cur = nxt; nxt = t; /* :107 */
        if (diff < tolerance) {           /* :109-116 */
            converged = 1;
            break;
        }
    }
    if (cur != scores) {
        memcpy(scores, cur, n * sizeof(double));
        free(cur);
    } else {
        free(nxt);
    }
    *iters_out = iters;
    *diff_out = diff;
    *conv_out = converged;
}

static void l2_normalize_seq(uint64_t n, double *v) { /* centrality.rs:168-175 */
    double sq = 0.0;
    for (uint64_t i = 0; i < n; i++) sq += v[i] * v[i];
    const double norm = sqrt(sq);
    if (norm > 0.0)
        for (uint64_t i = 0; i < n; i++) v[i] /= norm;
}

static double l2_diff_seq(uint64_t n, const double *a, const double *b) { /* :177-183 */
    double sq = 0.0;
    for (uint64_t i = 0; i < n; i++) sq += (a[i] - b[i]) * (a[i] - b[i]);
    return sqrt(sq);
}

/* hits, centrality.rs:114-166.  The reference returns (hub, authority);
 * iters_out, auth_diff and hub_diff report the loop count and the last L2
 * changes (0 iterations: diffs 0). */
void orc_hits(uint64_t n, const uint64_t *off, const uint32_t *nb, uint64_t max_iter,
              double tol, double *hub, double *auth, uint64_t *iters_out,
              double *auth_diff_out, double *hub_diff_out) {
    *iters_out = 0;
    *auth_diff_out = 0.0;
    *hub_diff_out = 0.0;
    if (n == 0) return; /* :116-118 */
    /* in-neighbour lists, u ascending (:121-128) */
    uint64_t *in_off = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    const uint64_t m = off[n];
    for (uint64_t j = 0; j < m; j++)
        if (nb[j] < n) in_off[nb[j] + 1]++;
    for (uint64_t v = 0; v < n; v++) in_off[v + 1] += in_off[v];
    uint32_t *in_src = (uint32_t *)malloc((in_off[n] ? in_off[n] : 1) * sizeof(uint32_t));
    uint64_t *cursor = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    memcpy(cursor, in_off, n * sizeof(uint64_t));
    for (uint64_t u = 0; u < n; u++)
        for (uint64_t j = off[u]; j < off[u + 1]; j++)
            if (nb[j] < n) in_src[cursor[nb[j]]++] = (uint32_t)u;
    free(cursor);
    const double init = 1.0 / sqrt((double)n); /* :130-132 */
    for (uint64_t i = 0; i < n; i++) hub[i] = auth[i] = init;
    double *new_auth = (double *)malloc(n * sizeof(double));
    double *new_hub = (double *)malloc(n * sizeof(double));
    for (uint64_t it = 0; it < max_iter; it++) { /* :134 */
        *iters_out = it + 1;
        for (uint64_t v = 0; v < n; v++) { /* :136-141 */
            double s = 0.0;
            for (uint64_t k = in_off[v]; k < in_off[v + 1]; k++) s += hub[in_src[k]];
            new_auth[v] = s;
        }
        l2_normalize_seq(n, new_auth); /* :142 */
        for (uint64_t u = 0; u < n; u++) { /* :145-152 */
            double s = 0.0;
            for (uint64_t j = off[u]; j < off[u + 1]; j++)
                if (nb[j] < n) s += new_auth[nb[j]];
            new_hub[u] = s;
        }
        l2_normalize_seq(n, new_hub); /* :153 */
        const double ad = l2_diff_seq(n, auth, new_auth); /* :156-157 */
        const double hd = l2_diff_seq(n, hub, new_hub);
        memcpy(auth, new_auth, n * sizeof(double)); /* :158-159 */
        memcpy(hub, new_hub, n * sizeof(double));
        *auth_diff_out = ad;
        *hub_diff_out = hd;
        if (ad < tol && hd < tol) break; /* :160-162 */
    }
    free(new_auth);
    free(new_hub);
    free(in_off);
    free(in_src);
}

/* top_k, topk.rs:17-42: a min-heap of (score, index) of at most k entries
 * (Reverse((NotNan, usize)) in a max-heap), scores that are not finite or
 * <= 0 skipped, the minimum replaced only by a strictly larger score; then
 * a sort by descending score.  The reference's sort is unstable; equal
 * scores are ordered here by ascending index.  Returns the entry count. */
static int tk_less(double sa, uint64_t ia, double sb, uint64_t ib) {
    return sa < sb || (sa == sb && ia < ib);
}

static void tk_sift_down(double *hs, uint64_t *hi, uint64_t len, uint64_t p) {
    for (;;) {
        uint64_t c = 2 * p + 1;
        if (c >= len) return;
        if (c + 1 < len && tk_less(hs[c + 1], hi[c + 1], hs[c], hi[c])) c++;
        if (!tk_less(hs[c], hi[c], hs[p], hi[p])) return;
        double ts = hs[p]; hs[p] = hs[c]; hs[c] = ts;
        uint64_t ti = hi[p]; hi[p] = hi[c]; hi[c] = ti;
        p = c;
    }
}

static void tk_sift_up(double *hs, uint64_t *hi, uint64_t c) {
    while (c > 0) {
        uint64_t p = (c - 1) / 2;
        if (!tk_less(hs[c], hi[c], hs[p], hi[p])) return;
        double ts = hs[p]; hs[p] = hs[c]; hs[c] = ts;
        uint64_t ti = hi[p]; hi[p] = hi[c]; hi[c] = ti;
        c = p;
    }
}

static double *tk_sort_scores;
static int tk_desc(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    const double sx = tk_sort_scores[x], sy = tk_sort_scores[y];
    if (sx != sy) return sx > sy ? -1 : 1;
    return x < y ? -1 : (x > y);
}

uint64_t orc_top_k(uint64_t n, const double *scores, uint64_t k, uint64_t *ids, double *vals) {
    if (k == 0 || n == 0) return 0; /* :18-20 */
    double *hs = (double *)malloc(k * sizeof(double));
    uint64_t *hi = (uint64_t *)malloc(k * sizeof(uint64_t));
    uint64_t len = 0;
    for (uint64_t i = 0; i < n; i++) { /* :22-35 */
        const double s = scores[i];
        if (!isfinite(s) || s <= 0.0) continue;
        if (len < k) {
            hs[len] = s;
            hi[len] = i;
            tk_sift_up(hs, hi, len);
            len++;
        } else if (s > hs[0]) {
            hs[0] = s;
            hi[0] = i;
            tk_sift_down(hs, hi, len, 0);
        }
    }
    tk_sort_scores = (double *)scores; /* :36-40 */
    qsort(hi, len, sizeof(uint64_t), tk_desc);
    for (uint64_t r = 0; r < len; r++) {
        ids[r] = hi[r];
        vals[r] = scores[hi[r]];
    }
    free(hs);
    free(hi);
    return len;
}
