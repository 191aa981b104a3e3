/*
 * egx_oracle.c -- CPU oracle for the PageRank path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the repository's smoke check and bench.py's cpu_baseline leg
 * may load this library, and only to check results or as the timed CPU
 * baseline.
 * Nothing in the eggpu_amd package links it or calls it.
 *
 * Every function restates a reference algorithm in plain C, in the
 * reference's own operation order, and is built with -ffp-contract=off so
 * that no multiply-add is fused behind our back:
 *
 *   orc_pagerank_push   Easy-Graph/easygraph/functions/centrality/pagerank.py:75-139
 *                       (_pagerank_power_iteration, the FP64 oracle)
 *   orc_build_inbound   Easy-Graph/gpu_easygraph/functions/centrality/pagerank.cu:286-347
 *                       (build_inbound_from_csr, the pull CSR built on the host)
 *   orc_pagerank_omp    Easy-Graph/cpp_easygraph/functions/pagerank/pagerank.cpp:21-122
 *                       (_pagerank, OpenMP push with atomic scatter; CPU baseline)
 *   orc_ppr_push        personalised PageRank with networkx 3.4.2 semantics
 *                       (teleport and dangling mass both follow p); the
 *                       reference has no personalisation (pagerank.py:14)
 *   orc_rmat_coo        this project's seeded R-MAT definition; the reference
 *                       ships no R-MAT generator.  The HIP generator in
 *                       eggpu_amd/csrc must reproduce it bit for bit.
 *
 * Pinning: see tests/test_oracle.py (networkx 3.4.2, the EasyGraph README
 * known-answer vector, closed forms, brute force on tiny graphs).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* pagerank.py:75-139                                                        */
/* rows: (src[e], dst[e], w[e]) in the order the Python oracle builds them   */
/* (G.edges order; for undirected graphs the caller passes the forward rows  */
/* without self-loops followed by the reversed rows, pagerank.py:87-91).     */
/* w == NULL means unweighted (every row weight is 1.0, pagerank.py:100-102). */
/* Returns 0, or -1 on allocation failure.                                   */
/* ------------------------------------------------------------------------ */
int orc_pagerank_push(int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                      const double *w, double alpha, int32_t max_iter, double tol,
                      double *out, int32_t *iters_out, double *err_out)
{
    if (iters_out) *iters_out = 0;
    if (err_out) *err_out = 0.0;
    if (n <= 0) return 0;
    if (m == 0) { /* pagerank.py:93-95 */
        for (int64_t i = 0; i < n; ++i) out[i] = 1.0 / (double)n;
        return 0;
    }
    /* out_adj[i] as a CSR that keeps append order (pagerank.py:97-107) */
    int64_t *off = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *adj_dst = (int32_t *)malloc((size_t)m * sizeof(int32_t));
    double *adj_w = (double *)malloc((size_t)m * sizeof(double));
    double *out_sums = (double *)calloc((size_t)n, sizeof(double));
    double *pr = (double *)malloc((size_t)n * sizeof(double));
    double *new_pr = (double *)malloc((size_t)n * sizeof(double));
    int64_t *fill = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!off || !adj_dst || !adj_w || !out_sums || !pr || !new_pr || !fill) {
        free(off); free(adj_dst); free(adj_w); free(out_sums); free(pr); free(new_pr); free(fill);
        return -1;
    }
    for (int64_t e = 0; e < m; ++e) off[src[e] + 1] += 1;
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    for (int64_t i = 0; i < n; ++i) fill[i] = off[i];
    for (int64_t e = 0; e < m; ++e) {
        int64_t s = src[e];
        double wf = w ? w[e] : 1.0;
        int64_t pos = fill[s]++;
        adj_dst[pos] = dst[e];
        adj_w[pos] = wf;
        out_sums[s] += wf;
    }
    for (int64_t i = 0; i < n; ++i) pr[i] = 1.0 / (double)n;

    int32_t it_done = 0;
    double err = 0.0;
    for (int32_t it = 0; it < max_iter; ++it) {
        for (int64_t i = 0; i < n; ++i) new_pr[i] = 0.0;
        double dangling_sum = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            if (out_sums[i] <= 1.0e-15) {
                dangling_sum += pr[i];
                continue;
            }
            double base = alpha * pr[i] / out_sums[i];
            for (int64_t p = off[i]; p < off[i + 1]; ++p) new_pr[adj_dst[p]] += base * adj_w[p];
        }
        double jump = (1.0 - alpha) / (double)n + alpha * dangling_sum / (double)n;
        err = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double val = new_pr[i] + jump;
            err += fabs(val - pr[i]);
            pr[i] = val;
        }
        it_done = it + 1;
        if (err < tol * (double)n) break;
    }
    /* pagerank.py:134-139 (Python's sum() on 3.10 adds left to right) */
    double norm = 0.0;
    for (int64_t i = 0; i < n; ++i) norm += pr[i];
    if (norm <= 0.0) {
        for (int64_t i = 0; i < n; ++i) out[i] = 1.0 / (double)n;
    } else {
        double inv = 1.0 / norm;
        for (int64_t i = 0; i < n; ++i) out[i] = pr[i] * inv;
    }
    if (iters_out) *iters_out = it_done;
    if (err_out) *err_out = err;
    free(off); free(adj_dst); free(adj_w); free(out_sums); free(pr); free(new_pr); free(fill);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Personalised PageRank, networkx 3.4.2 semantics (pagerank_scipy):         */
/*   x0 = 1/n;  x' = alpha*(x A + dang(x) p) + (1-alpha) p;                  */
/*   err = sum |x' - x|; stop when err < n*tol; then the final sum-normalise */
/*   of pagerank.py:134-139.  The push order is the one of orc_pagerank_push.*/
/* p must be a probability vector over the n nodes.                          */
/* ------------------------------------------------------------------------ */
int orc_ppr_push(int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                 const double *w, double alpha, int32_t max_iter, double tol,
                 const double *p, double *out, int32_t *iters_out, double *err_out)
{
    if (iters_out) *iters_out = 0;
    if (err_out) *err_out = 0.0;
    if (n <= 0) return 0;
    int64_t *off = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *adj_dst = (int32_t *)malloc(((size_t)m + 1) * sizeof(int32_t));
    double *adj_w = (double *)malloc(((size_t)m + 1) * sizeof(double));
    double *out_sums = (double *)calloc((size_t)n, sizeof(double));
    double *pr = (double *)malloc((size_t)n * sizeof(double));
    double *new_pr = (double *)malloc((size_t)n * sizeof(double));
    int64_t *fill = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!off || !adj_dst || !adj_w || !out_sums || !pr || !new_pr || !fill) {
        free(off); free(adj_dst); free(adj_w); free(out_sums); free(pr); free(new_pr); free(fill);
        return -1;
    }
    for (int64_t e = 0; e < m; ++e) off[src[e] + 1] += 1;
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    for (int64_t i = 0; i < n; ++i) fill[i] = off[i];
    for (int64_t e = 0; e < m; ++e) {
        int64_t s = src[e];
        double wf = w ? w[e] : 1.0;
        int64_t pos = fill[s]++;
        adj_dst[pos] = dst[e];
        adj_w[pos] = wf;
        out_sums[s] += wf;
    }
    for (int64_t i = 0; i < n; ++i) pr[i] = 1.0 / (double)n;
    int32_t it_done = 0;
    double err = 0.0;
    for (int32_t it = 0; it < max_iter; ++it) {
        for (int64_t i = 0; i < n; ++i) new_pr[i] = 0.0;
        double dangling_sum = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            if (out_sums[i] <= 1.0e-15) {
                dangling_sum += pr[i];
                continue;
            }
            double base = alpha * pr[i] / out_sums[i];
            for (int64_t q = off[i]; q < off[i + 1]; ++q) new_pr[adj_dst[q]] += base * adj_w[q];
        }
        double c = (1.0 - alpha) + alpha * dangling_sum;
        err = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double val = new_pr[i] + c * p[i];
            err += fabs(val - pr[i]);
            pr[i] = val;
        }
        it_done = it + 1;
        if (err < tol * (double)n) break;
    }
    double norm = 0.0;
    for (int64_t i = 0; i < n; ++i) norm += pr[i];
    if (norm <= 0.0) {
        for (int64_t i = 0; i < n; ++i) out[i] = 1.0 / (double)n;
    } else {
        double inv = 1.0 / norm;
        for (int64_t i = 0; i < n; ++i) out[i] = pr[i] * inv;
    }
    if (iters_out) *iters_out = it_done;
    if (err_out) *err_out = err;
    free(off); free(adj_dst); free(adj_w); free(out_sums); free(pr); free(new_pr); free(fill);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pagerank.cu:286-347 build_inbound_from_csr, including its clamping of     */
/* row bounds (:302-306), skipping of invalid endpoints (:310) and weight    */
/* sanitising (:311-312).  in_col/in_val must hold at least m entries.       */
/* Returns the number of stored inbound entries (in_row[n]).                 */
/* ------------------------------------------------------------------------ */
int64_t orc_build_inbound(int32_t n, int32_t m, const int32_t *V, const int32_t *E,
                          const double *W, double alpha, int32_t *in_row, int32_t *in_col,
                          float *in_val, int32_t *dangling01, double *out_weight_sum)
{
    int32_t *indeg = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *cur = (int32_t *)malloc(((size_t)n + 1) * sizeof(int32_t));
    if (!indeg || !cur) { free(indeg); free(cur); return -1; }
    for (int32_t u = 0; u < n; ++u) out_weight_sum[u] = 0.0;
    for (int32_t u = 0; u < n; ++u) {
        int32_t begin = V[u], end = V[u + 1];
        if (begin < 0) begin = 0;
        if (end > m) end = m;
        if (begin > end) begin = end;
        for (int32_t p = begin; p < end; ++p) {
            int32_t v = E[p];
            if (!(0 <= v && v < n)) continue;
            double w = W ? W[p] : 1.0;
            if (!isfinite(w) || w < 0.0) w = 0.0;
            out_weight_sum[u] += w;
            indeg[v] += 1;
        }
    }
    in_row[0] = 0;
    for (int32_t i = 0; i < n; ++i) in_row[i + 1] = in_row[i] + indeg[i];
    for (int32_t u = 0; u < n; ++u) dangling01[u] = (out_weight_sum[u] <= 0.0) ? 1 : 0;
    for (int32_t i = 0; i <= n; ++i) cur[i] = in_row[i];
    for (int32_t u = 0; u < n; ++u) {
        int32_t begin = V[u], end = V[u + 1];
        if (begin < 0) begin = 0;
        if (end > m) end = m;
        if (begin > end) begin = end;
        double denom = out_weight_sum[u];
        for (int32_t p = begin; p < end; ++p) {
            int32_t v = E[p];
            if (!(0 <= v && v < n)) continue;
            double w = W ? W[p] : 1.0;
            if (!isfinite(w) || w < 0.0) w = 0.0;
            int32_t pos = cur[v]++;
            in_col[pos] = u;
            in_val[pos] = (denom > 0.0) ? (float)(alpha * (w / denom)) : 0.0f;
        }
    }
    int64_t total = in_row[n];
    free(indeg); free(cur);
    return total;
}

/* ------------------------------------------------------------------------ */
/* pagerank.cpp:21-122 _pagerank: OpenMP push over CSR rows (u = 0..n-1      */
/* stands for the 1-based ids 1..N of Graph_L).  The reference's weighted    */
/* logic is used only when some |w-1| > 1e-9 (:54-67); dangling is outdeg==0 */
/* unweighted, outW < 1e-15 weighted (:78).  No final normalisation (:119).  */
/* ------------------------------------------------------------------------ */
int orc_pagerank_omp(int32_t n, const int32_t *V, const int32_t *E, const double *W,
                     double alpha, int32_t max_iter, double threshold, int32_t nthreads,
                     double *out, int32_t *iters_out)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double *outW = (double *)calloc((size_t)n, sizeof(double));
    double *oldPR = (double *)malloc((size_t)n * sizeof(double));
    double *newPR = (double *)calloc((size_t)n, sizeof(double));
    if (!outW || !oldPR || !newPR) { free(outW); free(oldPR); free(newPR); return -1; }
    int weighted = 0;
    if (W) {
#pragma omp parallel for reduction(| : weighted) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int32_t p = V[i]; p < V[i + 1]; ++p) {
                s += W[p];
                if (fabs(W[p] - 1.0) > 1e-9) weighted = 1;
            }
            outW[i] = s;
        }
    }
    for (int32_t i = 0; i < n; ++i) oldPR[i] = 1.0 / (double)n;
    int32_t cnt = 0, performed = 0;
    while (cnt < max_iter) {
        double dangling_sum = 0.0;
#pragma omp parallel for reduction(+ : dangling_sum) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            int dang = weighted ? (outW[i] < 1e-15) : (V[i + 1] == V[i]);
            if (dang) dangling_sum += oldPR[i];
        }
        if (!weighted) {
#pragma omp parallel for schedule(dynamic, 128)
            for (int32_t i = 0; i < n; ++i) {
                int32_t d = V[i + 1] - V[i];
                if (d == 0) continue;
                double out_val = (oldPR[i] / d) * alpha;
                for (int32_t p = V[i]; p < V[i + 1]; ++p) {
#pragma omp atomic
                    newPR[E[p]] += out_val;
                }
            }
        } else {
#pragma omp parallel for schedule(dynamic, 128)
            for (int32_t i = 0; i < n; ++i) {
                if (outW[i] < 1e-15) continue;
                double out_val = (oldPR[i] / outW[i]) * alpha;
                for (int32_t p = V[i]; p < V[i + 1]; ++p) {
#pragma omp atomic
                    newPR[E[p]] += out_val * W[p];
                }
            }
        }
        double diff_sum = 0.0;
        double jump_val = (1.0 - alpha) / n + (dangling_sum / n) * alpha;
#pragma omp parallel for reduction(+ : diff_sum) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            double final_pr = newPR[i] + jump_val;
            diff_sum += fabs(final_pr - oldPR[i]);
            oldPR[i] = final_pr;
            newPR[i] = 0.0;
        }
        performed += 1;
        if (diff_sum < threshold * n) break;
        cnt++;
    }
    for (int32_t i = 0; i < n; ++i) out[i] = oldPR[i];
    if (iters_out) *iters_out = performed;
    free(outW); free(oldPR); free(newPR);
    return 0;
}

int orc_omp_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Seeded R-MAT (this project's definition).                                 */
/* Edge e draws `scale` quadrant choices, most significant bit first.  Level */
/* l uses 32 bits of the 64-bit word mix64(seed + G*(16*e + l/2 + 1)), the   */
/* low half for even l and the high half for odd l.  Quadrant 0 (src bit 0,  */
/* dst bit 0) when u < tA, 1 (0,1) when u < tAB, 2 (1,0) when u < tABC,      */
/* else 3 (1,1).  The thresholds are floor(cumulative probability * 2^32).   */
/* ------------------------------------------------------------------------ */
static inline uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_rmat_coo(int32_t scale, int64_t e_begin, int64_t e_end, uint64_t seed, uint32_t tA,
                  uint32_t tAB, uint32_t tABC, int32_t *src, int32_t *dst)
{
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
#pragma omp parallel for schedule(static)
    for (int64_t e = e_begin; e < e_end; ++e) {
        uint32_t s = 0, d = 0;
        uint64_t word = 0;
        for (int32_t l = 0; l < scale; ++l) {
            if ((l & 1) == 0) word = orc_mix64(seed + G * ((uint64_t)e * 16u + (uint64_t)(l >> 1) + 1u));
            uint32_t u = (l & 1) ? (uint32_t)(word >> 32) : (uint32_t)word;
            uint32_t q = (u < tA) ? 0u : (u < tAB) ? 1u : (u < tABC) ? 2u : 3u;
            s = (s << 1) | (q >> 1);
            d = (d << 1) | (q & 1u);
        }
        src[e - e_begin] = (int32_t)s;
        dst[e - e_begin] = (int32_t)d;
    }
}

/* Erdos-Renyi G(n, M) draw with the same word stream: src = hi32 % n,      */
/* dst = lo32 % n of mix64(seed + G*(16*e + 1)).                            */
void orc_er_coo(int32_t n, int64_t e_begin, int64_t e_end, uint64_t seed, int32_t *src, int32_t *dst)
{
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
#pragma omp parallel for schedule(static)
    for (int64_t e = e_begin; e < e_end; ++e) {
        uint64_t word = orc_mix64(seed + G * ((uint64_t)e * 16u + 1u));
        src[e - e_begin] = (int32_t)((uint32_t)(word >> 32) % (uint32_t)n);
        dst[e - e_begin] = (int32_t)((uint32_t)word % (uint32_t)n);
    }
}

/* Edge weight in [lo, hi): u = top 24 bits of mix64(wseed ^ (src<<32 | dst)),
 * w = (float)(lo + (hi - lo) * (u * 2^-24)), double arithmetic, no fused ops. */
void orc_edge_weights(int64_t m, const int32_t *src, const int32_t *dst, uint64_t wseed,
                      double lo, double hi, double *w)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        uint64_t key = ((uint64_t)(uint32_t)src[e] << 32) | (uint64_t)(uint32_t)dst[e];
        uint64_t u = orc_mix64(wseed ^ key) >> 40;
        double t = (double)u * (1.0 / 16777216.0);
        double span = hi - lo;
        double prod = span * t;
        w[e] = (double)(float)(lo + prod);
    }
}
