/*
 * egx.h -- C ABI of the MI355X-native PageRank library (libegx.so).
 *
 * Plain pointers and sizes only.  Every entry point returns a status code
 * whose numeric values are those of EG_GPU_STATUS_CODE
 * (Easy-Graph/gpu_easygraph/common/err.h:7-15); no exception crosses the ABI.
 * Host pointers are borrowed for the duration of the call; device-resident
 * graphs (egx_graph) are owned by the library until egx_graph_free().
 */
#ifndef EGX_H
#define EGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EG_GPU_STATUS_CODE, Easy-Graph/gpu_easygraph/common/err.h:7-15 */
#define EGX_SUCC 0
#define EGX_FAILED_TO_ALLOCATE_HOST_MEM 1
#define EGX_FAILED_TO_ALLOCATE_DEVICE_MEM 2
#define EGX_DEVICE_ERR 3
#define EGX_UNKNOW_ERROR 4
#define EGX_UNSUPPORTED_GRAPH 5
#define EGX_NEGATIVE_CYCLE 6

/* Iterate storage of a solve. */
#define EGX_PRECISION_FP32 0 /* FP32 ranks and weights, FP64 dangling/L1/teleport */
#define EGX_PRECISION_FP64 1 /* FP64 ranks (weights stay FP32, as in the build) */

typedef struct egx_graph egx_graph; /* opaque, device resident */

/* Replaces gpu_easygraph::err_code_detail (Easy-Graph/gpu_easygraph/common/err.cpp:9-31):
 * same strings, "EasyGraph GPU: ..." prefix. */
const char *egx_err_code_detail(int status);

/* Detail text of the last failure on the calling thread ("" if none). */
const char *egx_last_error(void);

/* Library version string. */
const char *egx_version(void);

/* Number of visible devices (0 when no GPU); never fails. */
int egx_device_count(void);

/*
 * Replaces cuda_pagerank (Easy-Graph/gpu_easygraph/functions/centrality/pagerank.cuh:7-18)
 * and the C++ wrapper gpu_easygraph::pagerank (Easy-Graph/gpu_easygraph/gpu_easygraph.h:95-104).
 *
 * V[len_V+1], E[len_E]: out-CSR of the graph (Graph::gen_CSR order).  W[len_E]
 * edge weights or NULL for unit weights; weights that are all exactly 1 (the
 * reference's unweighted_W) are taken as NULL, which gives the same pull CSC
 * and runs the unweighted kernels.  len_V is the node count (the
 * reference wrapper passes V.size()-1).  graph_id plays the part of
 * active_device_graph_id() (device_graph_cache.h:34-48): the prepared device
 * graph is cached per host thread and reused while graph_id, the V/E/W
 * pointers, n, m and the first/last entries stay the same
 * (pagerank.cu:140-167); alpha no longer forces a rebuild, only a cheap
 * device recomputation of the normalised weights.
 * PR[len_V] (caller owned) receives sum-normalised FP64 ranks.  iterations and
 * kernel_seconds may be NULL.  Offsets must be monotonic with V[0]=0 and
 * V[len_V]=len_E (EGX_UNSUPPORTED_GRAPH otherwise); endpoints outside
 * [0, len_V) are skipped and weights that are negative or not finite count as
 * 0, as in pagerank.cu:282-312.  A thread that switches device while its
 * cache lives gets EGX_DEVICE_ERR (pagerank.cu:373-377).
 */
int egx_pagerank(const int32_t *V, const int32_t *E, const double *W, int32_t len_V,
                 int32_t len_E, double alpha, int32_t max_iter_num, double threshold,
                 uint64_t graph_id, double *PR, int32_t *iterations, double *kernel_seconds);

/* Prepared-graph cache of egx_pagerank: a per-thread LRU registry after
 * device_graph_cache.cu:53-168 (EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_ENTRIES,
 * default 4, at most 16; EASYGRAPH_GPU_DEVICE_CSR_CACHE_MAX_BYTES, default
 * max((free + owned)/4, 256 MB); EASYGRAPH_GPU_DISABLE_DEVICE_CSR_CACHE).
 * Control after cpp_gpu_device_csr_cache_reset/_stats
 * (Easy-Graph/cpp_easygraph/cpp_easygraph.cpp:97-113, device_graph_cache.cu:519-527). */
typedef struct {
    uint64_t hits;
    uint64_t misses;
    uint64_t builds;
    uint64_t value_refreshes; /* alpha changes served without a rebuild */
    uint64_t resident_bytes;  /* prepared graphs of the calling thread */
    uint64_t evictions;       /* least-recently-used entries dropped for the entry or byte limit */
    uint64_t entries;         /* prepared graphs held by the calling thread */
} egx_cache_stats;
void egx_pagerank_cache_reset(void);
int egx_pagerank_cache_stats(egx_cache_stats *out);

/* ------------------------------------------------------------------------ */
/* Device-resident graphs.  No counterpart in the reference, which keeps one  */
/* thread-local workspace (pagerank.cu:44-132); these give the serving pool,  */
/* the batch engines and benchmarks an explicit handle.                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n;               /* nodes */
    int64_t m;               /* stored inbound arcs (valid endpoints) */
    int64_t n_dangling;      /* nodes whose sanitised out-weight is 0 */
    int32_t weighted;        /* 1 when explicit weights were given */
    int32_t max_in_degree;
    int32_t max_out_degree;
    int32_t thread_row_max;  /* bin settings in use (egx_bin_opts) */
    int32_t wave_row_max;
    int32_t hub_segment;
    int32_t tile_items;
    int32_t reordered;       /* 1: iteration layout relabelled by out-degree */
    int32_t hub_rows_first;  /* hub rows, stored first in the iteration layout */
    int32_t tile_sort;       /* 1: tile arcs gathered in source order */
    int32_t hot_sources;     /* gathers past this relabelled id are non-temporal (0: none) */
    int32_t blocked;         /* 1: iterations run propagation-blocked (egx_bin_opts.blocked) */
    int64_t rows_thread;     /* rows reduced one thread per row */
    int64_t rows_wave;       /* rows reduced one wavefront per row */
    int64_t rows_hub;        /* rows of the hub bin (segmented, one workgroup per segment) */
    int64_t hub_segments;
    int64_t tiles;           /* LDS-staged tiles of non-hub rows */
    int64_t device_bytes;
    double build_seconds;    /* whole device build, host clock */
    double layout_seconds;   /* relabelling part of it (hipEvents) */
} egx_graph_info;

/* Degree bins of the SpMV.  Rows with in-degree above wave_row_max form the
 * hub bin (cut into hub_segment-arc segments, one workgroup each); the other
 * rows are split into tiles of tile_items rows+arcs whose gathered terms are
 * staged in LDS and reduced one thread per row when the tile's mean row
 * length is at most thread_row_max, else one wavefront per row.
 * tile_items + wave_row_max <= 7936. */
typedef struct {
    int32_t thread_row_max;
    int32_t wave_row_max;
    int32_t hub_segment;
    int32_t tile_items;
    int32_t reorder;        /* 1: relabel vertices by out-degree, 0: keep ids, <0: keep setting */
    int32_t tile_sort;      /* 1: gather tile arcs in source order, 0: row order, <0: keep setting */
    int32_t hot_sources;    /* > 0: gather sources past this relabelled id non-temporally;
                               0: keep setting; < 0: all gathers cached */
    int32_t blocked;        /* 1: propagation-blocked iterations (unweighted graphs, FP32
                               iterates): sources are scattered by destination bin
                               through LDS, then summed per bin in LDS, with no
                               random gathers; 0: pull SpMV; < 0: keep setting */
    int32_t blocked_shared; /* 1: size the blocked layout's scatter and sum items for
                               graphs solved together (egx_multi_pagerank): the
                               largest items, since the set fills the device; 0: size
                               them so that the graph alone fills it (the default);
                               < 0: keep setting */
} egx_bin_opts;

typedef struct {
    double alpha;
    int32_t max_iter;       /* clamped to >= 1, as pagerank.py:45-48 */
    double tol;             /* stop when L1(x' - x) < tol * n, pagerank.py:131 */
    int32_t precision;      /* EGX_PRECISION_* */
    int32_t check_every;    /* 0: adaptive host checks of the device stop flag */
} egx_solve_opts;

/* Build from a host out-CSR (same contract as egx_pagerank, all-ones weights
 * included). keep_csr keeps the out-CSR on the device so
 * egx_graph_download_csr can return it. */
int egx_graph_from_csr(const int32_t *V, const int32_t *E, const double *W, int32_t n,
                       int32_t m, double alpha, int32_t keep_csr, egx_graph **out);

/* Seeded directed R-MAT generated and built on the device: 2^scale nodes,
 * edge_factor * 2^scale draws, self-loops and duplicates removed.  When
 * weighted, w = fl32(wlo + (whi - wlo) * u24 * 2^-24) per arc. */
int egx_graph_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                   uint64_t seed, int32_t weighted, double wlo, double whi, uint64_t wseed,
                   double alpha, int32_t keep_csr, egx_graph **out);

/* R-MAT on n nodes (any n >= 2): draws descend ceil(log2 n) levels and a
 * draw with an endpoint >= n is dropped like a self-loop.  For n = 2^scale and
 * nedges = edge_factor * n this is egx_graph_rmat. */
int egx_graph_rmat_n(int32_t n, int64_t nedges, double a, double b, double c, uint64_t seed,
                     int32_t weighted, double wlo, double whi, uint64_t wseed, double alpha,
                     int32_t keep_csr, egx_graph **out);

/* eggpu-csr-v1 artifact files (int32 offsets[n+1], int32 indices[m], optional
 * float64 weights[m]; eggpu_bulk_graph.py:170-334) streamed straight into
 * HBM: 32 MB pread chunks through two pinned buffers, each copied while the
 * next is read (replaces cpp_graph_from_csr_files, graph_convert.cpp:447-526,
 * plus the first call's upload, pagerank.cu:497-516).  Validation runs on
 * the device: offsets must start at 0, end at m and not decrease (always);
 * with validate != 0 endpoints must lie in [0, n) and weights be finite
 * (otherwise invalid endpoints are skipped, as the reference build does).
 * weights_path NULL or "" means unit weights.  report may be NULL. */
typedef struct {
    int32_t in_range;       /* every endpoint in [0, n) */
    int32_t finite;         /* every weight finite */
    int32_t simple_sorted;  /* no self-loop, columns strictly increasing per row */
    int32_t pad;
    int64_t bytes;          /* file bytes moved to HBM */
    double read_seconds;    /* host pread time */
    double copy_seconds;    /* open to last byte in HBM (reads overlap copies) */
    double check_seconds;
    double build_seconds;   /* pull CSC + iteration layout */
} egx_ingest_report;

int egx_graph_from_csr_files(const char *offsets_path, const char *indices_path, const char *weights_path,
                             int32_t n, int32_t m, double alpha, int32_t validate, int32_t keep_csr,
                             egx_ingest_report *report, egx_graph **out);

/* Seeded Erdos-Renyi G(n, M) draw (M pairs, self-loops and duplicates removed). */
int egx_graph_er(int32_t n, int64_t nedges, uint64_t seed, int32_t weighted, double wlo,
                 double whi, uint64_t wseed, double alpha, int32_t keep_csr, egx_graph **out);

int egx_graph_info_get(const egx_graph *g, egx_graph_info *out);

/* Re-bin rows for the SpMV kernels (autotuner hook). Fields <= 0 keep the current value. */
int egx_graph_set_bins(egx_graph *g, const egx_bin_opts *opts);

/* Pull CSR (the reference's in_row/in_col/in_val/dangling01, pagerank.cu:286-347)
 * for the alpha of the last build/solve. Arrays sized n+1, m, m, n; any may be NULL. */
int egx_graph_download_csc(const egx_graph *g, int32_t *in_row, int32_t *in_col, float *in_val,
                           int32_t *dangling01);

/* Out-CSR kept by keep_csr (V[n+1], E[m_out], W[m_out] or NULL). */
int egx_graph_download_csr(const egx_graph *g, int32_t *V, int32_t *E, double *W,
                           int64_t *m_out);

/* Global PageRank on a device-resident graph. PR[n] may be NULL (keep the
 * result on the device).  kernel_seconds covers the iteration launches only
 * (hipEvents, as pagerank.cu:578,666-676). */
int egx_graph_pagerank(egx_graph *g, const egx_solve_opts *opts, double *PR, int32_t *iterations,
                       double *kernel_seconds);

/* egx_graph_pagerank with no allocation in the call.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-solve workspace that egx_graph_pagerank allocates on a graph's first
 * solve comes from a static 64 MiB device array instead (one per GPU, held
 * until the process ends): FP32 graphs up to about 4M nodes, FP64 up to
 * about 2M; larger graphs return EGX_UNSUPPORTED_GRAPH with the exact need in
 * egx_last_error.  opts as egx_graph_pagerank (check_every as there; no
 * per-launch timing, no graph replay).  PR_device: caller-owned device memory
 * on the graph's device, n doubles, written with the normalised ranks in
 * caller node order (the values egx_graph_pagerank returns, bit for bit).
 * iterations may be NULL.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU race, as do overlapping calls on one graph
 * through any entry point. */
int egx_graph_pagerank_noalloc(egx_graph *g, const egx_solve_opts *opts, double *PR_device,
                               int32_t *iterations);

/* Solve statistics; spmv/reduce sums are filled with EGX_SOLVE_TIME_LAUNCHES
 * (a hipEvent pair around every launch on the graph's stream). */
typedef struct {
    double kernel_seconds;  /* first to last iteration launch */
    double spmv_seconds;    /* sum of the SpMV launch durations (iterations performed) */
    double reduce_seconds;  /* sum of the reduce launch durations */
    int32_t iterations;
    int32_t spmv_launches;
    int64_t spmv_grid;      /* workgroups per SpMV launch */
} egx_solve_stats;

/* flags for egx_graph_pagerank_ex / egx_multi_pagerank */
#define EGX_SOLVE_TIME_LAUNCHES 1 /* time every launch (no hipGraph replay) */
#define EGX_SOLVE_HIP_GRAPH 2     /* capture check_every iterations once, replay the hipGraph */
#define EGX_SOLVE_XCD_PLACEMENT 4 /* multi-graph: put each small graph's work items on one XCD */

int egx_graph_pagerank_ex(egx_graph *g, const egx_solve_opts *opts, double *PR, egx_solve_stats *stats,
                          int32_t flags);

/* One personalised PageRank vector on the plain SpMV (no counterpart in the
 * reference, which has no personalisation, pagerank.py:14): the teleport and
 * the dangling mass go uniformly to the distinct seeds (networkx 3.4.2
 * semantics, as egx_graph_ppr), x0 = 1/n, stop when sum|x' - x| < tol*n,
 * result sum-normalised.  The single-vector form of egx_graph_ppr, with the
 * gather cost of one FP32 vector per arc. */
int egx_graph_ppr_single(egx_graph *g, const int32_t *seeds, int32_t nseeds, const egx_solve_opts *opts,
                         double *PR, egx_solve_stats *stats, int32_t flags);

/* Global PageRank on many device-resident graphs of one device at once (no
 * counterpart in the reference, which solves one graph per call,
 * pagerank.cu:591-664).  All graphs must be weighted or all unweighted; each
 * stops on its own tol test and gets the result it would get alone
 * (bit-identical: the same block code runs over a table of per-graph
 * arguments, one SpMV and one reduce launch per iteration for the whole set).
 * PR[i] (n_i doubles) and iterations[i] may be NULL; stats sums the set. */
int egx_multi_pagerank(egx_graph *const *graphs, int32_t count, const egx_solve_opts *opts,
                       double *const *PR, int32_t *iterations, egx_solve_stats *stats, int32_t flags);

void egx_graph_free(egx_graph *g);

/* ------------------------------------------------------------------------ */
/* Batched and personalised PageRank (no counterpart in the reference, whose  */
/* pagerank.py:14 has no personalisation).  Up to 64 columns share one graph: */
/* each is a global PageRank (no seeds) or a personalised one whose teleport  */
/* and dangling mass are spread uniformly over its seed nodes (networkx 3.4.2 */
/* pagerank(personalization=...) semantics), with its own alpha, tol,         */
/* max_iter, stop flag and iteration count.  Results of a column do not       */
/* depend on the other columns of the batch.                                  */
/* ------------------------------------------------------------------------ */
typedef struct egx_batch egx_batch; /* opaque, owns its stream and HBM state */

/* Open a batch session on g (g must outlive it and is not re-binned meanwhile). */
int egx_batch_open(egx_graph *g, egx_batch **out);

/* (Re)start column slot in [0, 64): x0 = 1/n, seeds are node ids (nseeds = 0:
 * global PageRank).  Safe between steps; other columns keep running. */
int egx_batch_assign(egx_batch *b, int32_t slot, const int32_t *seeds, int32_t nseeds, double alpha,
                     int32_t max_iter, double tol);

/* Free column slot (it stops taking part in later iterations). */
int egx_batch_release(egx_batch *b, int32_t slot);

/* Enqueue `iterations` iterations of every active column (asynchronous). */
int egx_batch_step(egx_batch *b, int32_t iterations);

/* Wait for the batch stream and read the per-column state; each array holds
 * 64 entries and may be NULL. */
int egx_batch_poll(egx_batch *b, int32_t *active, int32_t *done, int32_t *iterations, double *err);

/* Step until no active column is left or max_launches iterations ran,
 * polling every check_every iterations (0: 4).  kernel_seconds: hipEvents. */
int egx_batch_run(egx_batch *b, int32_t max_launches, int32_t check_every, double *kernel_seconds);

/* Sum-normalised FP64 ranks of column slot, in node id order (PR[n]). */
int egx_batch_result(egx_batch *b, int32_t slot, double *PR);

/* 1 when everything queued on the session has finished, 0 while work is in
 * flight, a negative status on error.  Never blocks. */
int egx_batch_ready(egx_batch *b);

void egx_batch_close(egx_batch *b);

/* One-shot: k columns, column j seeded by seeds[seed_offsets[j] .. seed_offsets[j+1])
 * (an empty range = global PageRank), alpha_j = alphas ? alphas[j] : alpha.
 * PR (k x n, row j = column j) and iterations (k) may be NULL.  The batch
 * session it runs in stays on the handle for the next call (its X buffers,
 * 8*n*width bytes, and captured step graphs); egx_graph_set_bins and
 * egx_graph_free close it.  Calls on one handle run one at a time. */
int egx_graph_ppr(egx_graph *g, int32_t k, const int32_t *seed_offsets, const int32_t *seeds,
                  const double *alphas, double alpha, int32_t max_iter, double tol, double *PR,
                  int32_t *iterations, double *kernel_seconds);

/* ------------------------------------------------------------------------ */
/* Katz centrality on the same SpMV engine.  No counterpart on the reference's */
/* GPU path; it restates the CPU function katz_centrality.py:7-110 (the     */
/* Python route of its @hybrid dispatch): A[u][v] = 1 for every arc u->v of */
/* the CSR (V, E), x = 1, x_new = alpha*A*x + b with b = beta[v] (beta may  */
/* be NULL: beta_scalar everywhere); the loop stops when sum|x_new - x| <   */
/* tol and keeps the iterate before that update; normalized != 0 divides by */
/* the L2 norm.  FP64 iterates and sums, alpha applied to each row sum.  *converged */
/* = 0 when max_iter passes without the test holding (the reference raises  */
/* RuntimeError); out (n doubles) is then unspecified.  Returns EGX_SUCC    */
/* whenever the iteration ran.                                              */
/* ------------------------------------------------------------------------ */
int egx_katz_csr(const int32_t *V, const int32_t *E, int32_t n, int32_t m, double alpha, const double *beta,
                 double beta_scalar, int32_t max_iter, double tol, int32_t normalized, double *out,
                 int32_t *iterations, int32_t *converged, double *kernel_seconds);

/* ------------------------------------------------------------------------ */
/* Eigenvector centrality on the same SpMV engine.  No counterpart on the    */
/* reference's GPU path; it restates the power iterations of the CPU         */
/* functions eigenvector.py:110-131 (Python route, nstart given) and         */
/* eigenvector.cpp:33-97 (C++ twin, cpp_eigenvector_centrality).  A[v][u] =  */
/* W[p] (1 when W is NULL) for every arc p: u -> v of the CSR (V, E), so row */
/* v of A x sums over v's in-arcs.  From the normalised start x_0:          */
/*   y = A x_{k-1} + shift x_{k-1};  if ||y||_2 < norm_floor: stop, keep     */
/*   x_{k-1};  x_k = y / ||y||_2;  if delta(x_k, x_{k-1}) < tol: stop        */
/* for k = 1..max_iter.  FP64 iterates and sums; weights are gathered as    */
/* FP32 values (the pull layout's value array).                              */
/* ------------------------------------------------------------------------ */
typedef struct egx_eig_opts {
    int32_t max_iter;   /* iterations k = 1..max_iter (0: x_0 is returned) */
    int32_t delta;      /* 1: sum|x_k - x_{k-1}| (eigenvector.cpp:86-93);
                           2: ||x_k - x_{k-1}||_2 (eigenvector.py:128)      */
    double tol;         /* >= 0 */
    double norm_floor;  /* > 0: 1e-12 (eigenvector.cpp:83), 1e-10 (eigenvector.py:122) */
    double shift;       /* >= 0: iterate A + shift I (0: the reference iteration) */
    int32_t start_norm; /* x_0 = x0 / sum|x0| (1, eigenvector.py:113), or
                           x0 / ||x0||_2, uniform 1/sqrt(n) when ||x0||_2 < 1e-12
                           (2, eigenvector.cpp:44-56)                       */
    int32_t pad;
} egx_eig_opts;

/* x0: n start values (caller order).  out: n doubles, the last normalised
 * iterate kept (x_k, or x_{k-1} on a norm stop).  *iterations: the k of the
 * stop (the reference loop's iteration count); *stop: 1 delta below tol,
 * 2 max_iter reached without it (eigenvector.py raises EasyGraphError,
 * eigenvector.cpp returns x), 3 norm below the floor (both return x_{k-1}).
 * Weights must be finite and non-negative; n = 0 writes nothing (*stop 0). */
int egx_eigenvector_csr(const int32_t *V, const int32_t *E, const double *W, int32_t n, int32_t m,
                        const double *x0, const egx_eig_opts *opts, double *out, int32_t *iterations,
                        int32_t *stop, double *kernel_seconds);

/* ------------------------------------------------------------------------ */
/* Serving (no counterpart in the reference, whose pagerank.py:141-179 solves */
/* one graph per call): a worker thread answers global and personalised      */
/* PageRank requests against a pool of device-resident graphs by continuous  */
/* batching.  Each graph with work gets one batch session (64 slots, own     */
/* stream); queued requests join running sessions at iteration-step          */
/* boundaries; sessions open while the HBM budget holds them and idle ones   */
/* close least-recently-used first; busy sessions on different graphs run    */
/* concurrently.  A result is bit-identical alone and in any mix.  A request */
/* identical to one still in flight on the same graph (same seed list,       */
/* alpha, max_iter and tol) shares that request's column and receives the    */
/* same result (request coalescing; coalesce_off = 1 turns it off).          */
/* ------------------------------------------------------------------------ */
typedef struct egx_server egx_server; /* opaque */

typedef struct {
    int64_t hbm_budget_bytes; /* session memory the server may hold; <= 0: 80% of free HBM at open */
    int32_t max_sessions;     /* concurrently open sessions (streams); <= 0: 64 */
    int32_t step_iters;       /* iterations enqueued per session between polls; <= 0: 4 */
    int32_t coalesce_off;     /* 1: every request takes a column of its own */
    int32_t pad;
} egx_server_opts;

typedef struct {
    int32_t iterations;
    int32_t width;            /* X row width (8..64 columns) of the session at completion */
    double queue_seconds;     /* submit -> admitted into a slot */
    double service_seconds;   /* admitted -> result in the caller's buffer */
    double latency_seconds;   /* submit -> result in the caller's buffer */
} egx_request_info;

typedef struct {
    int64_t submitted;
    int64_t completed;
    int64_t failed;
    int64_t steps;               /* scheduling rounds that launched work */
    int64_t sessions_opened;
    int64_t sessions_closed;
    int64_t peak_active_columns; /* requests resident on the device at once */
    int64_t peak_session_bytes;
    int64_t budget_bytes;
    int64_t largest_session_bytes; /* footprint charged for the pool's largest graph */
    int64_t coalesced;           /* requests answered by a column they shared with an identical one */
} egx_server_stats;

/* The pool's graphs must stay alive (and not be re-binned) until close. */
int egx_server_open(egx_graph *const *pool, int32_t count, const egx_server_opts *opts, egx_server **out);

/* Queue a request on pool[graph]: nseeds = 0 is global PageRank, else the
 * teleport and dangling mass go uniformly to the seed nodes.  PR (n doubles,
 * or NULL) receives the sum-normalised ranks; it must stay valid until
 * egx_server_wait returns for this request. */
int egx_server_submit(egx_server *s, int32_t graph, const int32_t *seeds, int32_t nseeds, double alpha,
                      int32_t max_iter, double tol, double *PR, int64_t *request_id);

/* Wait for a request (timeout_seconds < 0: no limit) and collect it; returns
 * the request's status.  A request is collected once. */
int egx_server_wait(egx_server *s, int64_t request_id, double timeout_seconds, egx_request_info *info);

int egx_server_stats_get(egx_server *s, egx_server_stats *out);

/* Stops the worker; requests still queued or running fail. */
void egx_server_close(egx_server *s);

#ifdef __cplusplus
}
#endif

#endif /* EGX_H */
