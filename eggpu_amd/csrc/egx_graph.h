// egx_graph.h -- the device-resident graph.
//
// Two layouts live in HBM:
//
// 1. The reference layout, bit-exact with build_inbound_from_csr
//    (Easy-Graph/gpu_easygraph/functions/centrality/pagerank.cu:286-347),
//    n nodes and m stored inbound arcs:
//      in_row  int32[n+1]  exclusive scan of in-degree            (:318-319)
//      in_col  int32[m]    sources, ascending within each row     (:328-346)
//      in_val  f32[m]      fl32(alpha*(w/outW[src])), weighted    (:344)
//      w_csc   f64[m]      sanitised weight per slot, weighted    (:311-312)
//      out_w   f64[n]      sanitised out-weight sum               (:311-313)
//      out_cnt int32[n]    valid out-arcs per node
//      deg     f32[n]      unweighted: out-degree; weighted: 1/0; 0 == dangling
//
// 2. The iteration layout that the SpMV streams: the same graph with
//    vertices relabelled by descending out-degree (ties by id), so the
//    sources that most arcs gather from sit in a short prefix of the
//    gathered vector and stay resident in each XCD's 4 MB L2.  perm[r] is
//    the original id of relabelled row r, rank[v] its inverse.  it_row /
//    it_col / it_val / it_w / it_outw / it_deg are the relabelled arrays.
//    With relabelling off, the iteration layout aliases the reference one.
//
// Bins index rows of the iteration layout, where hub rows (in-degree above
// the hub threshold) come first and every other row follows:
//      hsegs   HubSeg[]    fixed-size segments of hub rows, hrows HubRow[]
//      tiles   int32[]     merge-path split of the other rows: tile k holds the
//                          rows whose (row + arc) diagonal starts in
//                          [k*tile_items, (k+1)*tile_items)
//      t_col / t_pos / t_val  (tile_sort) each tile's arcs stably sorted by
//                          source, with the arc's slot in row order: lanes
//                          of one gather instruction then share cache lines
// Solve workspace: x[2], s[2] iterate and contribution (x/deg) vectors,
// double buffered, sized for FP64; partials (3 x grid, SoA), hub_part, out.
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "egx_common.h"

namespace egx {

struct HubSeg {
    int32_t row;    // destination row
    int32_t beg;    // first CSC slot of the segment
    int32_t end;    // one past the last slot
    int32_t slot;   // hub row index
    int32_t nseg;   // segments of this row
    int32_t first;  // index of the row's first segment in hub_part
};

struct HubRow {
    int32_t row;    // destination row
    int32_t first;  // first segment in hub_part
    int32_t nseg;   // number of segments
    int32_t pad;
};

struct PrState {
    double dang;      // dangling mass of the current iterate
    double err;       // L1(x' - x) of the last iteration
    double sum;       // sum of the last iterate
    int32_t iter;     // iterations performed
    int32_t done;     // stop flag, set on the device
    // eigenvector mode: the iterate buffers hold y = A x unnormalised, and
    // x = scale * y for the input buffer, scale_prev * y for the other one
    double scale;
    double scale_prev;
    int32_t eig_iters;  // iterations the reference loop counts at the stop
    int32_t eig_stop;   // 1 delta below tol, 2 max_iter, 3 norm below the floor
};

// LDS entries (of the iterate type) a tile may stage: tile_items + hub
// threshold must not exceed it.
constexpr int32_t kTileLds = 7936;

// Zeroed slots past the m arcs of the iteration arc arrays (in_col, p_col,
// p_nval): the width-64 SpMM reads arc indices and weights 8 at a time in
// scalar loads, up to two steps past a wavefront's last arc.
constexpr int64_t kArcPad = 64;

struct BinCfg {
    int32_t thread_row_max = 12;  // tiles with mean row length <= this: thread per row, else wavefront per row
    int32_t wave_row_max = 2048;  // hub threshold: rows with more arcs go to the hub bin
    int32_t hub_segment = 8192;   // arcs per hub segment (one workgroup)
    int32_t tile_items = 5888;    // rows + arcs per tile (merge-path split)
    int32_t reorder = 1;          // relabel by out-degree (hub rows always first)
    int32_t tile_sort = 0;        // gather each tile's arcs in source order (LDS scatter back to row order)
    int32_t hot_sources = 0;      // > 0: sources past this label are gathered non-temporally
    int32_t blocked = 0;          // propagation-blocked iterations (unweighted, FP32)
    int32_t blocked_shared = 0;   // blocked items sized for shared multi-graph launches
};

// Propagation-blocked iteration (bins.blocked).  Sources are cut into chunks
// of kPbChunk labels and destination rows into sub-bins of kPbRows rows.
// Every arc has a slot in a value array grouped by (sub-bin, chunk) and, in
// the same order inside each such run, (row, source).  Runs are padded to a
// multiple of kPbVec source slots (source kPbChunk, a zero LDS entry).  A
// scatter launch stages one chunk of s = x/deg in LDS and takes a
// contiguous range of that chunk's source slots, kPbVec at a time: each
// group's stretches of equal rows (pb_mask) are summed, one term per
// stretch, written from the group's first value slot (pb_vgrp) on.  Value
// slots are grouped by (sub-bin, chunk) and padded per run to a multiple of
// kPbVec (row 0, value 0).  A sum launch gives each sub-bin (or piece of a
// large one, at most kPbItem value slots) to one wavefront, which adds its
// terms into FP64 LDS row accumulators in slot order.  Sub-bins cut into
// several pieces hand per-piece row sums to the reduce launch as hub rows.
constexpr int32_t kPbChunk = 32768;
constexpr int32_t kPbRows = 2048;
constexpr int32_t kPbVec = 8;
// Groups per scatter window: a wavefront's 64 lanes take 64 consecutive
// groups; in the layouts of weighted graphs a row's stretch may run on
// across the groups of one window (one term), never across windows.
// Windows start at the chunk's first group plus a multiple of kPbWindow
// (scatter items are cut on that grid).
constexpr int32_t kPbWindow = 64;
constexpr int32_t kPbScatterItem = 262144;
constexpr int32_t kPbItem = 131072;
// PageRank graphs with at least this many arcs start with bins.blocked = 1
// (measured faster than the pull SpMV from R-MAT 18 up, weighted or not)
constexpr int64_t kPbDefaultArcs = (int64_t)1 << 23;

struct PbScatter {
    int32_t chunk;
    int32_t g0, g1;  // kPbVec-slot groups [g0, g1) of the chunk-major source array, all of this chunk
    int32_t pad;
};
struct PbSum {
    int32_t sub_bin;
    int32_t e0, e1;  // value slots [e0, e1)
    int32_t piece;   // -1: the whole sub-bin; else this piece's index among `pieces`
    int32_t pieces;
    int32_t part0;   // hub_part index of row 0 of the sub-bin (rows are `pieces` apart)
    int32_t pad0, pad1;
};

struct Graph {
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;

    int32_t n = 0;
    int64_t m = 0;          // stored inbound arcs
    int64_t n_dangling = 0;
    bool weighted = false;
    // adjacency values: every arc carries fl32(alpha * w) with unit weights
    // and no division by the source's out-weight (Katz: x' = alpha*A*x + b);
    // set before build_from_device_csr
    bool adjacency = false;
    // adjacency graph built with weights: the weighted kernels gather
    // fl32(alpha * w) (eigenvector centrality); false: unit weights
    bool adjacency_weights = false;
    // propagation-blocked layout (bins.blocked): pb_src uint16 chunk-local
    // source per slot in (chunk, sub-bin, row, source) order; per kPbVec-slot
    // group of it, pb_mask uint8 (bit j: a stretch of one row starts at slot
    // j) and pb_vgrp int32 (its first value slot); per value slot, in
    // (sub-bin, chunk) order, pb_dst uint16 (sub-bin-local row) and pb_vals
    // (the per-iteration FP32 term); weighted graphs add pb_w, the FP32 arc
    // value of every source slot (0 for padding), and pb_arc, its arc index
    // (-1 for padding) for refreshing pb_w after an alpha change
    bool pb_ready = false;
    DevBuf pb_src, pb_dst, pb_vals, pb_vgrp, pb_mask, pb_scatter, pb_sum, pb_hrows, pb_hub_part, pb_w, pb_arc;
    int64_t n_pb_scatter = 0, n_pb_sum = 0, n_pb_hrows = 0;
    // host copies of the items' sizes (groups per scatter item, value slots
    // per sum item), for ordering them in multi-graph launches
    std::vector<int32_t> pb_scatter_size, pb_sum_size;
    int32_t pb_shared = -1;  // item sizing the blocked layout was built with (BinCfg::blocked_shared)
    // the blocked layout merges stretches across groups (weighted solves);
    // only the kernels built for that merge read it
    bool pb_merged = false;
    double alpha_vals = 0.0;   // alpha baked into it_val
    double alpha_orig = 0.0;   // alpha baked into in_val
    int32_t max_in_degree = 0;
    int32_t max_out_degree = 0;
    double build_seconds = 0.0;
    double layout_seconds = 0.0;

    // reference layout
    DevBuf in_row, in_col, in_val, w_csc, out_w, out_cnt, deg;
    DevBuf csr_V, csr_E, csr_W;  // kept out-CSR (keep_csr)
    int64_t csr_m = 0;
    bool have_csr = false;

    // iteration layout
    bool reordered = false;
    DevBuf perm, rank;
    DevBuf p_row, p_col, p_val, p_w, p_outw, p_deg;
    DevBuf p_nval;                // w/outW per arc (no alpha), batched solves
    DevBuf tele;                  // personalised teleport weights (n doubles), single-vector solves
    int32_t open_batches = 0;     // batch sessions reading this graph
    const int32_t* it_row = nullptr;
    const int32_t* it_col = nullptr;
    float* it_val = nullptr;
    const double* it_w = nullptr;
    const double* it_outw = nullptr;
    const float* it_deg = nullptr;

    BinCfg bins;
    bool layout_built = false;
    int32_t layout_hub = -1;      // hub threshold the layout was built for
    int32_t layout_reorder = -1;
    int32_t n_hub = 0;            // rows [0, n_hub) of the iteration layout are hub rows
    DevBuf tiles;
    int64_t n_tiles = 0;
    // tile arcs in source order (tile_sort): t_col source, t_pos slot of the
    // arc in its tile's row-order LDS staging, t_val weight (weighted); same
    // arc index space as it_col, entries [row[n_hub], m) used
    DevBuf t_col, t_pos, t_val;
    bool tsorted = false;
    DevBuf hsegs, hrows, hub_part;
    int64_t n_hseg = 0, n_hrows = 0;
    int64_t rows_thread = 0, rows_wave = 0;

    // solve workspace (allocated on first use, sized for FP64)
    DevBuf x[2], s[2], state, partials, red_part, out;
    int64_t partial_blocks = 0;
    void* host_state = nullptr;  // pinned PrState
    // pinned staging of a packed PageRank result (kPackedHeader + 4 n bytes),
    // made on the first download and kept
    char* host_result = nullptr;

    ~Graph();
    size_t device_bytes() const;
};

// egx_build.hip.  With g.adjacency set, dW may be null (unit weights).
int build_from_device_csr(Graph& g, const int32_t* dV, const int32_t* dE, const double* dW, int32_t n,
                          int64_t m, double alpha);
int build_layout(Graph& g, const BinCfg& cfg);
int rebuild_bins(Graph& g, const BinCfg& cfg);
int refresh_alpha(Graph& g, double alpha);
int download_csc(Graph& g, int32_t* in_row, int32_t* in_col, float* in_val, int32_t* dangling01);

// egx_gen.hip: device generators producing an out-CSR (V, E, W) in device memory.
int gen_rmat_csr(int32_t n, int64_t nedges, uint64_t seed, uint32_t tA, uint32_t tAB, uint32_t tABC,
                 DevBuf& V, DevBuf& E, int64_t& m, hipStream_t s);
int gen_er_csr(int32_t n, int64_t nedges, uint64_t seed, DevBuf& V, DevBuf& E, int64_t& m, hipStream_t s);
int gen_edge_weights(const int32_t* dV, const int32_t* dE, int32_t n, int64_t m, uint64_t wseed, double lo,
                     double hi, DevBuf& W, hipStream_t s);

// egx_pagerank.hip
struct SolveParams {
    double alpha = 0.85;
    int32_t max_iter = 200;
    double tol = 1e-6;
    int32_t precision = EGX_PRECISION_FP32;
    int32_t check_every = 0;
    bool time_each_launch = false;  // hipEvent pair around every launch
    bool use_graph = false;         // replay captured runs of check_every iterations
    bool xcd_placement = false;     // multi-graph: keep a small graph's work items on one XCD
    // Katz mode (katz_centrality.py:81-110): x0 = 1, the teleport and
    // dangling terms are replaced by bias[row] (device, iteration-row order),
    // the stop test is sum|x' - x| < tol (not tol*n), and the result is the
    // iterate before the last update, unnormalised.
    const double* bias = nullptr;
    // personalised PageRank (networkx 3.4.2 semantics): per-row teleport
    // weights (device, iteration-row order, summing to 1) that receive both
    // the teleport and the dangling mass; null: uniform 1/n
    const double* tele = nullptr;
    // eigenvector mode (eigenvector.py:110-131, eigenvector.cpp:33-97): FP64,
    // adjacency graph, alpha 1.  x <- (A + eig_shift I) x / ||.||_2 from eig_x0
    // (device, caller order, already normalised); stop once the delta between
    // normalised iterates (eig_delta 1: L1, 2: L2) is below tol, after
    // max_iter iterations, or when ||A x|| < eig_floor (keeping x).  Each
    // iteration's delta is found in the following launch, so up to
    // max_iter + 1 launches run.
    int32_t eig_delta = 0;  // 0: not eigenvector mode
    double eig_floor = 0.0;
    double eig_shift = 0.0;
    const double* eig_x0 = nullptr;
    // per-solve workspace from caller memory instead of the graph's own
    // buffers (allocated on first use); with it the normalised result goes
    // to out_device (n doubles, caller order, device memory) and nothing is
    // copied to the host
    const struct SolveBufs* bufs = nullptr;
    double* out_device = nullptr;
};

// The per-solve workspace of one PageRank solve: two iterate buffers and two
// contribution buffers of n entries of the iterate type, the device state,
// per-block partial sums (3 per SpMV work item) and the reduce launch's
// partial sums (3 per reduce workgroup, at most solve_reduce_parts()).
struct SolveBufs {
    void* x[2];
    void* s[2];
    PrState* state;
    double* partials;
    double* red_part;
};
// Bytes a solve of graph g in this precision needs for its SolveBufs when
// they are carved from one region with carve_solve_bufs (256-byte aligned).
int64_t solve_scratch_bytes(const Graph& g, int32_t precision);
// Carves the SolveBufs of graph g out of base (at least
// solve_scratch_bytes(g, precision) bytes, 256-byte aligned).
void carve_solve_bufs(const Graph& g, int32_t precision, char* base, SolveBufs& b);
struct SolveStats {
    double kernel_seconds = 0.0;  // first to last iteration launch
    double spmv_seconds = 0.0;    // sum of k_pr_spmv launch durations
    double reduce_seconds = 0.0;  // sum of k_pr_reduce launch durations
    int32_t iterations = 0;
    int32_t spmv_launches = 0;
    int64_t spmv_grid = 0;
};
int solve_pagerank(Graph& g, const SolveParams& p, double* host_out, SolveStats* stats);
// g.tele := 1/k on the iteration rows of the k (distinct, in-range) seeds, 0 elsewhere.
int ensure_tele(Graph& g, const std::vector<int32_t>& seeds);
// Many graphs (same device, all weighted or all unweighted) in one launch per
// iteration on stream s; each graph stops on its own.
int solve_pagerank_multi(std::vector<Graph*>& gs, const SolveParams& p, double* const* host_out, int32_t* iters,
                         SolveStats* st, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1);

}  // namespace egx

struct egx_graph {
    egx::Graph g;
    // the batch session egx_graph_ppr keeps for its next call (its X buffers
    // and captured step graphs); closed before a re-bin and with the graph
    egx_batch* ppr_batch = nullptr;
    std::mutex ppr_mu;  // one egx_graph_ppr call at a time on this graph
    ~egx_graph();
};

namespace egx {
// A new, empty handle bound to the current device, with its stream and events.
int make_graph(std::unique_ptr<egx_graph>& out);

// egx_batch_result without the wait, and with half the copy: queues the
// packed result of `slot` on the batch's stream -- the column's FP64 sum s,
// then its FP32 iterate in caller order (n floats) -- copies it to host_out
// (pinned, at least kPackedHeader + 4 n bytes; may be null) and records
// `done` (may be null) after the copy.  unpack_result turns it into the
// doubles egx_batch_result returns, bit for bit.  host_out must stay valid
// until `done` completes.
constexpr size_t kPackedHeader = 16;  // the sum, padded so the floats start 16-byte aligned
int batch_result_enqueue(egx_batch* b, int32_t slot, void* host_out, hipEvent_t done);

// out[v] = (double)x[v] * (1.0 / s) if s > 0, else 1.0 / n: the arithmetic
// of the device's FP64 output (k_spmm_output), on a packed result.
inline void unpack_result(const void* packed, int32_t n, double* out) {
    const double s = *static_cast<const double*>(packed);
    const float* x = reinterpret_cast<const float*>(static_cast<const char*>(packed) + kPackedHeader);
    if (s > 0.0) {
        const double inv = 1.0 / s;
        for (int32_t v = 0; v < n; ++v) out[v] = (double)x[v] * inv;
    } else {
        const double u = 1.0 / (double)n;
        for (int32_t v = 0; v < n; ++v) out[v] = u;
    }
}

// One request's assignment for batch_assign_many.
struct BatchAssign {
    int32_t slot;
    const int32_t* seeds;  // caller node ids (null: global PageRank)
    int32_t nseeds;
    double alpha;
    int32_t max_iter;
    double tol;
};

// Assign several slots of a batch in one update (one copy and one launch)
// after widening the batch once, to its highest slot.  status[i] receives
// assignment i's code; one that fails its checks (slot range, seed range)
// is skipped and changes nothing.  Returns a code for the update itself.
int batch_assign_many(egx_batch* b, int32_t count, const BatchAssign* reqs, int* status);

// Release several slots in one update.
int batch_release_many(egx_batch* b, int32_t count, const int32_t* slots);

// Close the batch session egx_graph_ppr keeps on the handle, if any.
void drop_ppr_batch(egx_graph* gh);
}  // namespace egx

