// egx_pagerank.hip -- power iteration on the device-resident iteration layout.
//
// Three launches per iteration replace the reference's cuSPARSE SpMV, cuBLAS
// Sdot, two memsets, fused update/reduce kernel, scale kernel and two host
// round trips (pagerank.cu:591-664):
//
//   x'[v] = alpha * sum_{u->v} x[u]/outdeg[u] + jump          (unweighted)
//   x'[v] = sum_{u->v} in_val[u->v] * x[u] + jump             (weighted)
//   jump  = (1-alpha)/n + alpha*dang/n                        (pagerank.py:125)
//   err   = sum_v |x'[v] - x[v]|,  dang' = sum_{v dangling} x'[v]
//   stop when err < tol*n (pagerank.py:131), decided on the device.
//
// SpMV, one workgroup per work item:
//   * hub segment (rows with in-degree above the hub threshold): a fixed
//     slice of one row, 8 gathers in flight per thread, FP64 block sum
//     written to hub_part;
//   * tile (all other rows): the rows whose merge-path diagonal starts in
//     the tile.  Their arcs are contiguous; the block first stages every
//     gathered term into LDS (coalesced index loads, 8 independent gathers
//     per thread), then reduces rows from LDS -- one thread per row when the
//     tile's mean row length is at most thread_row_max, one wavefront per
//     row with a wave64 shuffle reduction otherwise -- and finalises them.
// Reduce (a few workgroups, then one wavefront) finalises hub rows from
// their segment partials in segment order and folds the per-block (err, sum,
// dang) slots in a fixed order: results are run-to-run deterministic and no
// launch uses a global atomic (device-scope atomics on one address serialise
// across the 8 XCDs).  Row sums, the teleport term, the dangling mass and the L1 residual
// are FP64; iterates are stored as T (float or double).  Unweighted graphs
// keep s = x/outdeg beside x, so each arc costs one 4-byte index and one
// gather.
//
// The ping-pong buffer of an iteration is chosen on the device from the
// graph's own iteration counter, so kernel arguments never change between
// iterations: a run of iterations can be captured once in a hipGraph and
// replayed, and many graphs -- each stopping on its own -- can share one
// launch through a table of per-graph arguments (k_pr_spmv_multi /
// k_pr_reduce_multi run the same block code as the one-graph kernels, so a
// graph's result is bit-identical alone and in a batch).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "egx_graph.h"

namespace egx {

namespace {

// 512 threads over one LDS tile: 8 wavefronts per SIMD at FP32 (LDS-limited to
// 5 with 256-thread blocks); the many small tiles of batched graphs gain most
constexpr int kThreads = 512;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kReduceThreads = 256;  // per reduce-part workgroup
constexpr int kRedItems = 512;       // hub rows + partial slots per reduce-part workgroup (target)
constexpr int kMaxRed = 1024;        // reduce-part workgroups per graph, at most
constexpr int kXcds = 8;             // MI355X: accelerator complex dies, one L2 each
constexpr int64_t kAffineBytes = (int64_t)2 << 20;  // gathered vector kept on one XCD up to this size
constexpr int kBatch = 8;  // gathers in flight per thread

template <typename T>
struct IterArgs {
    const int32_t* row;
    const int32_t* col;
    const float* val;
    const float* deg;
    const int32_t* tiles;
    const int32_t* tcol;    // tile arcs in source order (null: gather in row order)
    const uint16_t* tpos;   // their slots in the row-order LDS staging
    const float* tval;      // their weights (weighted)
    const HubSeg* hsegs;
    const HubRow* hrows;
    double* hub_part;
    T* xb[2];          // iterate buffers; state->iter parity selects the input
    T* sb[2];          // contribution buffers (unweighted)
    const T* x_in;     // bound per launch
    const T* s_in;
    T* x_out;
    T* s_out;
    PrState* state;
    double* partials;  // SoA: err[grid], sum[grid], dang[grid]
    double* red_part;  // SoA: err[n_red], sum[n_red], dang[n_red]
    int64_t n_hrows;
    int32_t n;
    int32_t thread_row_max;
    uint32_t nblk_h;
    uint32_t grid;
    uint32_t n_red;    // reduce-part workgroups
    int32_t cold_from; // first source label gathered non-temporally (n: none)
    const double* bias;  // per-row constant term replacing the teleport (Katz); null: PageRank
    const double* tele;  // per-row teleport weights summing to 1 (personalised); null: 1/n
    int32_t eig;         // eigenvector mode: 0 off, 1 L1 delta, 2 L2 delta
    double eig_floor;    // eigenvector: stop when ||A x||_2 < eig_floor
    double eig_shift;    // eigenvector: iterate A + eig_shift I
    const int32_t* perm; // iteration row -> caller id (null: identity); output only
    double* out;         // normalised result (n doubles); output only
    int64_t n_dangling;  // dangling rows, for the initial dangling mass
    double alpha;
    double tol_n;
    int32_t max_iter;
};

template <typename T>
__device__ __forceinline__ void bind(IterArgs<T>& a) {
    // selects, not an index: a runtime index would move the argument block to scratch
    const bool odd = (a.state->iter & 1) != 0;
    a.x_in = odd ? a.xb[1] : a.xb[0];
    a.x_out = odd ? a.xb[0] : a.xb[1];
    a.s_in = odd ? a.sb[1] : a.sb[0];
    a.s_out = odd ? a.sb[0] : a.sb[1];
}

// The gathered term of an arc from u: x[u] * in_val (weighted) or
// s[u] = x[u]/outdeg[u].
template <typename T, bool W>
__device__ __forceinline__ T term(const IterArgs<T>& a, int32_t u, float w) {
    // sources past the hot prefix are read non-temporally, so they do not
    // evict the prefix that most arcs gather from
    const T* v = W ? a.x_in : a.s_in;
    const T g = (u >= a.cold_from) ? __builtin_nontemporal_load(&v[u]) : v[u];
    return W ? (T)w * g : g;
}

template <typename T, bool W>
__device__ __forceinline__ void finalize_row(const IterArgs<T>& a, int32_t v, double acc, double jump, double& err,
                                             double& sum, double& dang) {
    if (a.eig) {
        // eigenvector: jump holds the input's scale; x = scale * y is the
        // normalised iterate, y' = A x + shift x is written unnormalised, and
        // the other buffer's old y (the iterate before x) gives the delta
        const double xc = jump * (double)a.x_in[v];
        const double yn = jump * acc + a.eig_shift * xc;
        const PrState* st = a.state;
        if (st->iter > 0) {
            const double d = xc - st->scale_prev * (double)a.x_out[v];
            err += (a.eig == 1) ? fabs(d) : d * d;
        }
        sum += yn * yn;
        a.x_out[v] = (T)yn;
        if (!W) a.s_out[v] = (T)yn;  // out-weight 1 on adjacency graphs
        return;
    }
    // uniform branches (kernel arguments): Katz bias, or personalised teleport
    // where jump holds (1-alpha) + alpha*dang
    if (a.bias) jump = a.bias[v];
    else if (a.tele) jump *= a.tele[v];
    const double xd = W ? (acc + jump) : (a.alpha * acc + jump);
    const T xn = (T)xd;
    const T xo = a.x_in[v];
    err += fabs((double)xn - (double)xo);
    sum += (double)xn;
    const float d = a.deg[v];
    if (d == 0.0f) dang += (double)xn;
    a.x_out[v] = xn;
    if (!W) a.s_out[v] = (d > 0.0f) ? (T)(xn / (T)d) : (T)0;
}

// (1-alpha)/n + alpha*dang/n (pagerank.py:125); with a personalised teleport
// p the teleport and the dangling mass both go to p (networkx 3.4.2), and
// this returns (1-alpha) + alpha*dang, scaled by p[v] per row.
template <typename T>
__device__ __forceinline__ double jump_of(const IterArgs<T>& a) {
    if (a.eig) return a.state->scale;
    if (a.tele) return (1.0 - a.alpha) + a.alpha * a.state->dang;
    const double n_d = (double)a.n;
    return (1.0 - a.alpha) / n_d + a.alpha * a.state->dang / n_d;
}

// Gather the terms of a tile's arcs [a0, a0+na) into lds[] in row order.
// S: the arcs are read in source order (neighbouring lanes then hit the same
// cache lines) and each term is written to its row-order slot.
template <typename T, bool W, bool S>
__device__ __forceinline__ void stage_tile(const IterArgs<T>& a, int32_t a0, int32_t na, T* lds) {
    const int32_t* col = S ? a.tcol : a.col;
    const float* val = S ? a.tval : a.val;
    const int t = threadIdx.x;
    for (int32_t base = 0; base < na; base += kThreads * kBatch) {
        int32_t u[kBatch];
        float w[kBatch];
        int32_t pos[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int32_t j = base + q * kThreads + t;
            const bool ok = j < na;
            u[q] = ok ? __builtin_nontemporal_load(&col[a0 + j]) : -1;
            w[q] = (W && ok) ? __builtin_nontemporal_load(&val[a0 + j]) : 0.0f;
            pos[q] = S ? (ok ? (int32_t)__builtin_nontemporal_load(&a.tpos[a0 + j]) : 0) : j;
        }
        T g[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) g[q] = (u[q] >= 0) ? term<T, W>(a, u[q], w[q]) : (T)0;
#pragma unroll
        for (int q = 0; q < kBatch; ++q)
            if (u[q] >= 0) lds[pos[q]] = g[q];
    }
}

// One SpMV work item.  col / val are streamed once per iteration:
// non-temporal loads keep them from evicting the gathered vector from L2.
template <typename T, bool W>
__device__ __forceinline__ void spmv_block(const IterArgs<T>& a, uint32_t b, T* lds, double* sh_red) {
    const double jump = jump_of(a);
    const int t = threadIdx.x;

    if (b < a.nblk_h) {
        const HubSeg sg = a.hsegs[b];
        double acc = 0.0;
        for (int32_t base = sg.beg; base < sg.end; base += kThreads * kBatch) {
            int32_t u[kBatch];
            float w[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
                const int32_t p = base + k * kThreads + t;
                const bool ok = p < sg.end;
                u[k] = ok ? __builtin_nontemporal_load(&a.col[p]) : -1;
                w[k] = (W && ok) ? __builtin_nontemporal_load(&a.val[p]) : 0.0f;
            }
            T g[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) g[k] = (u[k] >= 0) ? term<T, W>(a, u[k], w[k]) : (T)0;
#pragma unroll
            for (int k = 0; k < kBatch; ++k) acc += (double)g[k];
        }
        acc = wave_sum(acc);
        if (lane_id() == 0) sh_red[wave_id()] = acc;
        __syncthreads();
        if (t == 0) {
            double tot = 0.0;
            for (int w = 0; w < kWavesPerBlock; ++w) tot += sh_red[w];
            a.hub_part[b] = tot;
            a.partials[b] = 0.0;
            a.partials[(size_t)a.grid + b] = 0.0;
            a.partials[2 * (size_t)a.grid + b] = 0.0;
        }
        return;
    }

    // tile of whole rows
    const uint32_t k = b - a.nblk_h;
    const int32_t r0 = a.tiles[k], r1 = a.tiles[k + 1];
    const int32_t a0 = a.row[r0];
    const int32_t na = a.row[r1] - a0;
    const int32_t nr = r1 - r0;
    if (a.tpos) stage_tile<T, W, true>(a, a0, na, lds);
    else stage_tile<T, W, false>(a, a0, na, lds);
    __syncthreads();
    double err = 0.0, sum = 0.0, dang = 0.0;
    if ((int64_t)na > (int64_t)a.thread_row_max * nr) {
        // wavefront per row
        for (int32_t i = wave_id(); i < nr; i += kWavesPerBlock) {
            const int32_t v = r0 + i;
            const int32_t beg = a.row[v] - a0, end = a.row[v + 1] - a0;
            double acc = 0.0;
            for (int32_t j = beg + lane_id(); j < end; j += kWave) acc += (double)lds[j];
            acc = wave_sum(acc);
            if (lane_id() == 0) finalize_row<T, W>(a, v, acc, jump, err, sum, dang);
        }
    } else {
        // thread per row
        for (int32_t i = t; i < nr; i += kThreads) {
            const int32_t v = r0 + i;
            const int32_t beg = a.row[v] - a0, end = a.row[v + 1] - a0;
            double acc = 0.0;
            for (int32_t j = beg; j < end; ++j) acc += (double)lds[j];
            finalize_row<T, W>(a, v, acc, jump, err, sum, dang);
        }
    }
    block_sum3(err, sum, dang, sh_red);
    if (t == 0) {
        a.partials[b] = err;
        a.partials[(size_t)a.grid + b] = sum;
        a.partials[2 * (size_t)a.grid + b] = dang;
    }
}

// Reduce, in two launches.  Part r (of n_red workgroups) finalises its share
// of the hub rows from their segment partials (segment order) and folds its
// share of the per-block partial slots; the final launch (one wavefront)
// folds the n_red part sums in index order, updates the state and takes the
// stop decision err < tol*n (pagerank.py:131) on the device.  Every sum has
// a fixed order, so results are run-to-run deterministic.
template <typename T, bool W>
__device__ __forceinline__ void reduce_part_block(const IterArgs<T>& a, uint32_t r, double* sh_red) {
    const double jump = jump_of(a);
    const int t = threadIdx.x;
    const int64_t stride = (int64_t)a.n_red * kReduceThreads;
    const int64_t first = (int64_t)r * kReduceThreads + t;
    double e = 0.0, su = 0.0, dg = 0.0;
    for (int64_t h = first; h < a.n_hrows; h += stride) {
        const HubRow hr = a.hrows[h];
        double row = 0.0;
        for (int32_t q = 0; q < hr.nseg; ++q) row += a.hub_part[hr.first + q];
        finalize_row<T, W>(a, hr.row, row, jump, e, su, dg);
    }
    const double* pe = a.partials;
    const double* ps = a.partials + a.grid;
    const double* pd = a.partials + 2 * (size_t)a.grid;
    for (int64_t i = first; i < (int64_t)a.grid; i += stride) {
        e += pe[i];
        su += ps[i];
        dg += pd[i];
    }
    block_sum3(e, su, dg, sh_red);
    if (t == 0) {
        a.red_part[r] = e;
        a.red_part[a.n_red + r] = su;
        a.red_part[2 * (size_t)a.n_red + r] = dg;
    }
}

template <typename T>
__device__ __forceinline__ void reduce_final_wave(const IterArgs<T>& a) {
    const int l = lane_id();
    double e = 0.0, su = 0.0, dg = 0.0;
    for (uint32_t r = l; r < a.n_red; r += kWave) {
        e += a.red_part[r];
        su += a.red_part[a.n_red + r];
        dg += a.red_part[2 * (size_t)a.n_red + r];
    }
    e = wave_sum(e);
    su = wave_sum(su);
    dg = wave_sum(dg);
    if (l == 0 && a.eig) {
        // launch p found delta_p (x_p against x_{p-1}, p >= 1) and ||A x_p||;
        // the reference's iteration p tests the delta, iteration p+1 the norm
        PrState* st = a.state;
        const int32_t p = st->iter;
        const double delta = (a.eig == 2) ? sqrt(e) : e;
        const double nrm = sqrt(su);
        st->err = delta;
        st->sum = nrm;
        int32_t stop = 0, counted = p;
        if (p >= 1 && delta < a.tol_n) stop = 1;
        else if (p >= a.max_iter) stop = 2;
        else if (!(nrm >= a.eig_floor) || !(nrm > 0.0)) {
            stop = 3;
            counted = p + 1;
        }
        if (stop) {
            // x_p (input buffer, scale) is the result; iter keeps its parity
            st->eig_stop = stop;
            st->eig_iters = counted;
            st->done = 1;
        } else {
            st->scale_prev = st->scale;
            st->scale = 1.0 / nrm;
            st->iter = p + 1;
        }
    } else if (l == 0) {
        PrState* st = a.state;
        const int32_t it = st->iter + 1;
        st->err = e;
        st->sum = su;
        st->dang = dg;
        st->iter = it;
        st->done = (e < a.tol_n || it >= a.max_iter) ? 1 : 0;
    }
}

template <typename T, bool W>
__global__ __launch_bounds__(kThreads) void k_pr_spmv(IterArgs<T> a) {
    __shared__ T lds[kTileLds];
    __shared__ double sh_red[3 * kWavesPerBlock];
    if (a.state->done) return;  // uniform: written by an earlier launch
    bind(a);
    spmv_block<T, W>(a, blockIdx.x, lds, sh_red);
}

template <typename T, bool W>
__global__ __launch_bounds__(kReduceThreads) void k_pr_reduce(IterArgs<T> a) {
    __shared__ double sh_red[3 * (kReduceThreads / kWave)];
    if (a.state->done) return;
    bind(a);
    reduce_part_block<T, W>(a, blockIdx.x, sh_red);
}

template <typename T>
__global__ __launch_bounds__(kWave) void k_pr_final(IterArgs<T> a) {
    if (a.state->done) return;
    reduce_final_wave<T>(a);
}

// Propagation-blocked iteration (unweighted, FP32 iterates): the layout of
// PbPiece / PbScatter / PbSum in egx_graph.h.  No random global gathers:
// sources are read in chunks into LDS and their terms written out as runs,
// then every destination sub-bin sums its runs in LDS.
struct PbArgs {
    const uint16_t* src;
    const uint16_t* dst;
    float* vals;
    const int32_t* vgrp;
    const uint8_t* mask;
    const PbScatter* scatter;
    const PbSum* sums;
    const float* w;  // weighted graphs: FP32 arc value per source slot
};

constexpr int kPbThreads = 1024;  // one scatter workgroup per CU: its chunk takes 128 KB of LDS

// One scatter item: s of its chunk into LDS (16-byte loads; entry kPbChunk
// is the zero that padding slots read), then the item's groups of kPbVec
// source slots, one per lane: a 16-byte index load, kPbVec LDS reads, each
// stretch of equal rows summed in FP64 and emitted as one FP32 term.  A lane
// takes kPbGroups groups per batch, and works through two batches in turn:
// one batch's loads are in flight while the other is processed.  The terms
// of a wavefront's 64 groups usually form one contiguous range (groups of
// one run): they are then compacted in a per-wave LDS staging area (offsets
// from a wave prefix sum) and stored by consecutive lanes to consecutive
// slots.  Otherwise each lane stores its terms from its group's first value
// slot on.
constexpr int kPbGroups = 8;   // 4 measured 3% slower at R-MAT 24 (less of the latency covered)
constexpr int kPbGroupsW = 2;  // weighted: 32 more bytes per group (4 groups spill)
constexpr int kPbStage = 504;  // staged terms per wavefront (64 groups emit at most 512)
typedef uint32_t pb_u32x4 __attribute__((ext_vector_type(4)));
typedef float pb_f32x4 __attribute__((ext_vector_type(4)));

// Exclusive prefix sum over the wavefront of x in [0, 15], from one ballot
// per bit (no LDS traffic).
__device__ __forceinline__ int32_t wave_excl_scan_small(uint32_t x, int32_t& total) {
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint64_t bits = __ballot((x >> b) & 1u);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bits >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)bits, 0u));
        off += (int32_t)below << b;
        tot += __popcll(bits) << b;
    }
    total = tot;
    return off;
}

// The kPbVec staged sources of one group, times their arc values when
// weighted (W: the FP32 product, as in the pull kernel).
template <bool W>
__device__ __forceinline__ void pb_group_values(const float* sc, pb_u32x4 u, pb_f32x4 wa, pb_f32x4 wb,
                                                float (&v)[kPbVec]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        v[2 * h] = sc[w[h] & 0xffffu];
        v[2 * h + 1] = sc[w[h] >> 16];
    }
    if (W) {
        const float wt[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
#pragma unroll
        for (int k = 0; k < kPbVec; ++k) v[k] = wt[k] * v[k];
    }
}

// FP64 sum of the group's last stretch (from the last of mask bits 1..7,
// or slot 0, to the end).
__device__ __forceinline__ double pb_tail(const float (&v)[kPbVec], uint32_t mk) {
    double t = (double)v[0];
#pragma unroll
    for (int j = 1; j < kPbVec; ++j) t = ((mk >> j) & 1u) ? (double)v[j] : t + (double)v[j];
    return t;
}

// The terms of one group, written to o[0], o[1], ...: one per stretch
// (mask bit j >= 1: a stretch starts at slot j), summed in FP64.  Merged
// layouts (weighted graphs): the first stretch takes `carry`, the sum that
// the groups before carried into its row, and the last stretch has no term
// of its own when mask bit 0 is clear (it goes into the next group's first
// term).  Otherwise bit 0 is always set and there is no carry.
template <bool Merge, typename P>
__device__ __forceinline__ void pb_emit(const float (&v)[kPbVec], uint32_t mk, double carry, P* o) {
    int32_t k = 0;
    double run = Merge ? carry + (double)v[0] : (double)v[0];
#pragma unroll
    for (int j = 1; j < kPbVec; ++j) {
        if ((mk >> j) & 1u) {
            o[k++] = (float)run;
            run = (double)v[j];
        } else {
            run += (double)v[j];
        }
    }
    if (!Merge || (mk & 1u)) o[k] = (float)run;
}

// The same terms, written without branches for the LDS staging area: every
// slot writes the running sum of its stretch to the stretch's entry o[k], so
// a stretch's entry is written once per slot and, since a lane's LDS writes
// land in program order, ends with the stretch's full sum (the value pb_emit
// writes).  Entries past the group's popcount(mk) terms are never written.
template <bool Merge>
__device__ __forceinline__ void pb_emit_stage(const float (&v)[kPbVec], uint32_t mk, double carry, float* o) {
    const int32_t nt = __builtin_popcount(mk);
    int32_t k = 0;
    double run = Merge ? carry + (double)v[0] : (double)v[0];
    if (!Merge || nt > 0) o[0] = (float)run;
#pragma unroll
    for (int j = 1; j < kPbVec; ++j) {
        const uint32_t st = (mk >> j) & 1u;
        k += (int32_t)st;
        run = st ? (double)v[j] : run + (double)v[j];
        if (!Merge || k < nt) o[k] = (float)run;
    }
}

// A DPP move of a double (two 32-bit moves): lanes without a source, and
// the rows RowMask leaves out, read 0.
template <int Ctrl, int RowMask>
__device__ __forceinline__ double pb_dpp(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, Ctrl, RowMask, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), Ctrl, RowMask, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// One step of the chain scan: the lanes whose bit of B is set add the a
// that DPP control Ctrl brings them (0 where it brings none).
template <int Ctrl, int RowMask>
__device__ __forceinline__ void pb_chain_step(double& a, uint64_t B) {
    const double an = pb_dpp<Ctrl, RowMask>(a);
    a = __builtin_amdgcn_inverse_ballot_w64(B) ? a + an : a;
}

// The stretch sums that flow right across a wavefront's 64 consecutive
// groups (one window): lane l sends S_l = a_l + b_l * S_{l-1}, with a_l the
// sum of its last stretch when that stretch continues into the next group
// (mask bit 0 clear) and b_l = 1 when the whole group is that one stretch.
// An inclusive scan of the affine maps by DPP row shifts and row broadcasts
// (no LDS), in a fixed order, with the b of all lanes kept as one wave mask
// in scalar registers; returns S_{l-1}, the carry lane l adds to its first
// term (0 for lane 0).
__device__ __forceinline__ double pb_window_carry(double a, bool b) {
    uint64_t B = __ballot(b);
    // the lanes (l % 16) < d of every row, which a row shift by d leaves as they are
    constexpr uint64_t kRowHead1 = 0x0001000100010001ull, kRowHead2 = 0x0003000300030003ull;
    constexpr uint64_t kRowHead4 = 0x000f000f000f000full, kRowHead8 = 0x00ff00ff00ff00ffull;
    pb_chain_step<0x111, 0xf>(a, B);  // row_shr:1
    B &= (B << 1) | kRowHead1;
    pb_chain_step<0x112, 0xf>(a, B);  // row_shr:2
    B &= (B << 2) | kRowHead2;
    pb_chain_step<0x114, 0xf>(a, B);  // row_shr:4
    B &= (B << 4) | kRowHead4;
    pb_chain_step<0x118, 0xf>(a, B);  // row_shr:8
    B &= (B << 8) | kRowHead8;
    pb_chain_step<0x142, 0xa>(a, B);  // row_bcast:15: rows 1 and 3 take lanes 15 and 47
    B &= 0x0000ffff0000ffffull | (((B >> 15) & 1u) ? 0x00000000ffff0000ull : 0ull) |
         (((B >> 47) & 1u) ? 0xffff000000000000ull : 0ull);
    pb_chain_step<0x143, 0xc>(a, B);  // row_bcast:31: rows 2 and 3 take lane 31
    return pb_dpp<0x138, 0xf>(a);     // wave_shr:1
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The source slots of kPbGroups groups per lane, loaded together; the
// weighted arrays only when W.
// Weighted graphs carry 32 more bytes per group, so they keep fewer groups
// per batch (the double-buffered batches then fit in registers).
template <bool W>
struct PbScatterBatch {
    static constexpr int kG = W ? kPbGroupsW : kPbGroups;
    pb_u32x4 u[kG];
    pb_f32x4 wa[W ? kG : 1], wb[W ? kG : 1];
    int32_t vg[kG];
    uint32_t mk[kG];
};

// Every lane loads from an index inside the item (lanes past its end load
// its last group; the batch is told which lanes are past the end), so the
// loads are unconditional and nothing reads their results until the batch
// is processed: with loads under per-lane branches, or a select right after
// them, the compiler waits for every load in flight, the prefetched batch
// included.
template <bool W>
__device__ __forceinline__ void pb_scatter_load(const PbArgs& pb, int32_t g0, int32_t g1, int l,
                                                PbScatterBatch<W>& b) {
    const pb_u32x4* src = reinterpret_cast<const pb_u32x4*>(pb.src);
    const pb_f32x4* wts = reinterpret_cast<const pb_f32x4*>(pb.w);
#pragma unroll
    for (int q = 0; q < PbScatterBatch<W>::kG; ++q) {
        const int32_t gi = g0 + q * kPbThreads + l;
        const int32_t gc = gi < g1 ? gi : g1 - 1;
        b.u[q] = __builtin_nontemporal_load(&src[gc]);
        b.vg[q] = __builtin_nontemporal_load(&pb.vgrp[gc]);
        b.mk[q] = (uint32_t)__builtin_nontemporal_load(&pb.mask[gc]);
        if (W) {
            b.wa[q] = __builtin_nontemporal_load(&wts[2 * gc]);
            b.wb[q] = __builtin_nontemporal_load(&wts[2 * gc + 1]);
        }
    }
}

// The stretch terms of one batch: for each group, a wavefront prefix sum of
// the lanes' term counts; when the wavefront's terms form one contiguous
// range they are compacted in the wave's staging area and stored by
// consecutive lanes, otherwise each lane stores its own.
template <bool W>
__device__ __forceinline__ void pb_scatter_batch(const PbScatterBatch<W>& cur, int32_t g0, int32_t g1,
                                                 const float* sc, float* stg, float* vals, int l) {
#pragma unroll
    for (int q = 0; q < PbScatterBatch<W>::kG; ++q) {
        const bool ok = g0 + q * kPbThreads + l < g1;
        const int32_t vgq = cur.vg[q];
        const uint32_t mkq = ok ? cur.mk[q] : 0u;
        const pb_f32x4 waq = W ? cur.wa[q] : pb_f32x4{0.f, 0.f, 0.f, 0.f};
        const pb_f32x4 wbq = W ? cur.wb[q] : pb_f32x4{0.f, 0.f, 0.f, 0.f};
        // weighted graphs merge stretches across the window (the unweighted
        // kernel measured no faster for it: its registers are full); the
        // values are read where they are summed, so that the unweighted
        // kernel keeps its registers
        double carry = 0.0;
        if constexpr (W) {
            float v[kPbVec];
            pb_group_values<W>(sc, cur.u[q], waq, wbq, v);
            const bool goes_on = ok && !(mkq & 1u);
            carry = pb_window_carry(goes_on ? pb_tail(v, mkq) : 0.0, goes_on && (mkq >> 1) == 0u);
        }
        int32_t total = 0;
        const int32_t off = wave_excl_scan_small((uint32_t)__builtin_popcount(mkq), total);
        const int32_t first = __builtin_amdgcn_readfirstlane(vgq - off);  // lane 0: every lane runs here
        const bool contig = total > 0 && total <= kPbStage && __all(!ok || vgq - off == first);
        if (contig) {
            if (ok) {
                float v[kPbVec];
                pb_group_values<W>(sc, cur.u[q], waq, wbq, v);
                pb_emit_stage<W>(v, mkq, carry, stg + off);
            }
            wave_lds_sync();
            for (int32_t j = l; j < total; j += kWave) vals[first + j] = stg[j];
            wave_lds_sync();
        } else if (ok) {
            float v[kPbVec];
            pb_group_values<W>(sc, cur.u[q], waq, wbq, v);
            pb_emit<W>(v, mkq, carry, vals + vgq);
        }
    }
}

// Scatter item `item` of one graph, in a workgroup of kPbThreads threads
// with sc (kPbChunk + 4 floats) and stage (kPbThreads / kWave rows of
// kPbStage floats) in LDS.
template <typename T, bool W>
__device__ __forceinline__ void pb_scatter_item(const IterArgs<T>& a, const PbArgs& pb, uint32_t item, float* sc,
                                                float (*stage)[kPbStage]) {
    const PbScatter it = pb.scatter[item];
    const int l = lane_id();
    // every lane of a wavefront runs the loop the same number of times (the
    // bound depends on the wavefront's first thread only), so the shuffles
    // below see all 64 lanes
    const int32_t wave0 = __builtin_amdgcn_readfirstlane(it.g0 + (int32_t)(threadIdx.x - l));
    constexpr int32_t kStep = PbScatterBatch<W>::kG * kPbThreads;
    // the first batch of slots is in flight while the chunk is staged
    PbScatterBatch<W> nxt;
    pb_scatter_load<W>(pb, wave0, it.g1, l, nxt);
    const int32_t base = it.chunk * kPbChunk;
    const int32_t cnt = min(kPbChunk, a.n - base);
    // staged sources: x (weighted) or s = x/deg; 16-byte loads for the
    // whole quads (base is a multiple of kPbChunk, so they are aligned)
    const T* sv = W ? a.x_in : a.s_in;
    if constexpr (sizeof(T) == 4) {
        const int32_t nq = cnt / 4;
        const pb_f32x4* sv4 = reinterpret_cast<const pb_f32x4*>(sv + base);
        pb_f32x4* sc4 = reinterpret_cast<pb_f32x4*>(sc);
        for (int32_t i = threadIdx.x; i < nq; i += kPbThreads) sc4[i] = sv4[i];
        for (int32_t i = 4 * nq + threadIdx.x; i < cnt; i += kPbThreads) sc[i] = (float)sv[base + i];
    } else {
        for (int32_t i = threadIdx.x; i < cnt; i += kPbThreads) sc[i] = (float)sv[base + i];
    }
    if (threadIdx.x == 0) sc[kPbChunk] = 0.0f;
    __syncthreads();
    float* stg = stage[wave_id()];
    // software pipelined over two batches in turn: the next batch's loads
    // are issued before this batch's LDS reads and term stores, and neither
    // batch is copied
    PbScatterBatch<W> nx2;
    for (int32_t g0 = wave0; g0 < it.g1; g0 += 2 * kStep) {
        if (g0 + kStep < it.g1) pb_scatter_load<W>(pb, g0 + kStep, it.g1, l, nx2);
        pb_scatter_batch<W>(nxt, g0, it.g1, sc, stg, pb.vals, l);
        if (g0 + kStep >= it.g1) break;
        if (g0 + 2 * kStep < it.g1) pb_scatter_load<W>(pb, g0 + 2 * kStep, it.g1, l, nxt);
        pb_scatter_batch<W>(nx2, g0 + kStep, it.g1, sc, stg, pb.vals, l);
    }
}

template <typename T, bool W>
__global__ __launch_bounds__(kPbThreads) void k_pb_scatter(IterArgs<T> a, PbArgs pb) {
    __shared__ __align__(16) float sc[kPbChunk + 4];
    __shared__ float stage[kPbThreads / kWave][kPbStage];
    if (a.state->done) return;
    bind(a);
    pb_scatter_item<T, W>(a, pb, blockIdx.x, sc, stage);
}

// Many graphs, one launch: workgroup b runs scatter item item_local[b] of
// graph item_graph[b].
template <typename T, bool W>
__global__ __launch_bounds__(kPbThreads) void k_pb_scatter_multi(const IterArgs<T>* __restrict__ table,
                                                                 const PbArgs* __restrict__ pbs,
                                                                 const int32_t* __restrict__ item_graph,
                                                                 const uint32_t* __restrict__ item_local) {
    __shared__ __align__(16) float sc[kPbChunk + 4];
    __shared__ float stage[kPbThreads / kWave][kPbStage];
    const int32_t gi = item_graph[blockIdx.x];
    IterArgs<T> a = table[gi];
    if (a.state->done) return;
    bind(a);
    pb_scatter_item<T, W>(a, pbs[gi], item_local[blockIdx.x], sc, stage);
}

// One sum item (one workgroup of kPbSumWaves wavefronts): its value slots
// added into the sub-bin's LDS row accumulators, then the sub-bin's rows
// finalised (or, for a piece of a split sub-bin, its row sums left in
// hub_part for the reduce launch).  The accumulators are 64-bit unsigned
// fixed point in units of 2^-62: every FP32 term is converted exactly from
// its bits (truncated below 2^-62), and integer addition is associative, so
// the wavefronts (each taking a quarter of the item's slots) add in any
// order and the result is the same bits every run.  A lane adds the terms
// of each stretch of one row inside its 8-slot group in registers, then
// adds that with one LDS atomic.  Terms are non-negative and a row's sum is
// at most the total rank mass (about 1), far below the 4.0 the format
// holds; the truncation is below 2.2e-19 per term.  Padding slots add 0 to
// row 0.
constexpr int kPbSumWaves = 4;
constexpr int kPbSumGroups = 2;  // value groups per lane per batch (70 VGPRs: 7 wavefronts per SIMD)
constexpr int kPbSumThreads = kPbSumWaves * kWave;
constexpr double kPbUnfix = 1.0 / 4611686018427387904.0;  // 2^-62

// t * 2^62 as an integer, truncated, for FP32 t in [0, 4): t = m * 2^(e - 150)
// with m the significand (implicit bit included for normal numbers), so
// t * 2^62 = m * 2^(e - 88), one 64-bit shift.  Subnormals and zero give 0.
__device__ __forceinline__ unsigned long long pb_fix32(float t) {
    const uint32_t b = __float_as_uint(t);
    const int32_t e = (int32_t)(b >> 23);
    // subnormals and zero: significand 0, so every shift gives 0
    const unsigned long long m = (unsigned long long)((b & 0x7fffffu) | (e != 0 ? 0x800000u : 0u));
    const int32_t sh = e - 88;
    const unsigned long long up = m << (uint32_t)min(max(sh, 0), 40);
    const unsigned long long dn = m >> (uint32_t)min(max(-sh, 0), 63);
    return sh >= 0 ? up : dn;
}

// The value slots of kPbSumGroups 8-slot groups per lane, loaded together.
struct PbSumBatch {
    pb_u32x4 d[kPbSumGroups];
    pb_f32x4 va[kPbSumGroups], vb[kPbSumGroups];
};

// Value groups [base + g0, ...) of the whole value array (base: the
// wavefront's first group).  Unconditional loads, read only when the batch
// is processed (see pb_scatter_load): lanes past the end load group 0, which
// every layout has, and the processing skips them.
__device__ __forceinline__ void pb_sum_load(const pb_u32x4* dst, const pb_f32x4* val, int32_t base, int32_t g0,
                                            int32_t nv, int l, PbSumBatch& b) {
#pragma unroll
    for (int q = 0; q < kPbSumGroups; ++q) {
        const int32_t gi = g0 + q * kWave + l;
        const bool ok = gi < nv;
        const int32_t gc = ok ? base + gi : 0;
        b.d[q] = __builtin_nontemporal_load(&dst[gc]);
        b.va[q] = __builtin_nontemporal_load(&val[2 * gc]);
        b.vb[q] = __builtin_nontemporal_load(&val[2 * gc + 1]);
    }
}

// Finalise rows r0 + i, i = thread, thread + kPbSumThreads, ... < nr, from
// their fixed-point sums acc[i]; each thread loads the inputs of all its
// rows before it writes any of them.  The arithmetic is that of
// finalize_row.
template <typename T, bool W>
__device__ __forceinline__ void finalize_rows_fix(const IterArgs<T>& a, int32_t r0, int32_t nr,
                                                  const unsigned long long* acc, double jump, double& err,
                                                  double& sum, double& dang) {
    const int t = threadIdx.x;
    if (a.eig) {
        for (int32_t i = t; i < nr; i += kPbSumThreads)
            finalize_row<T, W>(a, r0 + i, (double)acc[i] * kPbUnfix, jump, err, sum, dang);
        return;
    }
    constexpr int kRowBatch = kPbRows / kPbSumThreads;
    T xo[kRowBatch];
    float d[kRowBatch];
    double jv[kRowBatch];
#pragma unroll
    for (int q = 0; q < kRowBatch; ++q) {
        const int32_t i = q * kPbSumThreads + t;
        const int32_t v = r0 + (i < nr ? i : 0);
        xo[q] = a.x_in[v];
        d[q] = a.deg[v];
        jv[q] = a.bias ? a.bias[v] : (a.tele ? jump * a.tele[v] : jump);
    }
#pragma unroll
    for (int q = 0; q < kRowBatch; ++q) {
        const int32_t i = q * kPbSumThreads + t;
        if (i < nr) {
            const int32_t v = r0 + i;
            const double rs = (double)acc[i] * kPbUnfix;
            const double xd = W ? (rs + jv[q]) : (a.alpha * rs + jv[q]);
            const T xn = (T)xd;
            err += fabs((double)xn - (double)xo[q]);
            sum += (double)xn;
            if (d[q] == 0.0f) dang += (double)xn;
            a.x_out[v] = xn;
            if (!W) a.s_out[v] = (d[q] > 0.0f) ? (T)(xn / (T)d[q]) : (T)0;
        }
    }
}

// Sum item `item` of one graph (partial slot `item`), in a workgroup of
// kPbSumThreads threads with acc (kPbRows) and sh_red (3 * kPbSumWaves) in LDS.
template <typename T, bool W>
__device__ __forceinline__ void pb_sum_item(const IterArgs<T>& a, const PbArgs& pb, uint32_t item,
                                            unsigned long long* acc, double* sh_red) {
    const PbSum it = pb.sums[item];
    const int t = threadIdx.x;
    const int l = lane_id();
    for (int32_t i = t; i < kPbRows; i += kPbSumThreads) acc[i] = 0ull;
    __syncthreads();
    // wavefront w takes value groups [q0, q1) of the item
    const int32_t ng = (it.e1 - it.e0) / kPbVec;
    const int32_t wv = __builtin_amdgcn_readfirstlane(wave_id());
    const int32_t q0 = (int32_t)(((int64_t)ng * wv) / kPbSumWaves);
    const int32_t q1 = (int32_t)(((int64_t)ng * (wv + 1)) / kPbSumWaves);
    const pb_u32x4* dst = reinterpret_cast<const pb_u32x4*>(pb.dst);
    const pb_f32x4* val = reinterpret_cast<const pb_f32x4*>(pb.vals);
    const int32_t base = it.e0 / kPbVec + q0;
    const int32_t nv = q1 - q0;
    // software pipelined: the next batch's loads are issued before this
    // batch is added (the loop bound is wavefront-uniform)
    PbSumBatch nxt;
    pb_sum_load(dst, val, base, 0, nv, l, nxt);
    for (int32_t g0 = 0; g0 < nv; g0 += kPbSumGroups * kWave) {
        const PbSumBatch cur = nxt;
        if (g0 + kPbSumGroups * kWave < nv) pb_sum_load(dst, val, base, g0 + kPbSumGroups * kWave, nv, l, nxt);
#pragma unroll
        for (int q = 0; q < kPbSumGroups; ++q) {
            if (g0 + q * kWave + l < nv) {
                const uint32_t w[4] = {cur.d[q].x, cur.d[q].y, cur.d[q].z, cur.d[q].w};
                const uint32_t r[8] = {w[0] & 0xffffu, w[0] >> 16, w[1] & 0xffffu, w[1] >> 16,
                                       w[2] & 0xffffu, w[2] >> 16, w[3] & 0xffffu, w[3] >> 16};
                const float v[8] = {cur.va[q].x, cur.va[q].y, cur.va[q].z, cur.va[q].w,
                                    cur.vb[q].x, cur.vb[q].y, cur.vb[q].z, cur.vb[q].w};
                // slots of one row are consecutive inside a run: add each
                // stretch in registers, one LDS add per stretch
                unsigned long long run = pb_fix32(v[0]);
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const unsigned long long f = pb_fix32(v[j + 1]);
                    const bool same = r[j + 1] == r[j];
                    if (!same) atomicAdd(&acc[r[j]], run);
                    run = same ? run + f : f;
                }
                atomicAdd(&acc[r[7]], run);
            }
        }
    }
    __syncthreads();
    const int32_t r0 = it.sub_bin * kPbRows;
    const int32_t nr = min(kPbRows, a.n - r0);
    double err = 0.0, sum = 0.0, dang = 0.0;
    if (it.piece < 0) {
        const double jump = jump_of(a);
        finalize_rows_fix<T, W>(a, r0, nr, acc, jump, err, sum, dang);
        block_sum3(err, sum, dang, sh_red);
    } else {
        for (int32_t i = t; i < nr; i += kPbSumThreads)
            a.hub_part[(size_t)it.part0 + (size_t)i * it.pieces + it.piece] = (double)acc[i] * kPbUnfix;
    }
    if (t == 0) {
        a.partials[item] = err;
        a.partials[(size_t)a.grid + item] = sum;
        a.partials[2 * (size_t)a.grid + item] = dang;
    }
}

template <typename T, bool W>
__global__ __launch_bounds__(kPbSumThreads) void k_pb_sum(IterArgs<T> a, PbArgs pb) {
    __shared__ unsigned long long acc[kPbRows];
    __shared__ double sh_red[3 * kPbSumWaves];
    if (a.state->done) return;
    bind(a);
    pb_sum_item<T, W>(a, pb, blockIdx.x, acc, sh_red);
}

// Many graphs, one launch: workgroup b runs sum item item_local[b] of graph
// item_graph[b].
template <typename T, bool W>
__global__ __launch_bounds__(kPbSumThreads) void k_pb_sum_multi(const IterArgs<T>* __restrict__ table,
                                                                const PbArgs* __restrict__ pbs,
                                                                const int32_t* __restrict__ item_graph,
                                                                const uint32_t* __restrict__ item_local) {
    __shared__ unsigned long long acc[kPbRows];
    __shared__ double sh_red[3 * kPbSumWaves];
    const int32_t gi = item_graph[blockIdx.x];
    IterArgs<T> a = table[gi];
    if (a.state->done) return;
    bind(a);
    pb_sum_item<T, W>(a, pbs[gi], item_local[blockIdx.x], acc, sh_red);
}

// Many graphs, one launch: block b works for graph blk_graph[b], as its
// local block blk_local[b]; blk_graph[b] < 0 is padding.
template <typename T, bool W>
__global__ __launch_bounds__(kThreads) void k_pr_spmv_multi(const IterArgs<T>* __restrict__ table,
                                                            const int32_t* __restrict__ blk_graph,
                                                            const uint32_t* __restrict__ blk_local) {
    __shared__ T lds[kTileLds];
    __shared__ double sh_red[3 * kWavesPerBlock];
    const int32_t gi = blk_graph[blockIdx.x];
    if (gi < 0) return;
    IterArgs<T> a = table[gi];
    if (a.state->done) return;
    bind(a);
    spmv_block<T, W>(a, blk_local[blockIdx.x], lds, sh_red);
}

template <typename T, bool W>
__global__ __launch_bounds__(kReduceThreads) void k_pr_reduce_multi(const IterArgs<T>* __restrict__ table,
                                                                    const int32_t* __restrict__ red_graph,
                                                                    const uint32_t* __restrict__ red_base) {
    __shared__ double sh_red[3 * (kReduceThreads / kWave)];
    const int32_t gi = red_graph[blockIdx.x];
    IterArgs<T> a = table[gi];
    if (a.state->done) return;
    bind(a);
    reduce_part_block<T, W>(a, blockIdx.x - red_base[gi], sh_red);
}

template <typename T>
__global__ __launch_bounds__(kWave) void k_pr_final_multi(const IterArgs<T>* __restrict__ table) {
    const IterArgs<T> a = table[blockIdx.x];
    if (a.state->done) return;
    reduce_final_wave<T>(a);
}

// Node-wise kernels over a table of graphs: graph i owns the blocks
// [nbase[i], nbase[i+1]) of the launch, 256 nodes each.
__device__ __forceinline__ int32_t graph_of_block(const uint32_t* __restrict__ nbase, int32_t count, uint32_t b) {
    int32_t lo = 0, hi = count;  // last i with nbase[i] <= b
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (nbase[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// x0 = 1/n, s0 = x0/deg and the initial state of every graph of the table
// (the multi-graph form of k_pr_init plus the host-written PrState).
template <typename T>
__global__ void k_pr_init_multi(const IterArgs<T>* __restrict__ table, const uint32_t* __restrict__ nbase,
                                int32_t count) {
    const int32_t gi = graph_of_block(nbase, count, blockIdx.x);
    const IterArgs<T>& a = table[gi];
    const int32_t v = (int32_t)(blockIdx.x - nbase[gi]) * blockDim.x + threadIdx.x;
    const T x0 = (T)(1.0 / (double)a.n);
    if (v < a.n) {
        a.xb[0][v] = x0;
        const float d = a.deg[v];
        a.sb[0][v] = (d > 0.0f) ? (T)(x0 / (T)d) : (T)0;
    }
    if (blockIdx.x == nbase[gi] && threadIdx.x == 0) {
        PrState st{};
        st.dang = (double)a.n_dangling * (double)x0;
        st.sum = 1.0;
        *a.state = st;
    }
}

// states[i] = *table[i].state: one copy to the host instead of one per graph.
template <typename T>
__global__ void k_gather_states(const IterArgs<T>* __restrict__ table, int32_t count, PrState* __restrict__ states) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) states[i] = *table[i].state;
}

// k_pr_output for every graph of the table.
template <typename T>
__global__ void k_pr_output_multi(const IterArgs<T>* __restrict__ table, const uint32_t* __restrict__ nbase,
                                  int32_t count) {
    const int32_t gi = graph_of_block(nbase, count, blockIdx.x);
    const IterArgs<T>& a = table[gi];
    const int32_t r = (int32_t)(blockIdx.x - nbase[gi]) * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    const double norm = a.state->sum;
    const T* x = (a.state->iter & 1) ? a.xb[1] : a.xb[0];
    const int32_t v = a.perm ? a.perm[r] : r;
    a.out[v] = (norm > 0.0) ? (double)x[r] * (1.0 / norm) : 1.0 / (double)a.n;
}

__global__ void k_fill_f64(double* __restrict__ out, int32_t n, double v) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// x0 = init (1/n for PageRank, 1 for Katz; as T), s0 = x0/deg.
template <typename T>
__global__ void k_pr_init(const float* __restrict__ deg, int32_t n, double init, T* __restrict__ x,
                          T* __restrict__ s) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const T x0 = (T)init;
    x[v] = x0;
    const float d = deg[v];
    s[v] = (d > 0.0f) ? (T)(x0 / (T)d) : (T)0;
}

// out[perm[r]] = x[r] * (1/sum) in FP64, uniform when sum <= 0
// (pagerank.py:134-139); perm maps iteration rows to caller node ids.
template <typename T>
__global__ void k_pr_output(const T* __restrict__ x0, const T* __restrict__ x1, const PrState* __restrict__ st,
                            const int32_t* __restrict__ perm, int32_t n, double* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double norm = st->sum;
    const T* x = (st->iter & 1) ? x1 : x0;
    const int32_t v = perm ? perm[r] : r;
    out[v] = (norm > 0.0) ? (double)x[r] * (1.0 / norm) : 1.0 / (double)n;
}

// eigenvector: x[r] = x0[perm[r]] (caller order, normalised), s = x
// (out-weight 1); state: no iterations, scale 1.
template <typename T>
__global__ void k_eig_init(const double* __restrict__ x0, const int32_t* __restrict__ perm, int32_t n,
                           T* __restrict__ x, T* __restrict__ s) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const T v = (T)x0[perm ? perm[r] : r];
    x[r] = v;
    s[r] = v;
}

// eigenvector result: out[perm[r]] = scale * y[r] for the stopping launch's
// input buffer (its iterate x_p).
template <typename T>
__global__ void k_eig_output(const T* __restrict__ x0, const T* __restrict__ x1, const PrState* __restrict__ st,
                             const int32_t* __restrict__ perm, int32_t n, double* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const T* x = (st->iter & 1) ? x1 : x0;
    const int32_t v = perm ? perm[r] : r;
    out[v] = st->scale * (double)x[r];
}

// Katz result: out[perm[r]] = the iterate before the last update
// (katz_centrality.py:95-99 keeps x when the test on x_new passes), in FP64.
template <typename T>
__global__ void k_katz_output(const T* __restrict__ x0, const T* __restrict__ x1, const PrState* __restrict__ st,
                              const int32_t* __restrict__ perm, int32_t n, double* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const T* x = (st->iter & 1) ? x0 : x1;
    const int32_t v = perm ? perm[r] : r;
    out[v] = (double)x[r];
}

inline unsigned nblk(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

int ensure_workspace(Graph& g, uint32_t grid) {
    const size_t nb = (size_t)std::max(g.n, 1) * 8;
    for (int i = 0; i < 2; ++i) {
        if (g.x[i].bytes < nb) EGX_TRY(g.x[i].alloc(nb));
        if (g.s[i].bytes < nb) EGX_TRY(g.s[i].alloc(nb));
    }
    if (g.out.bytes < nb) EGX_TRY(g.out.alloc(nb));
    if (g.state.bytes < sizeof(PrState)) EGX_TRY(g.state.alloc(sizeof(PrState)));
    if ((int64_t)grid > g.partial_blocks) {
        EGX_TRY(g.partials.alloc((size_t)grid * 3 * 8));
        g.partial_blocks = grid;
    }
    if (g.red_part.bytes < (size_t)kMaxRed * 3 * 8) EGX_TRY(g.red_part.alloc((size_t)kMaxRed * 3 * 8));
    if (!g.host_state) EGX_HIP_TRY(hipHostMalloc(&g.host_state, sizeof(PrState)), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
    return EGX_SUCC;
}

// Owns the per-launch timing events of one solve.
struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    ~LaunchTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    int reserve(size_t count) {
        while (ev.size() < count) {
            hipEvent_t e = nullptr;
            EGX_HIP_TRY(hipEventCreate(&e), EGX_DEVICE_ERR);
            ev.push_back(e);
        }
        return EGX_SUCC;
    }
};

// A captured run of iterations, replayed until every graph has stopped.
// Destroyed only after the stream it was launched on has drained.
struct CapturedRun {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    ~CapturedRun() {
        if (exec && stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

// Fills the launch arguments of graph g (refreshing its values for p.alpha
// on the graph's own stream; *refreshed reports whether that queued work).
template <typename T>
int fill_args(Graph& g, const SolveParams& p, IterArgs<T>& a, bool* refreshed = nullptr) {
    const uint32_t grid = (uint32_t)g.n_hseg + (uint32_t)g.n_tiles;
    if (grid == 0) {
        set_last_error("graph has no work items");
        return EGX_UNKNOW_ERROR;
    }
    const double before = g.alpha_vals;
    EGX_TRY(refresh_alpha(g, p.alpha));
    if (refreshed) *refreshed = g.weighted && before != g.alpha_vals;
    if (!p.bufs) EGX_TRY(ensure_workspace(g, grid));
    const int32_t n = g.n;
    a = IterArgs<T>{};
    a.row = g.it_row;
    a.col = g.it_col;
    a.val = g.it_val;
    a.deg = g.it_deg;
    a.tiles = g.tiles.as<int32_t>();
    if (g.tsorted) {
        a.tcol = g.t_col.as<int32_t>();
        a.tpos = g.t_pos.as<uint16_t>();
        a.tval = g.weighted ? g.t_val.as<float>() : nullptr;
    }
    a.hsegs = g.hsegs.as<HubSeg>();
    a.hrows = g.hrows.as<HubRow>();
    a.hub_part = g.hub_part.as<double>();
    if (p.bufs) {
        a.xb[0] = static_cast<T*>(p.bufs->x[0]);
        a.xb[1] = static_cast<T*>(p.bufs->x[1]);
        a.sb[0] = static_cast<T*>(p.bufs->s[0]);
        a.sb[1] = static_cast<T*>(p.bufs->s[1]);
        a.state = p.bufs->state;
        a.partials = p.bufs->partials;
        a.red_part = p.bufs->red_part;
    } else {
        a.xb[0] = g.x[0].as<T>();
        a.xb[1] = g.x[1].as<T>();
        a.sb[0] = g.s[0].as<T>();
        a.sb[1] = g.s[1].as<T>();
        a.state = g.state.as<PrState>();
        a.partials = g.partials.as<double>();
        a.red_part = g.red_part.as<double>();
    }
    a.n_hrows = g.n_hrows;
    a.n = n;
    a.thread_row_max = g.bins.thread_row_max;
    a.cold_from = (g.bins.hot_sources > 0 && g.bins.hot_sources < n) ? g.bins.hot_sources : n;
    a.nblk_h = (uint32_t)g.n_hseg;
    a.grid = grid;
    a.n_red = (uint32_t)std::min<int64_t>(kMaxRed, std::max<int64_t>(1, (g.n_hrows + grid + kRedItems - 1) / kRedItems));
    a.bias = p.bias;
    a.tele = p.tele;
    a.eig = p.eig_delta;
    a.eig_floor = p.eig_floor;
    a.eig_shift = p.eig_shift;
    a.perm = g.reordered ? g.perm.as<int32_t>() : nullptr;
    a.out = p.out_device ? p.out_device : g.out.as<double>();
    a.n_dangling = g.n_dangling;
    a.alpha = p.alpha;
    // Katz and eigenvector: absolute test
    a.tol_n = std::max(0.0, p.tol) * ((p.bias || p.eig_delta) ? 1.0 : (double)n);
    a.max_iter = p.max_iter;
    return EGX_SUCC;
}

// fill_args, then reset graph g's iterate to x0 and its state on stream s.
template <typename T>
int prepare_graph(Graph& g, const SolveParams& p, hipStream_t s, IterArgs<T>& a) {
    EGX_TRY(fill_args<T>(g, p, a));
    const int32_t n = g.n;
    if (p.eig_delta) {
        k_eig_init<T><<<nblk(n, 256), 256, 0, s>>>(p.eig_x0, g.reordered ? g.perm.as<int32_t>() : nullptr, n,
                                                   a.xb[0], a.sb[0]);
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        PrState* st0 = static_cast<PrState*>(g.host_state);
        *st0 = PrState{};
        st0->scale = 1.0;
        st0->scale_prev = 1.0;
        EGX_HIP_TRY(hipMemcpyAsync(a.state, st0, sizeof(PrState), hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
        return EGX_SUCC;
    }
    k_pr_init<T><<<nblk(n, 256), 256, 0, s>>>(g.it_deg, n, p.bias ? 1.0 : 1.0 / (double)n, a.xb[0], a.sb[0]);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    // staged through the pinned state buffer: the copy may run after return
    PrState* st0 = static_cast<PrState*>(g.host_state);
    *st0 = PrState{};
    st0->dang = (double)g.n_dangling * (double)(T)(1.0 / (double)n);
    st0->sum = 1.0;
    EGX_HIP_TRY(hipMemcpyAsync(a.state, st0, sizeof(PrState), hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

// The packed form of the FP32 PageRank result: the iterate's sum (FP64) in
// the header, then the FP32 iterate in caller order, half the bytes of the
// FP64 result; unpack_result finishes it on the host with k_pr_output's
// arithmetic, bit for bit.
__global__ void k_pr_output_packed(const float* __restrict__ x0, const float* __restrict__ x1,
                                   const PrState* __restrict__ st, const int32_t* __restrict__ perm, int32_t n,
                                   char* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) *reinterpret_cast<double*>(out) = st->sum;
    if (r >= n) return;
    const float* x = (st->iter & 1) ? x1 : x0;
    const int32_t v = perm ? perm[r] : r;
    reinterpret_cast<float*>(out + kPackedHeader)[v] = x[r];
}

// Writes the result of graph g to a.out (and copies it to host_out):
// PageRank's normalised iterate, Katz's unnormalised previous iterate, or
// the eigenvector iterate.
template <typename T>
int emit_result(const Graph& g, const IterArgs<T>& a, hipStream_t s, double* host_out, bool katz = false,
                bool eig = false) {
    const int32_t* perm = g.reordered ? g.perm.as<int32_t>() : nullptr;
    if (eig)
        k_eig_output<T><<<nblk(g.n, 256), 256, 0, s>>>(a.xb[0], a.xb[1], a.state, perm, g.n, a.out);
    else if (katz)
        k_katz_output<T><<<nblk(g.n, 256), 256, 0, s>>>(a.xb[0], a.xb[1], a.state, perm, g.n, a.out);
    else
        k_pr_output<T><<<nblk(g.n, 256), 256, 0, s>>>(a.xb[0], a.xb[1], a.state, perm, g.n, a.out);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    if (host_out)
        EGX_HIP_TRY(hipMemcpyAsync(host_out, a.out, (size_t)g.n * 8, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

template <typename T, bool W>
int solve_typed(Graph& g, const SolveParams& p, double* host_out, SolveStats* st) {
    hipStream_t s = g.stream;
    IterArgs<T> a;
    EGX_TRY(prepare_graph<T>(g, p, s, a));
    // propagation-blocked iterations: FP32 iterates on a graph with the
    // blocked layout (weighted: with its per-slot arc values); split
    // sub-bins go through the hub-row reduce
    constexpr bool kBlockable = std::is_same<T, float>::value;
    const bool blocked = kBlockable && g.pb_ready && (!W || g.pb_w.p != nullptr) && g.pb_merged == W;
    PbArgs pb{};
    if (blocked) {
        if (!p.bufs) {
            EGX_TRY(ensure_workspace(g, (uint32_t)g.n_pb_sum));
            a.partials = g.partials.as<double>();
        }
        a.hrows = g.pb_hrows.as<HubRow>();
        a.hub_part = g.pb_hub_part.as<double>();
        a.n_hrows = g.n_pb_hrows;
        a.nblk_h = 0;
        a.grid = (uint32_t)g.n_pb_sum;
        a.n_red = (uint32_t)std::min<int64_t>(
            kMaxRed, std::max<int64_t>(1, (a.n_hrows + a.grid + kRedItems - 1) / kRedItems));
        pb.src = g.pb_src.as<uint16_t>();
        pb.dst = g.pb_dst.as<uint16_t>();
        pb.vals = g.pb_vals.as<float>();
        pb.vgrp = g.pb_vgrp.as<int32_t>();
        pb.mask = g.pb_mask.as<uint8_t>();
        pb.scatter = g.pb_scatter.as<PbScatter>();
        pb.sums = g.pb_sum.as<PbSum>();
        pb.w = W ? g.pb_w.as<float>() : nullptr;
    }
    const uint32_t grid = a.grid;
    // the SpMV of one iteration: pull kernel, or scatter + sum
    auto launch_spmv = [&]() {
        if constexpr (kBlockable) {
            if (blocked) {
                k_pb_scatter<T, W><<<(uint32_t)g.n_pb_scatter, kPbThreads, 0, s>>>(a, pb);
                k_pb_sum<T, W><<<(uint32_t)g.n_pb_sum, kPbSumThreads, 0, s>>>(a, pb);
                return;
            }
        }
        k_pr_spmv<T, W><<<grid, kThreads, 0, s>>>(a);
    };

    // eigenvector mode finds iteration k's delta in launch k+1
    const int32_t max_launch = p.max_iter + (p.eig_delta ? 1 : 0);
    LaunchTimer timer;
    if (p.time_each_launch) EGX_TRY(timer.reserve((size_t)4 * max_launch));
    CapturedRun cap;
    cap.stream = s;
    EGX_HIP_TRY(hipEventRecord(g.ev0, s), EGX_DEVICE_ERR);
    int32_t launched = 0;
    int32_t chunk = p.check_every > 0 ? p.check_every : 4;
    const bool can_stop_early = p.tol > 0.0 || p.eig_delta != 0;
    PrState* hs = static_cast<PrState*>(g.host_state);
    while (launched < max_launch) {
        const int32_t todo = can_stop_early ? std::min(chunk, max_launch - launched) : (max_launch - launched);
        if (p.use_graph && !p.time_each_launch && can_stop_early && todo == p.check_every) {
            // replay one captured chunk of `check_every` iterations
            if (!cap.exec) {
                EGX_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), EGX_DEVICE_ERR);
                for (int32_t q = 0; q < todo; ++q) {
                    launch_spmv();
                    k_pr_reduce<T, W><<<a.n_red, kReduceThreads, 0, s>>>(a);
                    k_pr_final<T><<<1, kWave, 0, s>>>(a);
                }
                EGX_HIP_TRY(hipStreamEndCapture(s, &cap.graph), EGX_DEVICE_ERR);
                EGX_HIP_TRY(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0), EGX_DEVICE_ERR);
            }
            EGX_HIP_TRY(hipGraphLaunch(cap.exec, s), EGX_DEVICE_ERR);
            launched += todo;
        } else {
            for (int32_t q = 0; q < todo; ++q, ++launched) {
                if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched], s), EGX_DEVICE_ERR);
                launch_spmv();
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
                if (p.time_each_launch) {
                    EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 1], s), EGX_DEVICE_ERR);
                    EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 2], s), EGX_DEVICE_ERR);
                }
                k_pr_reduce<T, W><<<a.n_red, kReduceThreads, 0, s>>>(a);
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
                k_pr_final<T><<<1, kWave, 0, s>>>(a);
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
                if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 3], s), EGX_DEVICE_ERR);
            }
        }
        if (launched >= max_launch) break;
        EGX_HIP_TRY(hipMemcpyAsync(hs, a.state, sizeof(PrState), hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
        if (hs->done) break;
        if (p.check_every <= 0) chunk = std::min(chunk * 2, 64);
    }
    EGX_HIP_TRY(hipEventRecord(g.ev1, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(hs, a.state, sizeof(PrState), hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventSynchronize(g.ev1), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    const int32_t iters = hs->iter;
    if (st) {
        float ms = 0.0f;
        EGX_HIP_TRY(hipEventElapsedTime(&ms, g.ev0, g.ev1), EGX_DEVICE_ERR);
        st->kernel_seconds = (double)ms * 1e-3;
        st->iterations = iters;
        st->spmv_grid = grid;
        if (p.time_each_launch) {
            // launches after the stop flag return at once; time the real ones
            for (int32_t q = 0; q < iters; ++q) {
                float a_ms = 0.0f, b_ms = 0.0f;
                EGX_HIP_TRY(hipEventElapsedTime(&a_ms, timer.ev[4 * q], timer.ev[4 * q + 1]), EGX_DEVICE_ERR);
                EGX_HIP_TRY(hipEventElapsedTime(&b_ms, timer.ev[4 * q + 2], timer.ev[4 * q + 3]), EGX_DEVICE_ERR);
                st->spmv_seconds += (double)a_ms * 1e-3;
                st->reduce_seconds += (double)b_ms * 1e-3;
            }
            st->spmv_launches = iters;
        }
    }
    if constexpr (std::is_same<T, float>::value) {
        // a PageRank result for the host only crosses PCIe packed (FP32
        // iterate and its sum) and is finished here; a.out, the graph's own
        // buffer of n >= 4 doubles, has room for the packed form
        if (host_out && p.out_device == nullptr && p.bias == nullptr && p.eig_delta == 0 && g.n >= 4) {
            const size_t bytes = kPackedHeader + (size_t)g.n * 4;
            if (!g.host_result)
                EGX_HIP_TRY(hipHostMalloc((void**)&g.host_result, bytes), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
            k_pr_output_packed<<<nblk(g.n, 256), 256, 0, s>>>(a.xb[0], a.xb[1], a.state,
                                                             g.reordered ? g.perm.as<int32_t>() : nullptr, g.n,
                                                             reinterpret_cast<char*>(a.out));
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            EGX_HIP_TRY(hipMemcpyAsync(g.host_result, a.out, bytes, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
            EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
            unpack_result(g.host_result, g.n, host_out);
            return EGX_SUCC;
        }
    }
    EGX_TRY(emit_result<T>(g, a, s, host_out, p.bias != nullptr, p.eig_delta != 0));
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

template <typename T, bool W>
int solve_multi_typed(std::vector<Graph*>& gs, const SolveParams& p, double* const* host_out, int32_t* iters,
                      SolveStats* st, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    const int32_t count = (int32_t)gs.size();
    std::vector<IterArgs<T>> table((size_t)count);
    std::vector<uint32_t> base((size_t)count);
    uint64_t total = 0;
    // work still queued on a graph's own stream (none after the synchronous
    // entry points) must land before this launch reads the graph
    for (Graph* g : gs) {
        const hipError_t q = hipStreamQuery(g->stream);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            EGX_HIP_TRY(hipStreamSynchronize(g->stream), EGX_DEVICE_ERR);
        } else if (q != hipSuccess) {
            set_last_error(hipGetErrorString(q));
            return EGX_DEVICE_ERR;
        }
    }
    std::vector<Graph*> refreshed;
    std::vector<uint32_t> nbase((size_t)count + 1, 0);  // node blocks of the init / output launches
    for (int32_t i = 0; i < count; ++i) {
        bool r = false;
        EGX_TRY(fill_args<T>(*gs[(size_t)i], p, table[(size_t)i], &r));
        if (r) refreshed.push_back(gs[(size_t)i]);
        base[(size_t)i] = (uint32_t)total;
        total += table[(size_t)i].grid;
        nbase[(size_t)i + 1] = nbase[(size_t)i] + nblk(gs[(size_t)i]->n, 256);
    }
    // Members with a propagation-blocked layout (FP32 iterates; weighted ones
    // with their per-slot arc values) iterate as they do alone: their
    // scatter items and sum items go into two table-driven launches, and
    // only the others' work items into the pull launch.
    constexpr bool kBlockable = std::is_same<T, float>::value;
    std::vector<PbArgs> pbt((size_t)count);
    std::vector<char> is_pb((size_t)count, 0);
    std::vector<int32_t> sc_graph, sum_graph;
    std::vector<uint32_t> sc_local, sum_local;
    struct PbItem {
        int32_t size, graph;
        uint32_t local;
    };
    std::vector<PbItem> sc_items, sum_items;
    if constexpr (kBlockable) {
        for (int32_t i = 0; i < count; ++i) {
            Graph& g = *gs[(size_t)i];
            if (!g.pb_ready || (W && g.pb_w.p == nullptr) || g.pb_merged != W) continue;
            is_pb[(size_t)i] = 1;
            IterArgs<T>& a = table[(size_t)i];
            total -= a.grid;
            EGX_TRY(ensure_workspace(g, (uint32_t)g.n_pb_sum));
            a.partials = g.partials.as<double>();
            a.hrows = g.pb_hrows.as<HubRow>();
            a.hub_part = g.pb_hub_part.as<double>();
            a.n_hrows = g.n_pb_hrows;
            a.nblk_h = 0;
            a.grid = (uint32_t)g.n_pb_sum;
            a.n_red = (uint32_t)std::min<int64_t>(
                kMaxRed, std::max<int64_t>(1, (a.n_hrows + a.grid + kRedItems - 1) / kRedItems));
            PbArgs& pb = pbt[(size_t)i];
            pb.src = g.pb_src.as<uint16_t>();
            pb.dst = g.pb_dst.as<uint16_t>();
            pb.vals = g.pb_vals.as<float>();
            pb.vgrp = g.pb_vgrp.as<int32_t>();
            pb.mask = g.pb_mask.as<uint8_t>();
            pb.scatter = g.pb_scatter.as<PbScatter>();
            pb.sums = g.pb_sum.as<PbSum>();
            pb.w = W ? g.pb_w.as<float>() : nullptr;
            for (int64_t k = 0; k < g.n_pb_scatter; ++k)
                sc_items.push_back({g.pb_scatter_size[(size_t)k], i, (uint32_t)k});
            for (int64_t k = 0; k < g.n_pb_sum; ++k) sum_items.push_back({g.pb_sum_size[(size_t)k], i, (uint32_t)k});
        }
        // largest items first, so the small ones fill in behind them rather
        // than a large one running alone at the end of the launch
        auto by_size = [](const PbItem& x, const PbItem& y) { return x.size > y.size; };
        std::stable_sort(sc_items.begin(), sc_items.end(), by_size);
        std::stable_sort(sum_items.begin(), sum_items.end(), by_size);
        for (const PbItem& it : sc_items) {
            sc_graph.push_back(it.graph);
            sc_local.push_back(it.local);
        }
        for (const PbItem& it : sum_items) {
            sum_graph.push_back(it.graph);
            sum_local.push_back(it.local);
        }
    }
    if (total > 0x7FFFFFFFull || sc_graph.size() > 0x7FFFFFFFull || sum_graph.size() > 0x7FFFFFFFull) {
        set_last_error("too many work items for one launch");
        return EGX_UNSUPPORTED_GRAPH;
    }
    // Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs
    // on XCD b % 8), and each XCD has its own L2.  With XCD placement
    // (EGX_SOLVE_XCD_PLACEMENT) a graph whose gathered vector fits in one L2
    // gets all its work items on one XCD (the least loaded, largest graphs
    // first), so that vector is cached once instead of eight times; larger
    // graphs are spread over every XCD.  Otherwise work items go out graph
    // after graph.
    std::vector<std::vector<std::pair<int32_t, uint32_t>>> lanes(kXcds);
    {
        uint64_t next_lane = 0;
        std::vector<int32_t> order((size_t)count);
        for (int32_t i = 0; i < count; ++i) order[(size_t)i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return table[(size_t)x].grid > table[(size_t)y].grid;
        });
        std::vector<uint64_t> load(kXcds, 0);
        const bool placement = p.xcd_placement;
        for (int32_t gi : order) {
            if (is_pb[(size_t)gi]) continue;
            const uint32_t grid = table[(size_t)gi].grid;
            const bool affine = placement && (int64_t)gs[(size_t)gi]->n * (int64_t)sizeof(T) <= kAffineBytes;
            if (affine) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                for (uint32_t k = 0; k < grid; ++k) lanes[(size_t)x].emplace_back(gi, k);
                load[(size_t)x] += grid;
            } else if (placement) {
                for (uint32_t k = 0; k < grid; ++k) {
                    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                    lanes[(size_t)x].emplace_back(gi, k);
                    load[(size_t)x] += 1;
                }
            } else {
                // plain order: graph after graph, work items dealt round-robin
                for (uint32_t k = 0; k < grid; ++k) {
                    const int x = (int)(next_lane++ % kXcds);
                    lanes[(size_t)x].emplace_back(gi, k);
                }
            }
        }
    }
    size_t depth = 0;
    for (const auto& l : lanes) depth = std::max(depth, l.size());
    total = (uint64_t)depth * kXcds;
    if (total > 0x7FFFFFFFull) {
        set_last_error("too many work items for one launch");
        return EGX_UNSUPPORTED_GRAPH;
    }
    std::vector<int32_t> blk((size_t)total, -1);
    std::vector<uint32_t> blk_local((size_t)total, 0);
    for (int x = 0; x < kXcds; ++x)
        for (size_t q = 0; q < lanes[(size_t)x].size(); ++q) {
            blk[q * kXcds + (size_t)x] = lanes[(size_t)x][q].first;
            blk_local[q * kXcds + (size_t)x] = lanes[(size_t)x][q].second;
        }
    // reduce-part workgroups: red_graph / red_base play the part of blk / base
    std::vector<uint32_t> rbase((size_t)count);
    uint64_t rtotal = 0;
    for (int32_t i = 0; i < count; ++i) {
        rbase[(size_t)i] = (uint32_t)rtotal;
        rtotal += table[(size_t)i].n_red;
    }
    std::vector<int32_t> rblk((size_t)rtotal);
    for (int32_t i = 0; i < count; ++i)
        std::fill(rblk.begin() + rbase[(size_t)i], rblk.begin() + rbase[(size_t)i] + table[(size_t)i].n_red, i);
    DevBuf d_table, d_blk, d_base, d_rblk, d_rbase, d_nbase, d_states, d_pbt, d_scg, d_scl, d_sumg, d_suml;
    const uint32_t n_sc = (uint32_t)sc_graph.size(), n_sum = (uint32_t)sum_graph.size();
    if (n_sc > 0) {
        EGX_TRY(d_pbt.alloc(sizeof(PbArgs) * (size_t)count));
        EGX_TRY(d_scg.alloc((size_t)n_sc * 4));
        EGX_TRY(d_scl.alloc((size_t)n_sc * 4));
        EGX_TRY(d_sumg.alloc((size_t)n_sum * 4));
        EGX_TRY(d_suml.alloc((size_t)n_sum * 4));
        EGX_HIP_TRY(hipMemcpyAsync(d_pbt.p, pbt.data(), sizeof(PbArgs) * (size_t)count, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(d_scg.p, sc_graph.data(), (size_t)n_sc * 4, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(d_scl.p, sc_local.data(), (size_t)n_sc * 4, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(d_sumg.p, sum_graph.data(), (size_t)n_sum * 4, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(d_suml.p, sum_local.data(), (size_t)n_sum * 4, hipMemcpyHostToDevice, s),
                    EGX_DEVICE_ERR);
    }
    EGX_TRY(d_table.alloc(sizeof(IterArgs<T>) * (size_t)count));
    EGX_TRY(d_nbase.alloc(nbase.size() * 4));
    EGX_TRY(d_states.alloc(sizeof(PrState) * (size_t)count));
    EGX_HIP_TRY(hipMemcpyAsync(d_nbase.p, nbase.data(), nbase.size() * 4, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    EGX_TRY(d_blk.alloc(blk.size() * 4));
    EGX_TRY(d_base.alloc(blk_local.size() * 4));
    EGX_TRY(d_rblk.alloc(rblk.size() * 4));
    EGX_TRY(d_rbase.alloc(rbase.size() * 4));
    EGX_HIP_TRY(hipMemcpyAsync(d_rblk.p, rblk.data(), rblk.size() * 4, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(d_rbase.p, rbase.data(), rbase.size() * 4, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(d_table.p, table.data(), sizeof(IterArgs<T>) * (size_t)count, hipMemcpyHostToDevice, s),
                EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(d_blk.p, blk.data(), blk.size() * 4, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(d_base.p, blk_local.data(), blk_local.size() * 4, hipMemcpyHostToDevice, s),
                EGX_DEVICE_ERR);
    // per-graph states are gathered into one pinned copy for polling
    void* pinned = nullptr;
    EGX_HIP_TRY(hipHostMalloc(&pinned, sizeof(PrState) * (size_t)count), EGX_FAILED_TO_ALLOCATE_HOST_MEM);
    struct PinnedGuard {
        void* p;
        ~PinnedGuard() { (void)hipHostFree(p); }
    } pg{pinned};
    PrState* hs = static_cast<PrState*>(pinned);

    const int32_t chunk = p.check_every > 0 ? p.check_every : 8;
    // the SpMV of one iteration for every member: blocked members' scatter,
    // the others' pull work items, then the blocked members' sums
    auto launch_spmv = [&]() {
        if constexpr (kBlockable) {
            if (n_sc > 0)
                k_pb_scatter_multi<T, W><<<n_sc, kPbThreads, 0, s>>>(d_table.as<IterArgs<T>>(), d_pbt.as<PbArgs>(),
                                                                      d_scg.as<int32_t>(), d_scl.as<uint32_t>());
        }
        if (total > 0)
            k_pr_spmv_multi<T, W><<<(unsigned)total, kThreads, 0, s>>>(d_table.as<IterArgs<T>>(), d_blk.as<int32_t>(),
                                                                      d_base.as<uint32_t>());
        if constexpr (kBlockable) {
            if (n_sum > 0)
                k_pb_sum_multi<T, W><<<n_sum, kPbSumThreads, 0, s>>>(d_table.as<IterArgs<T>>(), d_pbt.as<PbArgs>(),
                                                                     d_sumg.as<int32_t>(), d_suml.as<uint32_t>());
        }
    };
    CapturedRun cap;
    cap.stream = s;
    // alpha refreshes run on each graph's own stream
    for (Graph* g : refreshed) EGX_HIP_TRY(hipStreamSynchronize(g->stream), EGX_DEVICE_ERR);
    // every iterate and state reset in one launch
    k_pr_init_multi<T><<<nbase[(size_t)count], 256, 0, s>>>(d_table.as<IterArgs<T>>(), d_nbase.as<uint32_t>(),
                                                              count);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipEventRecord(ev0, s), EGX_DEVICE_ERR);
    int32_t launched = 0;
    while (launched < p.max_iter) {
        const int32_t todo = std::min(chunk, p.max_iter - launched);
        if (p.use_graph && todo == chunk) {
            if (!cap.exec) {
                EGX_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), EGX_DEVICE_ERR);
                for (int32_t q = 0; q < chunk; ++q) {
                    launch_spmv();
                    k_pr_reduce_multi<T, W><<<(unsigned)rtotal, kReduceThreads, 0, s>>>(
                        d_table.as<IterArgs<T>>(), d_rblk.as<int32_t>(), d_rbase.as<uint32_t>());
                    k_pr_final_multi<T><<<(unsigned)count, kWave, 0, s>>>(d_table.as<IterArgs<T>>());
                }
                EGX_HIP_TRY(hipStreamEndCapture(s, &cap.graph), EGX_DEVICE_ERR);
                EGX_HIP_TRY(hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0), EGX_DEVICE_ERR);
            }
            EGX_HIP_TRY(hipGraphLaunch(cap.exec, s), EGX_DEVICE_ERR);
        } else {
            for (int32_t q = 0; q < todo; ++q) {
                launch_spmv();
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
                k_pr_reduce_multi<T, W><<<(unsigned)rtotal, kReduceThreads, 0, s>>>(
                    d_table.as<IterArgs<T>>(), d_rblk.as<int32_t>(), d_rbase.as<uint32_t>());
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
                k_pr_final_multi<T><<<(unsigned)count, kWave, 0, s>>>(d_table.as<IterArgs<T>>());
                EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            }
        }
        launched += todo;
        if (launched >= p.max_iter) break;
        k_gather_states<T><<<nblk(count, 256), 256, 0, s>>>(d_table.as<IterArgs<T>>(), count,
                                                             d_states.as<PrState>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(hs, d_states.p, sizeof(PrState) * (size_t)count, hipMemcpyDeviceToHost, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
        bool all_done = true;
        for (int32_t i = 0; i < count && all_done; ++i) all_done = hs[i].done != 0;
        if (all_done) break;
    }
    EGX_HIP_TRY(hipEventRecord(ev1, s), EGX_DEVICE_ERR);
    k_gather_states<T><<<nblk(count, 256), 256, 0, s>>>(d_table.as<IterArgs<T>>(), count, d_states.as<PrState>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(hs, d_states.p, sizeof(PrState) * (size_t)count, hipMemcpyDeviceToHost, s),
                EGX_DEVICE_ERR);
    // every normalised result in one launch; copies only where asked for
    k_pr_output_multi<T><<<nbase[(size_t)count], 256, 0, s>>>(d_table.as<IterArgs<T>>(), d_nbase.as<uint32_t>(),
                                                                count);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    for (int32_t i = 0; i < count; ++i)
        if (host_out && host_out[i])
            EGX_HIP_TRY(hipMemcpyAsync(host_out[i], gs[(size_t)i]->out.p, (size_t)gs[(size_t)i]->n * 8,
                                       hipMemcpyDeviceToHost, s),
                        EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    if (st) {
        float ms = 0.0f;
        EGX_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1), EGX_DEVICE_ERR);
        st->kernel_seconds = (double)ms * 1e-3;
        st->spmv_grid = (int64_t)total;
        st->spmv_launches = launched;
        st->iterations = 0;
        for (int32_t i = 0; i < count; ++i) st->iterations = std::max(st->iterations, hs[i].iter);
    }
    if (iters)
        for (int32_t i = 0; i < count; ++i) iters[i] = hs[i].iter;
    return EGX_SUCC;
}

// tele[rank[v]] = 1/k for the k seeds v (caller ids), 0 elsewhere.
__global__ void k_tele_seeds(const int32_t* __restrict__ seeds, int32_t k, const int32_t* __restrict__ rank,
                             double w, double* __restrict__ tele) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int32_t v = seeds[i];
    tele[rank ? rank[v] : v] = w;
}

constexpr int64_t kSolveAlign = 256;
int64_t align_up(int64_t v) { return (v + kSolveAlign - 1) / kSolveAlign * kSolveAlign; }
// SpMV work items of g's largest launch (partial slots needed)
int64_t solve_grid(const Graph& g) {
    return std::max<int64_t>(std::max<int64_t>(g.n_hseg + g.n_tiles, g.n_pb_sum), 1);
}

}  // namespace

int64_t solve_scratch_bytes(const Graph& g, int32_t precision) {
    const int64_t elem = precision == EGX_PRECISION_FP64 ? 8 : 4;
    const int64_t vec = align_up((int64_t)std::max(g.n, 1) * elem);
    return 4 * vec + align_up((int64_t)sizeof(PrState)) + align_up(solve_grid(g) * 3 * 8) +
           align_up((int64_t)kMaxRed * 3 * 8);
}

void carve_solve_bufs(const Graph& g, int32_t precision, char* base, SolveBufs& b) {
    const int64_t elem = precision == EGX_PRECISION_FP64 ? 8 : 4;
    const int64_t vec = align_up((int64_t)std::max(g.n, 1) * elem);
    char* p = base;
    for (int i = 0; i < 2; ++i) {
        b.x[i] = p;
        p += vec;
        b.s[i] = p;
        p += vec;
    }
    b.state = reinterpret_cast<PrState*>(p);
    p += align_up((int64_t)sizeof(PrState));
    b.partials = reinterpret_cast<double*>(p);
    p += align_up(solve_grid(g) * 3 * 8);
    b.red_part = reinterpret_cast<double*>(p);
}


int ensure_tele(Graph& g, const std::vector<int32_t>& seeds) {
    hipStream_t s = g.stream;
    const size_t nb = (size_t)std::max(g.n, 1) * 8;
    if (g.tele.bytes < nb) EGX_TRY(g.tele.alloc(nb));
    EGX_HIP_TRY(hipMemsetAsync(g.tele.p, 0, nb, s), EGX_DEVICE_ERR);
    const int32_t k = (int32_t)seeds.size();
    if (k == 0) return EGX_SUCC;
    DevBuf d;
    EGX_TRY(d.alloc((size_t)k * 4));
    EGX_HIP_TRY(hipMemcpyAsync(d.p, seeds.data(), (size_t)k * 4, hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);
    k_tele_seeds<<<nblk(k, 256), 256, 0, s>>>(d.as<int32_t>(), k, g.reordered ? g.rank.as<int32_t>() : nullptr,
                                              1.0 / (double)k, g.tele.as<double>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);  // d is freed on return
    return EGX_SUCC;
}

int solve_pagerank(Graph& g, const SolveParams& p_in, double* host_out, SolveStats* st) {
    SolveParams p = p_in;
    if (p.max_iter < 1) p.max_iter = 1;
    if (!(p.tol >= 0.0)) p.tol = 0.0;
    if (st) *st = SolveStats{};
    if (g.n <= 0) return EGX_SUCC;
    if (g.m == 0) {
        // no stored arcs: uniform ranks, as pagerank.py:93-95 and pagerank.cu:535-539
        if (host_out)
            for (int32_t i = 0; i < g.n; ++i) host_out[i] = 1.0 / (double)g.n;
        if (p.out_device) {
            k_fill_f64<<<nblk(g.n, 256), 256, 0, g.stream>>>(p.out_device, g.n, 1.0 / (double)g.n);
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            EGX_HIP_TRY(hipStreamSynchronize(g.stream), EGX_DEVICE_ERR);
        }
        return EGX_SUCC;
    }
    // adjacency graphs have out-weight 1 everywhere, so the unweighted path
    // (s = x/1, x' = alpha * row sum + b) applies alpha in FP64 to each sum
    // unless they carry weights (eigenvector centrality with a weight key)
    const bool w = g.weighted && (!g.adjacency || g.adjacency_weights);
    if (p.eig_delta && (p.precision != EGX_PRECISION_FP64 || !g.adjacency)) {
        set_last_error("eigenvector mode runs in FP64 on an adjacency graph");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (p.precision == EGX_PRECISION_FP64)
        return w ? solve_typed<double, true>(g, p, host_out, st) : solve_typed<double, false>(g, p, host_out, st);
    return w ? solve_typed<float, true>(g, p, host_out, st) : solve_typed<float, false>(g, p, host_out, st);
}

int solve_pagerank_multi(std::vector<Graph*>& gs, const SolveParams& p_in, double* const* host_out, int32_t* iters,
                         SolveStats* st, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    SolveParams p = p_in;
    if (p.max_iter < 1) p.max_iter = 1;
    if (!(p.tol >= 0.0)) p.tol = 0.0;
    if (st) *st = SolveStats{};
    // graphs without arcs (or nodes) are answered here and not launched
    std::vector<Graph*> work;
    std::vector<double*> outs;
    std::vector<int32_t> idx;
    bool weighted = false;
    for (size_t i = 0; i < gs.size(); ++i) {
        Graph& g = *gs[i];
        if (iters) iters[i] = 0;
        if (g.n <= 0) continue;
        if (g.m == 0) {
            if (host_out && host_out[i])
                for (int32_t v = 0; v < g.n; ++v) host_out[i][v] = 1.0 / (double)g.n;
            continue;
        }
        if (!work.empty() && g.weighted != weighted) {
            set_last_error("one multi-graph launch takes either weighted or unweighted graphs");
            return EGX_UNSUPPORTED_GRAPH;
        }
        weighted = g.weighted;
        work.push_back(&g);
        outs.push_back(host_out ? host_out[i] : nullptr);
        idx.push_back((int32_t)i);
    }
    if (work.empty()) return EGX_SUCC;
    std::vector<int32_t> it(work.size(), 0);
    int rc;
    if (p.precision == EGX_PRECISION_FP64)
        rc = weighted ? solve_multi_typed<double, true>(work, p, outs.data(), it.data(), st, s, ev0, ev1)
                      : solve_multi_typed<double, false>(work, p, outs.data(), it.data(), st, s, ev0, ev1);
    else
        rc = weighted ? solve_multi_typed<float, true>(work, p, outs.data(), it.data(), st, s, ev0, ev1)
                      : solve_multi_typed<float, false>(work, p, outs.data(), it.data(), st, s, ev0, ev1);
    if (rc != EGX_SUCC) return rc;
    if (iters)
        for (size_t k = 0; k < work.size(); ++k) iters[idx[k]] = it[k];
    return EGX_SUCC;
}

}  // namespace egx
