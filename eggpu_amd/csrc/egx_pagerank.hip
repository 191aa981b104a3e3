// egx_pagerank.hip -- power iteration on the device-resident iteration layout.
//
// Two launches per iteration replace the reference's cuSPARSE SpMV, cuBLAS
// Sdot, two memsets, fused update/reduce kernel, scale kernel and two host
// round trips (pagerank.cu:591-664):
//
//   x'[v] = alpha * sum_{u->v} x[u]/outdeg[u] + jump          (unweighted)
//   x'[v] = sum_{u->v} in_val[u->v] * x[u] + jump             (weighted)
//   jump  = (1-alpha)/n + alpha*dang/n                        (pagerank.py:125)
//   err   = sum_v |x'[v] - x[v]|,  dang' = sum_{v dangling} x'[v]
//   stop when err < tol*n (pagerank.py:131), decided on the device.
//
// k_pr_spmv, one workgroup per work item:
//   * hub segment (rows with in-degree above the hub threshold): a fixed
//     slice of one row, 8 gathers in flight per thread, FP64 block sum
//     written to hub_part;
//   * tile (all other rows): the rows whose merge-path diagonal starts in
//     the tile.  Their arcs are contiguous; the block first stages every
//     gathered term into LDS (coalesced index loads, 8 independent gathers
//     per thread), then reduces rows from LDS -- one thread per row when the
//     tile's mean row length is at most thread_row_max, one wavefront per
//     row with a wave64 shuffle reduction otherwise -- and finalises them.
// k_pr_reduce (one workgroup) finalises hub rows from their segment partials
// in segment order and folds the per-block (err, sum, dang) slots in index
// order: results are run-to-run deterministic and no launch uses a global
// atomic (device-scope atomics on one address serialise across the 8 XCDs).
// Row sums, the teleport term, the dangling mass and the L1 residual are
// FP64; iterates are stored as T (float or double).  Unweighted graphs keep
// s = x/outdeg beside x, so each arc costs one 4-byte index and one gather.
#include <algorithm>
#include <cmath>
#include <vector>

#include "egx_graph.h"

namespace egx {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kReduceThreads = 1024;
constexpr int kBatch = 8;  // gathers in flight per thread

template <typename T>
struct IterArgs {
    const int32_t* row;
    const int32_t* col;
    const float* val;
    const float* deg;
    const int32_t* tiles;
    const HubSeg* hsegs;
    const HubRow* hrows;
    double* hub_part;
    const T* x_in;
    const T* s_in;
    T* x_out;
    T* s_out;
    PrState* state;
    double* partials;  // SoA: err[grid], sum[grid], dang[grid]
    int64_t n_hrows;
    int32_t n;
    int32_t thread_row_max;
    uint32_t nblk_h;
    uint32_t grid;
    double alpha;
    double tol_n;
    int32_t max_iter;
};

// The gathered term of an arc from u: x[u] * in_val (weighted) or
// s[u] = x[u]/outdeg[u].
template <typename T, bool W>
__device__ __forceinline__ T term(const IterArgs<T>& a, int32_t u, float w) {
    if (W) return (T)w * a.x_in[u];
    return a.s_in[u];
}

template <typename T, bool W>
__device__ __forceinline__ void finalize_row(const IterArgs<T>& a, int32_t v, double acc, double jump, double& err,
                                             double& sum, double& dang) {
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

template <typename T>
__device__ __forceinline__ double jump_of(const IterArgs<T>& a) {
    const double n_d = (double)a.n;
    return (1.0 - a.alpha) / n_d + a.alpha * a.state->dang / n_d;  // pagerank.py:125
}

// col / val are streamed once per iteration: non-temporal loads keep them
// from evicting the gathered vector from L2.
template <typename T, bool W>
__global__ __launch_bounds__(kThreads) void k_pr_spmv(IterArgs<T> a) {
    __shared__ T lds[kTileLds];
    __shared__ double sh_red[3 * kWavesPerBlock];
    if (a.state->done) return;  // uniform: written by an earlier launch
    const double jump = jump_of(a);
    const uint32_t b = blockIdx.x;
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
    for (int32_t base = 0; base < na; base += kThreads * kBatch) {
        int32_t u[kBatch];
        float w[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int32_t j = base + q * kThreads + t;
            const bool ok = j < na;
            u[q] = ok ? __builtin_nontemporal_load(&a.col[a0 + j]) : -1;
            w[q] = (W && ok) ? __builtin_nontemporal_load(&a.val[a0 + j]) : 0.0f;
        }
        T g[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) g[q] = (u[q] >= 0) ? term<T, W>(a, u[q], w[q]) : (T)0;
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int32_t j = base + q * kThreads + t;
            if (j < na) lds[j] = g[q];
        }
    }
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

// Launch 2 of an iteration (one workgroup): finalise hub rows from their
// segment partials (segment order), fold every partial slot in index order,
// then update the state and take the stop decision err < tol*n
// (pagerank.py:131) on the device.
template <typename T, bool W>
__global__ __launch_bounds__(kReduceThreads) void k_pr_reduce(IterArgs<T> a) {
    __shared__ double sh_red[3 * (kReduceThreads / kWave)];
    if (a.state->done) return;
    const double jump = jump_of(a);
    const int t = threadIdx.x;
    double e = 0.0, su = 0.0, dg = 0.0;
    for (int64_t h = t; h < a.n_hrows; h += kReduceThreads) {
        const HubRow hr = a.hrows[h];
        double row = 0.0;
        for (int32_t q = 0; q < hr.nseg; ++q) row += a.hub_part[hr.first + q];
        finalize_row<T, W>(a, hr.row, row, jump, e, su, dg);
    }
    const double* pe = a.partials;
    const double* ps = a.partials + a.grid;
    const double* pd = a.partials + 2 * (size_t)a.grid;
#pragma unroll 8
    for (uint32_t i = t; i < a.grid; i += kReduceThreads) {
        e += pe[i];
        su += ps[i];
        dg += pd[i];
    }
    block_sum3(e, su, dg, sh_red);
    if (t == 0) {
        PrState* st = a.state;
        const int32_t it = st->iter + 1;
        st->err = e;
        st->sum = su;
        st->dang = dg;
        st->iter = it;
        st->done = (e < a.tol_n || it >= a.max_iter) ? 1 : 0;
    }
}

// x0 = 1/n (as T), s0 = x0/deg.
template <typename T>
__global__ void k_pr_init(const float* __restrict__ deg, int32_t n, T* __restrict__ x, T* __restrict__ s) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const T x0 = (T)(1.0 / (double)n);
    x[v] = x0;
    const float d = deg[v];
    s[v] = (d > 0.0f) ? (T)(x0 / (T)d) : (T)0;
}

// out[perm[r]] = x[r] * (1/sum) in FP64, uniform when sum <= 0
// (pagerank.py:134-139); perm maps iteration rows to caller node ids.
template <typename T>
__global__ void k_pr_output(const T* __restrict__ x, const PrState* __restrict__ st, const int32_t* __restrict__ perm,
                            int32_t n, double* __restrict__ out) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double norm = st->sum;
    const int32_t v = perm ? perm[r] : r;
    out[v] = (norm > 0.0) ? (double)x[r] * (1.0 / norm) : 1.0 / (double)n;
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

template <typename T, bool W>
int solve_typed(Graph& g, const SolveParams& p, double* host_out, SolveStats* st) {
    hipStream_t s = g.stream;
    const int32_t n = g.n;
    const uint32_t nblk_h = (uint32_t)g.n_hseg;
    const uint32_t grid = nblk_h + (uint32_t)g.n_tiles;
    if (grid == 0) {
        set_last_error("graph has no work items");
        return EGX_UNKNOW_ERROR;
    }
    EGX_TRY(ensure_workspace(g, grid));

    T* xb[2] = {g.x[0].as<T>(), g.x[1].as<T>()};
    T* sb[2] = {g.s[0].as<T>(), g.s[1].as<T>()};
    k_pr_init<T><<<nblk(n, 256), 256, 0, s>>>(g.it_deg, n, xb[0], sb[0]);
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    PrState st0{};
    st0.dang = (double)g.n_dangling * (double)(T)(1.0 / (double)n);
    st0.sum = 1.0;
    EGX_HIP_TRY(hipMemcpyAsync(g.state.p, &st0, sizeof(PrState), hipMemcpyHostToDevice, s), EGX_DEVICE_ERR);

    IterArgs<T> a{};
    a.row = g.it_row;
    a.col = g.it_col;
    a.val = g.it_val;
    a.deg = g.it_deg;
    a.tiles = g.tiles.as<int32_t>();
    a.hsegs = g.hsegs.as<HubSeg>();
    a.hrows = g.hrows.as<HubRow>();
    a.hub_part = g.hub_part.as<double>();
    a.state = g.state.as<PrState>();
    a.partials = g.partials.as<double>();
    a.n_hrows = g.n_hrows;
    a.n = n;
    a.thread_row_max = g.bins.thread_row_max;
    a.nblk_h = nblk_h;
    a.grid = grid;
    a.alpha = p.alpha;
    a.tol_n = std::max(0.0, p.tol) * (double)n;
    a.max_iter = p.max_iter;

    LaunchTimer timer;
    if (p.time_each_launch) EGX_TRY(timer.reserve((size_t)4 * p.max_iter));
    EGX_HIP_TRY(hipEventRecord(g.ev0, s), EGX_DEVICE_ERR);
    int32_t launched = 0;
    int32_t chunk = p.check_every > 0 ? p.check_every : 4;
    const bool can_stop_early = p.tol > 0.0;
    PrState* hs = static_cast<PrState*>(g.host_state);
    while (launched < p.max_iter) {
        const int32_t todo = can_stop_early ? std::min(chunk, p.max_iter - launched) : (p.max_iter - launched);
        for (int32_t q = 0; q < todo; ++q, ++launched) {
            const int cur = launched & 1;
            a.x_in = xb[cur];
            a.s_in = sb[cur];
            a.x_out = xb[cur ^ 1];
            a.s_out = sb[cur ^ 1];
            if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched], s), EGX_DEVICE_ERR);
            k_pr_spmv<T, W><<<grid, kThreads, 0, s>>>(a);
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 1], s), EGX_DEVICE_ERR);
            if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 2], s), EGX_DEVICE_ERR);
            k_pr_reduce<T, W><<<1, kReduceThreads, 0, s>>>(a);
            EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
            if (p.time_each_launch) EGX_HIP_TRY(hipEventRecord(timer.ev[4 * launched + 3], s), EGX_DEVICE_ERR);
        }
        if (launched >= p.max_iter) break;
        EGX_HIP_TRY(hipMemcpyAsync(hs, g.state.p, sizeof(PrState), hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
        if (hs->done) break;
        if (p.check_every <= 0) chunk = std::min(chunk * 2, 64);
    }
    EGX_HIP_TRY(hipEventRecord(g.ev1, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpyAsync(hs, g.state.p, sizeof(PrState), hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
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
    // the final iterate sits in the buffer written by launch iters-1
    T* xf = xb[iters & 1];
    k_pr_output<T><<<nblk(n, 256), 256, 0, s>>>(xf, g.state.as<PrState>(),
                                                 g.reordered ? g.perm.as<int32_t>() : nullptr, n,
                                                 g.out.as<double>());
    EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    if (host_out)
        EGX_HIP_TRY(hipMemcpyAsync(host_out, g.out.p, (size_t)n * 8, hipMemcpyDeviceToHost, s), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

}  // namespace

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
        return EGX_SUCC;
    }
    EGX_TRY(refresh_alpha(g, p.alpha));
    if (p.precision == EGX_PRECISION_FP64)
        return g.weighted ? solve_typed<double, true>(g, p, host_out, st)
                          : solve_typed<double, false>(g, p, host_out, st);
    return g.weighted ? solve_typed<float, true>(g, p, host_out, st) : solve_typed<float, false>(g, p, host_out, st);
}

}  // namespace egx
