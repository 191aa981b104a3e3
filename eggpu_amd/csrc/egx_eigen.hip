// egx_eigen.hip -- eigenvector centrality on the pull SpMV engine.
//
// The reference computes eigenvector centrality on the CPU only, by two
// power iterations:
//   * easygraph/functions/centrality/eigenvector.py:110-131 (Python route,
//     nstart given): x_0 = x0 / sum|x0|; y = A x; ||y|| < 1e-10 keeps x and
//     stops; x' = y / ||y||; stop when ||x' - x||_2 < tol; EasyGraphError
//     after max_iter iterations;
//   * cpp_easygraph/functions/centrality/eigenvector.cpp:33-97 (C++ twin):
//     x_0 = x0 * (1/||x0||_2) (uniform 1/sqrt(n) when ||x0|| < 1e-12);
//     ||y|| < 1e-12 stops; x' = y * (1/||y||); stop when sum|x' - x| < tol;
//     x is returned after max_iter iterations without an error.
// In both A[v][u] = w(u, v): row v of A x sums over v's in-arcs, which is
// the pull layout the build makes from the caller's CSR as it stands.
//
// The graph is built in adjacency mode at alpha 1 (out-weight 1 for every
// node; with weights the value array holds fl32(w)).  The SpMV launches keep
// the iterate unnormalised (y) with its scale 1/||y|| in the device state and
// find iteration k's delta in launch k+1, from the buffer that launch
// overwrites; every stop decision is taken on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "egx_graph.h"

extern "C" int egx_eigenvector_csr(const int32_t* V, const int32_t* E, const double* W, int32_t n, int32_t m,
                                   const double* x0, const egx_eig_opts* opts, double* out, int32_t* iterations,
                                   int32_t* stop, double* kernel_seconds) {
    using namespace egx;
    if (iterations) *iterations = 0;
    if (stop) *stop = 0;
    if (kernel_seconds) *kernel_seconds = 0.0;
    if (!opts || n < 0 || m < 0 || (n > 0 && (!V || !out || !x0)) || (m > 0 && !E)) {
        set_last_error("egx_eigenvector_csr: bad arguments");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const egx_eig_opts o = *opts;
    if (o.max_iter < 0 || (o.delta != 1 && o.delta != 2) || (o.start_norm != 1 && o.start_norm != 2) ||
        !(o.tol >= 0.0) || !(o.norm_floor > 0.0) || !std::isfinite(o.norm_floor) || !(o.shift >= 0.0) ||
        !std::isfinite(o.shift)) {
        set_last_error("egx_eigenvector_csr: bad options (delta and start_norm are 1 or 2; tol >= 0; "
                       "norm_floor > 0; shift >= 0)");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (n == 0) return EGX_SUCC;
    if (V[0] != 0 || V[n] != m) {
        set_last_error("invalid CSR offsets: V[0] must be 0 and V[n] must equal the arc count");
        return EGX_UNSUPPORTED_GRAPH;
    }
    for (int32_t u = 0; u < n; ++u)
        if (V[u + 1] < V[u]) {
            set_last_error("invalid CSR offsets: offsets must be monotonic");
            return EGX_UNSUPPORTED_GRAPH;
        }
    for (int32_t p = 0; p < m; ++p)
        if (E[p] < 0 || E[p] >= n) {
            set_last_error("invalid CSR index: endpoint is outside [0, num_nodes)");
            return EGX_UNSUPPORTED_GRAPH;
        }
    if (W)
        for (int32_t p = 0; p < m; ++p)
            if (!std::isfinite(W[p]) || W[p] < 0.0) {
                set_last_error("eigenvector centrality takes finite, non-negative weights");
                return EGX_UNSUPPORTED_GRAPH;
            }

    // the normalised start, caller order
    std::vector<double> x((size_t)n);
    if (o.start_norm == 1) {
        double s1 = 0.0;
        for (int32_t v = 0; v < n; ++v) s1 += std::fabs(x0[v]);
        if (!(s1 > 0.0) || !std::isfinite(s1)) {
            set_last_error("eigenvector centrality: the start vector's absolute sum must be finite and positive");
            return EGX_UNSUPPORTED_GRAPH;
        }
        for (int32_t v = 0; v < n; ++v) x[(size_t)v] = x0[v] / s1;
    } else {
        double ss = 0.0;
        for (int32_t v = 0; v < n; ++v) ss += x0[v] * x0[v];
        const double nrm = std::sqrt(ss);
        if (!std::isfinite(nrm)) {
            set_last_error("eigenvector centrality: the start vector must be finite");
            return EGX_UNSUPPORTED_GRAPH;
        }
        if (nrm < 1e-12) {
            for (int32_t v = 0; v < n; ++v) x[(size_t)v] = 1.0 / std::sqrt((double)n);
        } else {
            const double inv = 1.0 / nrm;
            for (int32_t v = 0; v < n; ++v) x[(size_t)v] = x0[v] * inv;
        }
    }
    for (int32_t v = 0; v < n; ++v)
        if (!std::isfinite(x[(size_t)v])) {
            set_last_error("eigenvector centrality: the start vector must be finite");
            return EGX_UNSUPPORTED_GRAPH;
        }

    if (o.max_iter == 0) {
        std::memcpy(out, x.data(), (size_t)n * 8);
        if (stop) *stop = 2;
        return EGX_SUCC;
    }
    if (m == 0) {
        // A = 0: y = shift x.  With no shift the first norm test stops;
        // otherwise x_1 = x_0 / ||x_0|| and every later iterate equals x_1.
        double ss = 0.0;
        for (int32_t v = 0; v < n; ++v) ss += x[(size_t)v] * x[(size_t)v];
        const double ny = o.shift * std::sqrt(ss);
        if (!(ny >= o.norm_floor) || !(ny > 0.0)) {
            std::memcpy(out, x.data(), (size_t)n * 8);
            if (iterations) *iterations = 1;
            if (stop) *stop = 3;
            return EGX_SUCC;
        }
        std::vector<double> x1((size_t)n);
        double d = 0.0;
        for (int32_t v = 0; v < n; ++v) {
            x1[(size_t)v] = (o.shift * x[(size_t)v]) * (1.0 / ny);
            const double e = x1[(size_t)v] - x[(size_t)v];
            d += o.delta == 1 ? std::fabs(e) : e * e;
        }
        if (o.delta == 2) d = std::sqrt(d);
        // iteration 2 sees ||shift x_1|| = shift and then x_2 = x_1 (delta 0)
        int32_t k = 1, st = 1;
        if (!(d < o.tol)) {
            if (o.max_iter == 1) st = 2;
            else if (!(o.shift >= o.norm_floor)) {
                k = 2;
                st = 3;
            } else if (o.tol > 0.0) k = 2;
            else {
                k = o.max_iter;
                st = 2;
            }
        }
        std::memcpy(out, x1.data(), (size_t)n * 8);
        if (iterations) *iterations = k;
        if (stop) *stop = st;
        return EGX_SUCC;
    }

    std::unique_ptr<egx_graph> h;
    EGX_TRY(make_graph(h));
    Graph& g = h->g;
    g.adjacency = true;
    DevBuf dV, dE, dW, dx;
    EGX_TRY(dV.alloc(((size_t)n + 1) * 4));
    EGX_TRY(dE.alloc((size_t)m * 4));
    EGX_HIP_TRY(hipMemcpy(dV.p, V, ((size_t)n + 1) * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    EGX_HIP_TRY(hipMemcpy(dE.p, E, (size_t)m * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    if (W) {
        EGX_TRY(dW.alloc((size_t)m * 8));
        EGX_HIP_TRY(hipMemcpy(dW.p, W, (size_t)m * 8, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
    }
    EGX_TRY(build_from_device_csr(g, dV.as<int32_t>(), dE.as<int32_t>(), W ? dW.as<double>() : nullptr, n, m, 1.0));
    dV.release();
    dE.release();
    dW.release();
    EGX_TRY(dx.alloc((size_t)n * 8));
    EGX_HIP_TRY(hipMemcpy(dx.p, x.data(), (size_t)n * 8, hipMemcpyHostToDevice), EGX_DEVICE_ERR);

    SolveParams p;
    p.alpha = 1.0;
    p.max_iter = o.max_iter;
    p.tol = o.tol;
    p.precision = EGX_PRECISION_FP64;
    p.eig_delta = o.delta;
    p.eig_floor = o.norm_floor;
    p.eig_shift = o.shift;
    p.eig_x0 = dx.as<double>();
    SolveStats st;
    EGX_TRY(solve_pagerank(g, p, out, &st));
    const PrState* hs = static_cast<const PrState*>(g.host_state);
    if (iterations) *iterations = hs->eig_iters;
    if (stop) *stop = hs->eig_stop;
    if (kernel_seconds) *kernel_seconds = st.kernel_seconds;
    return EGX_SUCC;
}
