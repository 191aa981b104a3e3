// egx_katz.hip -- Katz centrality on the pull SpMV engine.
//
// The reference computes Katz centrality on the CPU only
// (easygraph/functions/centrality/katz_centrality.py:7-110, the Python route
// of its @hybrid dispatch):
//
//   A[u][v] = 1 for every arc u->v of the adjacency (self-loops included)
//   x = 1;  repeat: x_new = alpha * A x + b;  if sum|x_new - x| < tol: stop
//           (x keeps the iterate BEFORE this update);  x = x_new
//   no stop within max_iter: RuntimeError;  normalized: x / ||x||_2
//
// Row u of A x sums x over u's out-neighbours, so the pull layout is built
// from the transposed arcs: the device build transposes its input once more,
// which turns the host transpose back into "row u reads E[V[u]..V[u+1])".
// The graph is built in adjacency mode (unit weights, out-weight 1 for every
// node), so the unweighted kernels compute x'[u] = alpha * (sum of x over the
// row) + b[u] with alpha applied in FP64, as alpha * A @ x + b does; iterates
// and every sum are FP64, and the per-row bias b replaces PageRank's
// teleport and dangling terms.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "egx_graph.h"

namespace egx {
namespace {

// bias[r] = beta[perm[r]] (beta in caller order, r an iteration row), or the scalar.
__global__ void k_katz_bias(const double* __restrict__ beta, double beta_scalar, const int32_t* __restrict__ perm,
                            int32_t n, double* __restrict__ bias) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t v = perm ? perm[r] : r;
    bias[r] = beta ? beta[v] : beta_scalar;
}

inline unsigned nblk(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace
}  // namespace egx

extern "C" int egx_katz_csr(const int32_t* V, const int32_t* E, int32_t n, int32_t m, double alpha,
                            const double* beta, double beta_scalar, int32_t max_iter, double tol, int32_t normalized,
                            double* out, int32_t* iterations, int32_t* converged, double* kernel_seconds) {
    using namespace egx;
    if (iterations) *iterations = 0;
    if (converged) *converged = 0;
    if (kernel_seconds) *kernel_seconds = 0.0;
    if (n < 0 || m < 0 || (n > 0 && (!V || !out)) || (m > 0 && !E)) {
        set_last_error("egx_katz_csr: bad arguments");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (max_iter < 1) return EGX_SUCC;  // the loop never runs: not converged
    if (n == 0) {
        // sum|x_new - x| over no nodes is 0
        if (converged) *converged = tol > 0.0 ? 1 : 0;
        if (iterations) *iterations = 1;
        return EGX_SUCC;
    }
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
    auto b_at = [&](int32_t v) { return beta ? beta[v] : beta_scalar; };
    if (m == 0) {
        // x_1 = b; the test after update k compares x_k with x_{k-1}
        double e0 = 0.0;
        for (int32_t v = 0; v < n; ++v) e0 += std::fabs(b_at(v) - 1.0);
        const bool first = e0 < tol;
        if (!first && (max_iter < 2 || !(0.0 < tol))) return EGX_SUCC;  // the second test sees 0
        for (int32_t v = 0; v < n; ++v) out[v] = first ? 1.0 : b_at(v);
        if (iterations) *iterations = first ? 1 : 2;
    } else {
        // transposed arcs, stable in source order: row v of (Vt, Et) lists the u with u->v
        std::vector<int32_t> Vt((size_t)n + 1, 0), Et((size_t)m), pos;
        for (int32_t p = 0; p < m; ++p) Vt[(size_t)E[p] + 1] += 1;
        for (int32_t v = 0; v < n; ++v) Vt[(size_t)v + 1] += Vt[(size_t)v];
        pos.assign(Vt.begin(), Vt.end() - 1);
        for (int32_t u = 0; u < n; ++u)
            for (int32_t p = V[u]; p < V[u + 1]; ++p) Et[(size_t)pos[(size_t)E[p]]++] = u;

        std::unique_ptr<egx_graph> h;
        EGX_TRY(make_graph(h));
        Graph& g = h->g;
        g.adjacency = true;
        DevBuf dV, dE, dB, bias;
        EGX_TRY(dV.alloc(((size_t)n + 1) * 4));
        EGX_TRY(dE.alloc((size_t)m * 4));
        EGX_HIP_TRY(hipMemcpy(dV.p, Vt.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpy(dE.p, Et.data(), (size_t)m * 4, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
        EGX_TRY(build_from_device_csr(g, dV.as<int32_t>(), dE.as<int32_t>(), nullptr, n, m, alpha));
        dV.release();
        dE.release();
        EGX_TRY(bias.alloc((size_t)n * 8));
        if (beta) {
            EGX_TRY(dB.alloc((size_t)n * 8));
            EGX_HIP_TRY(hipMemcpy(dB.p, beta, (size_t)n * 8, hipMemcpyHostToDevice), EGX_DEVICE_ERR);
        }
        k_katz_bias<<<nblk(n, 256), 256, 0, g.stream>>>(beta ? dB.as<double>() : nullptr, beta_scalar,
                                                         g.reordered ? g.perm.as<int32_t>() : nullptr, n,
                                                         bias.as<double>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        SolveParams p;
        p.alpha = alpha;
        p.max_iter = max_iter;
        p.tol = tol;
        p.precision = EGX_PRECISION_FP64;
        p.bias = bias.as<double>();
        SolveStats st;
        EGX_TRY(solve_pagerank(g, p, out, &st));
        const PrState* hs = static_cast<const PrState*>(g.host_state);
        if (iterations) *iterations = hs->iter;
        if (kernel_seconds) *kernel_seconds = st.kernel_seconds;
        if (!(hs->err < tol)) return EGX_SUCC;  // stopped by max_iter
    }
    if (converged) *converged = 1;
    if (normalized) {
        double ss = 0.0;
        for (int32_t v = 0; v < n; ++v) ss += out[v] * out[v];
        const double norm = std::sqrt(ss);
        if (norm > 0.0)
            for (int32_t v = 0; v < n; ++v) out[v] /= norm;
    }
    return EGX_SUCC;
}
