// egx_gen.hip -- seeded synthetic graphs generated on the device.
//
// The reference ships no R-MAT generator (its graph_generator/ only has
// Erdos-Renyi on Python's global random).  These are this project's own
// definitions, counter based so that any edge can be drawn independently:
//
//   word(e, j) = mix64(seed + G * (16*e + j + 1)),  G = 0x9E3779B97F4A7C15
//   R-MAT level l of edge e uses 32 bits of word(e, l/2) (low half for even
//   l, high half for odd l); quadrant 0/1/2/3 = (src,dst) bits 00/01/10/11
//   chosen by comparing with floor(cumulative probability * 2^32).
//   ER: src = hi32(word(e,0)) % n, dst = lo32(word(e,0)) % n.
//
// Self-loops and duplicate arcs are removed by sorting the packed keys
// (src << B | dst) and keeping first occurrences; the result is an out-CSR
// sorted by (src, dst), the order Graph::gen_CSR produces
// (Easy-Graph/cpp_easygraph/classes/graph.cpp:824-886).
#include <algorithm>

#include "egx_graph.h"

namespace egx {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// Draw e descends `scale` quadrant levels; a draw that lands on a self-loop
// or on an endpoint >= n (n not a power of two) becomes the sentinel key.
__global__ void k_rmat_keys(int32_t scale, int32_t n, int64_t nedges, uint64_t seed, uint32_t tA, uint32_t tAB,
                            uint32_t tABC, uint64_t* __restrict__ keys) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nedges) return;
    uint32_t s = 0, d = 0;
    uint64_t word = 0;
    for (int32_t l = 0; l < scale; ++l) {
        if ((l & 1) == 0) word = mix64(seed + kGolden * ((uint64_t)e * 16u + (uint64_t)(l >> 1) + 1u));
        const uint32_t u = (l & 1) ? (uint32_t)(word >> 32) : (uint32_t)word;
        const uint32_t q = (u < tA) ? 0u : (u < tAB) ? 1u : (u < tABC) ? 2u : 3u;
        s = (s << 1) | (q >> 1);
        d = (d << 1) | (q & 1u);
    }
    const uint64_t sentinel = (2 * scale >= 64) ? ~0ull : ((1ull << (2 * scale)) - 1ull);
    const bool drop = (s == d) || s >= (uint32_t)n || d >= (uint32_t)n;
    keys[e] = drop ? sentinel : (((uint64_t)s << scale) | (uint64_t)d);
}

__global__ void k_er_keys(int32_t n, int32_t bits, int64_t nedges, uint64_t seed, uint64_t* __restrict__ keys) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nedges) return;
    const uint64_t word = mix64(seed + kGolden * ((uint64_t)e * 16u + 1u));
    const uint32_t s = (uint32_t)(word >> 32) % (uint32_t)n;
    const uint32_t d = (uint32_t)word % (uint32_t)n;
    const uint64_t sentinel = (1ull << (2 * bits)) - 1ull;
    keys[e] = (s == d) ? sentinel : (((uint64_t)s << bits) | (uint64_t)d);
}

__global__ void k_unique_flags(const uint64_t* __restrict__ keys, int64_t n, uint64_t sentinel,
                               uint32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    flags[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

__global__ void k_compact_keys(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                               const uint32_t* __restrict__ pos, int64_t n, uint64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) out[pos[i]] = keys[i];
}

__global__ void k_keys_to_E(const uint64_t* __restrict__ keys, int64_t m, int32_t bits, int32_t* __restrict__ E) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    E[i] = (int32_t)(keys[i] & ((1ull << bits) - 1ull));
}

__global__ void k_weights(const int32_t* __restrict__ V, const int32_t* __restrict__ E, int32_t n, int64_t m,
                          uint64_t wseed, double lo, double hi, double* __restrict__ W) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    int32_t l = 0, h = n + 1;
    while (l < h) {
        int32_t mid = (l + h) >> 1;
        if ((int64_t)V[mid] <= p) l = mid + 1;
        else h = mid;
    }
    const uint32_t src = (uint32_t)(l - 1);
    const uint64_t key = ((uint64_t)src << 32) | (uint64_t)(uint32_t)E[p];
    const uint64_t u = mix64(wseed ^ key) >> 40;
    const double t = (double)u * (1.0 / 16777216.0);
    const double span = __dsub_rn(hi, lo);
    const double prod = __dmul_rn(span, t);
    W[p] = (double)(float)__dadd_rn(lo, prod);
}

inline unsigned nblk(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// keys -> unique sorted keys -> (V, E)
int keys_to_csr(DevBuf& keys, int64_t nedges, int32_t n, int32_t bits, DevBuf& V, DevBuf& E, int64_t& m,
                hipStream_t s) {
    const uint64_t sentinel = (2 * bits >= 64) ? ~0ull : ((1ull << (2 * bits)) - 1ull);
    {
        DevBuf alt;
        EGX_TRY(alt.alloc((size_t)std::max<int64_t>(nedges, 1) * 8));
        EGX_TRY(radix_sort_u64(keys.as<uint64_t>(), alt.as<uint64_t>(), nullptr, nullptr, nedges, 2 * bits, s));
    }
    DevBuf flags, pos, ukeys;
    EGX_TRY(flags.alloc((size_t)std::max<int64_t>(nedges, 1) * 4));
    EGX_TRY(pos.alloc((size_t)std::max<int64_t>(nedges, 1) * 4));
    m = 0;
    if (nedges > 0) {
        k_unique_flags<<<nblk(nedges, 256), 256, 0, s>>>(keys.as<uint64_t>(), nedges, sentinel, flags.as<uint32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
        EGX_TRY(exclusive_scan_u32(flags.as<uint32_t>(), pos.as<uint32_t>(), nedges, s));
        uint32_t last[2] = {0, 0};
        EGX_HIP_TRY(hipMemcpyAsync(&last[0], pos.as<uint32_t>() + (nedges - 1), 4, hipMemcpyDeviceToHost, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipMemcpyAsync(&last[1], flags.as<uint32_t>() + (nedges - 1), 4, hipMemcpyDeviceToHost, s),
                    EGX_DEVICE_ERR);
        EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
        m = (int64_t)last[0] + last[1];
    }
    if (m >= (int64_t)0x7FFFFFFF) {
        set_last_error("generated graph exceeds the signed 32-bit CSR limit");
        return EGX_UNSUPPORTED_GRAPH;
    }
    EGX_TRY(ukeys.alloc((size_t)std::max<int64_t>(m, 1) * 8));
    if (nedges > 0) {
        k_compact_keys<<<nblk(nedges, 256), 256, 0, s>>>(keys.as<uint64_t>(), flags.as<uint32_t>(),
                                                          pos.as<uint32_t>(), nedges, ukeys.as<uint64_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    flags.release();
    pos.release();
    keys.release();
    EGX_TRY(V.alloc(((size_t)n + 1) * 4));
    EGX_TRY(E.alloc((size_t)std::max<int64_t>(m, 1) * 4));
    EGX_TRY(lower_bound_iota_u64(ukeys.as<uint64_t>(), m, bits, V.as<int32_t>(), (int64_t)n + 1, s));
    if (m > 0) {
        k_keys_to_E<<<nblk(m, 256), 256, 0, s>>>(ukeys.as<uint64_t>(), m, bits, E.as<int32_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

}  // namespace

int gen_rmat_csr(int32_t n, int64_t nedges, uint64_t seed, uint32_t tA, uint32_t tAB, uint32_t tABC, DevBuf& V,
                 DevBuf& E, int64_t& m, hipStream_t s) {
    if (n < 2 || n > (1 << 30)) {
        set_last_error("R-MAT node count must be in [2, 2^30]");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int32_t scale = bits_for((uint64_t)n - 1);
    if (nedges < 0 || nedges > (int64_t)0xFFFFFFFFll) {
        set_last_error("R-MAT draw count must be in [0, 2^32)");
        return EGX_UNSUPPORTED_GRAPH;
    }
    DevBuf keys;
    EGX_TRY(keys.alloc((size_t)std::max<int64_t>(nedges, 1) * 8));
    if (nedges > 0) {
        k_rmat_keys<<<nblk(nedges, 256), 256, 0, s>>>(scale, n, nedges, seed, tA, tAB, tABC, keys.as<uint64_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    return keys_to_csr(keys, nedges, n, scale, V, E, m, s);
}

int gen_er_csr(int32_t n, int64_t nedges, uint64_t seed, DevBuf& V, DevBuf& E, int64_t& m, hipStream_t s) {
    if (n < 1) {
        set_last_error("ER graph needs n >= 1");
        return EGX_UNSUPPORTED_GRAPH;
    }
    if (nedges < 0 || nedges > (int64_t)0xFFFFFFFFll) {
        set_last_error("ER draw count must be in [0, 2^32)");
        return EGX_UNSUPPORTED_GRAPH;
    }
    const int32_t bits = std::max(1, bits_for((uint64_t)(n - 1)));
    DevBuf keys;
    EGX_TRY(keys.alloc((size_t)std::max<int64_t>(nedges, 1) * 8));
    if (nedges > 0) {
        k_er_keys<<<nblk(nedges, 256), 256, 0, s>>>(n, bits, nedges, seed, keys.as<uint64_t>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    return keys_to_csr(keys, nedges, n, bits, V, E, m, s);
}

int gen_edge_weights(const int32_t* dV, const int32_t* dE, int32_t n, int64_t m, uint64_t wseed, double lo,
                     double hi, DevBuf& W, hipStream_t s) {
    EGX_TRY(W.alloc((size_t)std::max<int64_t>(m, 1) * 8));
    if (m > 0) {
        k_weights<<<nblk(m, 256), 256, 0, s>>>(dV, dE, n, m, wseed, lo, hi, W.as<double>());
        EGX_CHECK_LAUNCH(EGX_DEVICE_ERR);
    }
    EGX_HIP_TRY(hipStreamSynchronize(s), EGX_DEVICE_ERR);
    return EGX_SUCC;
}

}  // namespace egx
